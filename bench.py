#!/usr/bin/env python3
"""Benchmark of the SNaC time-step hot path on one MI355X.

Workload: the FFT-accelerated periodic channel, plain FP64 -- 256x128x256
cells (x periodic, y wall-normal tanh-stretched with factor 1.8, z periodic
and the FFT axis), Re_b = 5600 with the laminar body force bforce(1) = 12/5600,
laminar Poiseuille profile plus seeded uniform +-10% noise.  A step is one
full RK3 time step (3 substages: fused momentum RHS + update, ghost fills,
fillps, r2r transforms along z and x, tridiagonal solves along y, inverse
transforms, correc + pressure update).  Inputs are resident in HBM before
the timed region; dt is fixed from the CFL check at the start.

Prints one JSON line: Mcell-steps/s, the per-phase split (Poisson ms/step),
the roofline of the dominant kernel (fused rk_mom, 104 algorithmic bytes
per cell) from HIP events on its own stream, and the CPU port's rate on the
host cores (a bounded sample of the same workload).

--precision mixed runs the MG-PCG V-cycle in FP32 (MG-PCG path only; the
channel's direct solver and the stencils stay FP64).  --precision learned
(MG-PCG path) first records the plain solver's (rhs, solution) pairs over
--learn-steps steps, trains the network on them for --learn-epochs epochs,
then times steps with the network preconditioning the flexible CG (V-cycle
fallback when its residual reduction stalls; eager launches); the line adds
the network's MFMA rate (one forward over every mode, HIP events) against
the 2.5 PFLOP/s dense BF16 peak.

--impdiff runs the channel with implicit diffusion (the reference's
IMPDIFF=1 build): each substage adds three Helmholtz solves (one per velocity
component) on the same tensor-product plan; not the headline workload, and
no CPU sample (the oracle's direct Helmholtz solves are too slow at this
size).

--workload bfs runs the three-block backward-facing step instead (inlet
192x64x128 with a Poiseuille inflow, lower 320x64x128, upper 320x96x128, z
periodic with 128 modes, y tanh-stretched per block, outflow at x+): the
union is not a box, so the pressure goes through the batched MG-PCG path
(per-mode relative residual 1e-13); the CPU sample is skipped there.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "Mcell-steps/s + Poisson ms/step, learned/mixed vs plain FP64, equal div tol"
HBM_PEAK_GBS = 8000.0
RK_MOM_BYTES_PER_CELL = 104  # R u,v,w,p,3x dudtrko + W u,v,w,3x dudtrko, FP64
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "rk_mom_channel_256x128x256_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=["channel", "bfs"], default="channel")
    ap.add_argument("--size", type=int, nargs=3, default=[256, 128, 256])
    ap.add_argument("--solver", choices=["auto", "tensor", "mg"], default="auto",
                    help="pressure solver: the channel's direct tensor-product solver (auto) or MG-PCG")
    ap.add_argument("--tol", type=float, default=1e-13,
                    help="MG-PCG relative residual per mode (&hypre hypre_tol), bfs workload")
    ap.add_argument("--learn-steps", type=int, default=20)
    ap.add_argument("--learn-epochs", type=int, default=4)
    ap.add_argument("--precision", choices=["fp64", "mixed", "learned"], default="fp64",
                    help="mixed: FP32 MG V-cycle under flexible CG (MG-PCG path only)")
    ap.add_argument("--impdiff", action="store_true", help="implicit diffusion (channel workload)")
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps for the CPU baseline (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cpu count)")
    ap.add_argument("--calib", action="store_true", help="also run the plain copy kernel (counter calibration)")
    return ap.parse_args()


def channel_setup(tmp, size, tol):
    import cases
    d, b = cases.write_case(tmp, [cases.channel_block(size)], 5600.0, cfl=0.5, bforce=(12.0 / 5600.0, 0.0, 0.0))
    txt = open(d).read().replace("hypre_tol = 1.e-13", f"hypre_tol = {tol!r}")
    open(d, "w").write(txt)
    return d, b


def bfs_setup(tmp, tol):
    import cases
    d, b = cases.write_case(tmp, cases.bfs_blocks(), 1000.0, cfl=0.5)
    txt = open(d).read().replace("hypre_tol = 1.e-13", f"hypre_tol = {tol!r}")
    open(d, "w").write(txt)
    return d, b


def bfs_ic(ctx, seed=20251202, amp=0.1):
    """Poiseuille u = 6 y (1 - y) below the step height (inlet and lower
    blocks), rest at zero, seeded +-amp noise on u, v, w interiors."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for b in range(ctx.nblocks):
        sh = ctx.shape(b, "u")
        yc = ctx.metric(b, 8)
        u = np.zeros(sh); v = np.zeros(sh); w = np.zeros(sh)
        inner = (slice(1, -1),) * 3
        if yc[1:-1].max() <= 1.0:
            u[inner] = (6.0 * yc[1:-1] * (1.0 - yc[1:-1]))[None, :, None]
        for a in (u, v, w):
            a[inner] += amp * rng.uniform(-1, 1, a[inner].shape)
        out.append((u, v, w))
    return out


def cpu_baseline(d, b, size, steps, threads):
    import cases
    import nmlparse
    import oracle_py
    n = oracle_py.set_threads(threads)
    oc = oracle_py.OracleCase(nmlparse.load_case(d, b, 3))
    u, v, w = cases.channel_ic(oc.shape(0, "u"), oc.metric(0, 8))
    for f, a in (("u", u), ("v", v), ("w", w)):
        oc.set(0, f, a)
    oc.init_bounds()
    dt = 0.5 * oc.chkdt()
    t0 = time.perf_counter()
    for _ in range(steps):
        oc.step(dt)
    el = time.perf_counter() - t0
    cells = size[0] * size[1] * size[2]
    return {"value": cells * steps / el / 1e6, "unit": "Mcell-steps/s", "cores": int(n), "kind": "port",
            "sample": f"{steps} full RK3 steps of the {size[0]}x{size[1]}x{size[2]} channel on the C oracle "
                      f"(OpenMP, tensor Poisson solver, radix-2 transforms), {el:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the time step runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding the blocks over GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import cases
    import snac_amd

    size = args.size
    bfs = args.workload == "bfs"
    with tempfile.TemporaryDirectory() as tmp:
        if bfs:
            d, b = bfs_setup(tmp, args.tol)
        else:
            d, b = channel_setup(tmp, size, args.tol)
        print(f"[bench] {args.workload} {'' if bfs else size} on device 0", file=sys.stderr, flush=True)
        ctx = snac_amd.Context(d, b, 3, device=0)
        if args.solver != "auto":
            ctx.set_solver(args.solver)
        if args.precision != "learned":
            ctx.set_precision(args.precision)
        if args.impdiff:
            ctx.set_impdiff(True)
        cells = sum(int(np.prod(np.array(ctx.shape(q, "u")) - 2)) for q in range(ctx.nblocks))
        if bfs:
            for q, fields in enumerate(bfs_ic(ctx)):
                for f, a in zip(("u", "v", "w"), fields):
                    ctx.set(q, f, a)
        else:
            u, v, w = cases.channel_ic(ctx.shape(0, "u"), ctx.metric(0, 8))
            for f, a in (("u", u), ("v", v), ("w", w)):
                ctx.set(0, f, a)
        ctx.init_bounds()
        tune_choices, tune_ms = ctx.autotune()
        dt = min(0.5 * ctx.chkdt(), ctx.dtmax)
        copy_gbs = ctx.stream_copy(1 << 27, reps=10 if args.calib else 3)
        learned = None
        if args.precision == "learned":
            ng = int(ctx.block_info(0)["n"][2])  # modes along z
            ctx.learned_collect(ng * 3 * args.learn_steps)
            t0 = time.perf_counter()
            ctx.steps(dt, args.learn_steps)
            ctx.sync()
            t1 = time.perf_counter()
            hist = ctx.learned_train(args.learn_epochs, batch=16, lr=2e-3)
            t2 = time.perf_counter()
            ctx.set_precision("learned")
            fwd_ms, gflop = ctx.learned_time(10)
            learned = {"pairs": ng * 3 * args.learn_steps, "collect_s": round(t1 - t0, 2),
                       "train_epochs": args.learn_epochs, "train_s": round(t2 - t1, 2),
                       "loss_per_epoch": [round(float(x), 4) for x in hist],
                       "network_forward_ms": round(fwd_ms, 3), "network_gflop": round(gflop, 1),
                       "roofline_net": {"bound": "mfma", "achieved": round(gflop / fwd_ms, 1), "peak": 2500.0,
                                        "unit": "TFLOP/s", "frac": round(gflop / fwd_ms / 2500.0, 4)}}
            print(f"[bench] learned: trained {args.learn_epochs} epochs, forward {fwd_ms:.3f} ms", file=sys.stderr,
                  flush=True)
        ctx.steps(dt, args.warmup)
        ctx.sync()
        print(f"[bench] warmup done, dt={dt:.3e}", file=sys.stderr, flush=True)
        # timed region: the production path, one hipGraph-captured RK3 step
        # replayed K times on the context's stream (eager launches on the
        # MG-PCG path, whose CG loop checks convergence on the host)
        t0 = time.perf_counter()
        ctx.steps(dt, args.steps)
        ctx.sync()
        el = time.perf_counter() - t0
        divtot, divmax = ctx.chkdiv()
        cg_iters, cg_relres = ctx.solver_stats()
        if learned is not None:
            learned["fallback_solves"] = ctx.learned_stats()[1]
        print(f"[bench] timed {args.steps} steps in {el:.3f} s, divmax={divmax:.2e}", file=sys.stderr, flush=True)
        # second timed pass: eager launches with HIP events around every
        # phase (on the stream the kernels run on) for the per-phase split
        # and the dominant kernel's average duration
        nph = max(1, min(args.steps, 50))
        ctx.set_timing(True)
        t1 = time.perf_counter()
        ctx.steps(dt, nph)
        ctx.sync()
        el_eager = time.perf_counter() - t1
        ph = ctx.get_timing()
        ctx.set_timing(False)
        ctx.close()

        ms_step = el * 1e3 / args.steps
        value = cells * args.steps / el / 1e6
        phases = {k: round(v[0] / nph, 4) for k, v in ph.items()}
        poisson = sum(ph[k][0] for k in ("r2r_fwd", "tridiag", "r2r_bwd")) / nph
        rk_ms, rk_n = ph["rk_mom"]
        rk_avg_s = rk_ms / rk_n * 1e-3
        achieved = RK_MOM_BYTES_PER_CELL * cells / rk_avg_s / 1e9
        traffic = None
        if not bfs and os.path.exists(TRAFFIC_FILE):
            tr = json.load(open(TRAFFIC_FILE))
            if tr.get("grid") == list(size):
                # HBM bytes per launch of the variant the autotuner chose in
                # this run (the last autotune choice is the rk_mom variant)
                traffic = tr.get("variants", {}).get(str(tune_choices[-1]), {}).get("bytes_per_launch")
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": "rk_mom (fused momentum RHS + RK3 update, autotuned variant; 104 B/cell)",
                    "rk_variant": int(tune_choices[-1]) if not bfs else None, "avg_launch_ms": round(rk_avg_s * 1e3, 4),
                    "stream_copy_GBs": round(copy_gbs, 1)}

        cpu = None
        if args.cpu_steps > 0 and not bfs and not args.impdiff:
            thr = args.cpu_threads or min(16, os.cpu_count() or 1)
            print(f"[bench] CPU baseline: {args.cpu_steps} oracle steps on {thr} threads", file=sys.stderr,
                  flush=True)
            cpu = cpu_baseline(d, b, size, args.cpu_steps, thr)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": {"fp64": "f64", "mixed": "f64+f32", "learned": "f64+f32+bf16"}[args.precision],
        "data": "synthetic (laminar Poiseuille + seeded +-10% uniform noise, seed 20251202)",
        "config": ({"workload": "plain FP64 FFT-accelerated periodic channel" +
                                ("; implicit diffusion (three Helmholtz solves per substage)" if args.impdiff else ""),
                    "grid": list(size), "stretching": "y tanh two-end, factor 1.8", "fft_axis": "z",
                    "Re_b": 5600, "forcing": "bforce(1)=12/5600",
                    "poisson": ("r2r along z and x, tridiagonal along y (exact)" if args.solver != "mg" else
                                f"DFT along z, batched MG-preconditioned CG per mode, relative residual {args.tol:g}"),
                    "parallelism": "1 GPU",
                    "autotune": {"choices": tune_choices, "best_ms": tune_ms}} if not bfs else
                   {"workload": "plain FP64 backward-facing step, 3 blocks on one GPU",
                    "blocks": [[192, 64, 128], [320, 64, 128], [320, 96, 128]], "cells": cells,
                    "stretching": "y tanh two-end per block, factor 1.2", "fft_axis": "z (128 modes)",
                    "Re": 1000, "inflow": "Poiseuille (inflow_type 1)", "outflow": "x+ velocity N, pressure D",
                    "poisson": "DFT along z, batched MG-preconditioned CG per mode (zebra y-line relaxation, "
                               f"x semi-coarsening, re-discretised coarse operators), relative residual {args.tol:g}; "
                               "the step graph runs a fixed CG count measured on the first step",
                    "cg_iterations_last_solve": cg_iters, "cg_max_relres_last_solve": cg_relres,
                    "precision": {"fp64": "FP64 throughout",
                                  "mixed": "FP32 V-cycle (smoothing, transfers, coarse solves) under flexible CG "
                                           "with FP64 vectors, operator, dot products and null-space removal",
                                  "learned": "BF16 network preconditioner (FP32 accumulation) under flexible CG "
                                             "with FP64 vectors, operator, dot products and null-space removal; "
                                             "FP32 V-cycle fallback"}[args.precision],
                    "learned": learned,
                    "parallelism": "1 GPU"}),
        "poisson_ms_per_step": round(poisson, 4),
        "phase_ms_per_step": phases,
        "phase_timing": f"eager launches with per-phase HIP events, {nph} steps, "
                        f"{el_eager * 1e3 / nph:.4f} ms/step",
        "divtot": divtot, "divmax": divmax,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

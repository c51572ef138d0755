#!/usr/bin/env python3
"""Benchmark of the SNaC time-step hot path on one MI355X.

Workload (default): the FFT-accelerated periodic channel 512x256x256 (x
periodic, y wall-normal tanh-stretched with factor 1.8, z periodic and the
FFT axis), Re_b = 5600 with the laminar body force bforce(1) = 12/5600,
laminar Poiseuille profile plus seeded uniform +-10% noise.  A step is one
full RK3 time step (3 substages: fused momentum RHS + update, ghost fills,
fillps, r2r transform along z, the per-mode 2D Helmholtz solves, inverse
transform, correc + pressure update).  Inputs are resident in HBM before
the timed region; dt is fixed from the CFL check at the start and is the
same for every form.

Forms, each timed on its own context in the same run (--forms):
  plain   the plain FP64 pipeline: batched MG-preconditioned CG per Fourier
          mode (zebra y-line relaxation, x semi-coarsening), every mode to
          the relative residual --tol (&hypre hypre_tol);
  mixed   the same solve with the V-cycle and the CG search directions in
          FP32 under a flexible CG whose residuals, operator, dot products
          and null-space removal stay FP64 (same --tol), and the momentum
          right-hand side in FP32 (RK3 accumulation and projection FP64);
  learned the plain form's FP64 stencils with the trained network
          preconditioning the flexible CG of the mixed form (FP32 V-cycle
          fallback when its residual reduction stalls; pairs recorded by the
          plain solver over --learn-steps steps, then --learn-epochs epochs
          of training; eager launches);
  direct  FP64 with the pressure solved directly (r2r along z and x,
          tridiagonal lines along y) -- the same discrete systems, not the
          plain form's method;
  direct_mixed, direct_impdiff: the direct solver with the FP32 momentum
          right-hand side / with implicit diffusion;
  impdiff, mixed_impdiff: the plain / mixed form with implicit diffusion,
          each field's Helmholtz problem solved by its own MG-PCG plan
          (eager launches: the host checks those CG loops).
The line's value is the mixed (or learned) form's Mcell-steps/s; the plain
form's figures and the ratio sit beside it, with every form's divergence.

Roofline objects: `roofline` is the dominant kernel of the headline form:
on the MG forms the one with more time per step of the CG update pass
(`roofline_update`: 40 B per unknown FP64, 44 B mixed) and the finest MG
level's zebra colour sweep (`roofline_line0`: 1.5 x unknowns x element size
per sweep -- b and x of half the lines, x of the other half), rk_mom on the
direct forms; `roofline_stencil` is the fused rk_mom of the same form (104
B/cell FP64, 80 B/cell with the FP32 right-hand side); all from HIP events
on the context's stream during an eager pass after the timed region.  cpu_baseline: the C oracle (FP64, OpenMP, direct Poisson solver)
on a bounded sample of the same channel.

--workload bfs runs the three-block backward-facing step (inlet 192x64x128
with a Poiseuille inflow, lower 320x64x128, upper 320x96x128, z periodic
with 128 modes, outflow at x+): the MG forms only (plain, mixed, learned,
impdiff, mixed_impdiff).
"""
import argparse
import json
import os
import platform
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "Mcell-steps/s + Poisson ms/step, learned/mixed vs plain FP64, equal div tol"
HBM_PEAK_GBS = 8000.0
RK_BYTES = {False: 104, True: 80}  # per cell: FP64 / FP32 right-hand side (FP32 old tendencies)
# the CG update pass per unknown: p read, x and r read and written (FP64),
# plus (FP32 preconditioners) z_old read and the FP32 copy of r written;
# p is FP32 in the mixed form
UPDATE_BYTES = {"fp64": 8 + 16 + 16, "mixed": 4 + 16 + 16 + 4 + 4, "learned": 8 + 16 + 16 + 4 + 4}
PROFILES = os.path.join(ROOT, "profiles")

FORMS = {
    # name: (solver, precision, impdiff, dtype)
    "plain": ("mg", "fp64", False, "f64"),
    "mixed": ("mg", "mixed", False, "f64+f32"),
    "learned": ("mg", "learned", False, "f64+f32+bf16"),
    "direct": ("tensor", "fp64", False, "f64"),
    "direct_mixed": ("tensor", "mixed", False, "f64+f32"),
    "direct_impdiff": ("tensor", "fp64", True, "f64"),
    "impdiff": ("mg", "fp64", True, "f64"),
    "mixed_impdiff": ("mg", "mixed", True, "f64+f32"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the metric's protocol: 100 timed steps after 10 warm-up steps
    ap.add_argument("--steps", type=int, default=100,
                    help="timed graph-replayed steps (0: none, for eager-only profiling runs)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=["channel", "bfs"], default="channel")
    ap.add_argument("--size", type=int, nargs=3, default=[512, 256, 256])
    ap.add_argument("--forms", default=None,
                    help="comma-separated forms (default: plain,mixed,direct; bfs: plain,mixed)")
    ap.add_argument("--tol", type=float, default=1e-10, help="MG-PCG relative residual per mode (&hypre hypre_tol)")
    ap.add_argument("--eager-steps", type=int, default=3, help="steps of the eager pass with per-phase HIP events")
    ap.add_argument("--mg-line-variant", type=int, default=None,
                    help="MG line relaxation kernel (0 rows per lane, 1 LDS-staged, 2 LDS-staged over modes, "
                         "3 / 4 as 2 with 16 / 8 lanes per line, -1 each level's fastest: the default)")
    ap.add_argument("--mg-fast-rcp", type=int, default=None, help="1: MG eliminations with refined reciprocals")
    ap.add_argument("--mg-warm", type=int, default=None,
                    help="1: MG-PCG pressure solves start from the previous solution (po) instead of zero")
    ap.add_argument("--mg-tail-lds", type=int, default=None,
                    help="0: the fused coarse end of the V-cycle with its vectors in global memory instead of LDS")
    ap.add_argument("--learn-steps", type=int, default=20)
    ap.add_argument("--learn-epochs", type=int, default=4)
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps for the CPU baseline (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, usable cpus)")
    return ap.parse_args()


def case_setup(tmp, workload, size, tol):
    import cases
    if workload == "bfs":
        d, b = cases.write_case(tmp, cases.bfs_blocks(), 1000.0, cfl=0.5)
    else:
        d, b = cases.write_case(tmp, [cases.channel_block(size)], 5600.0, cfl=0.5, bforce=(12.0 / 5600.0, 0.0, 0.0))
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


def set_ic(ctx, workload):
    import cases
    if workload == "bfs":
        for q, fields in enumerate(bfs_ic(ctx)):
            for f, a in zip(("u", "v", "w"), fields):
                ctx.set(q, f, a)
    else:
        u, v, w = cases.channel_ic(ctx.shape(0, "u"), ctx.metric(0, 8))
        for f, a in (("u", u), ("v", v), ("w", w)):
            ctx.set(0, f, a)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


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
    return {"value": round(cells * steps / el / 1e6, 3), "unit": "Mcell-steps/s", "cores": int(n), "kind": "port",
            "sample": f"{steps} full RK3 steps of the {size[0]}x{size[1]}x{size[2]} channel on the C oracle "
                      f"(FP64, OpenMP on {n} threads, direct Poisson solver: r2r along z and x, tridiagonal "
                      f"lines along y), {el:.1f} s",
            "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}


def box_unknowns(workload, cells):
    """Unknowns of the MG-PCG vectors: the bounding box of the union of
    blocks (the channel is its own box; the step's blocks span 512 x 160 x
    128 with the region below the inlet inactive)."""
    return (192 + 320) * (64 + 96) * 128 if workload == "bfs" else cells


def traffic_for(kernel, workload, size, form):
    """HBM bytes per launch from the committed PMC summary of this kernel,
    workload and form (FETCH_SIZE and WRITE_SIZE runs), if there is one."""
    tag = "bfs" if workload == "bfs" else "channel_" + "x".join(map(str, size))
    f = os.path.join(PROFILES, f"{kernel}_{tag}_traffic.json")
    if not os.path.exists(f):
        return None
    return json.load(open(f)).get(form, {}).get("bytes_per_launch")


def run_form(name, d, b, args, cells, dt_ref):
    import snac_amd
    solver, prec, impdiff, dtype = FORMS[name]
    ctx = snac_amd.Context(d, b, 3, device=0)
    try:
        if args.mg_line_variant is not None:
            ctx.set_mg_line_variant(args.mg_line_variant)
        if args.mg_fast_rcp is not None:
            ctx.set_mg_fast_rcp(args.mg_fast_rcp)
        if args.mg_tail_lds is not None:
            ctx.set_mg_tail_lds(args.mg_tail_lds)
        if args.mg_warm is not None:
            ctx.set_mg_warm_start(args.mg_warm)
        ctx.set_solver(solver)
        if prec != "learned":
            ctx.set_precision(prec)
        if impdiff:
            ctx.set_impdiff(True)
        set_ic(ctx, args.workload)
        ctx.init_bounds()
        dt = dt_ref if dt_ref is not None else min(0.5 * ctx.chkdt(), ctx.dtmax)
        tune_choices, tune_ms = ctx.autotune()
        learned = None
        if prec == "learned":
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
        ctx.steps(dt, args.warmup)
        ctx.sync()
        ctx.solver_counters()
        print(f"[bench] {name}: warmup done, dt={dt:.3e}", file=sys.stderr, flush=True)
        # timed region: one hipGraph-captured RK3 step replayed K times on
        # the context's stream (on the MG-PCG path the host resumes a solve
        # left above the tolerance after its captured iterations)
        t0 = time.perf_counter()
        ctx.steps(dt, args.steps)
        ctx.sync()
        el = time.perf_counter() - t0
        divtot, divmax = ctx.chkdiv()
        cg_iters, cg_relres = ctx.solver_stats()
        solves, continued, unconverged = ctx.solver_counters()
        print(f"[bench] {name}: timed {args.steps} steps in {el:.3f} s, divmax={divmax:.2e}", file=sys.stderr,
              flush=True)
        # eager pass: HIP events around every phase, on the stream the
        # kernels run on
        ne = max(1, args.eager_steps)
        ctx.set_timing(True)
        t1 = time.perf_counter()
        ctx.steps(dt, ne)
        ctx.sync()
        el_eager = time.perf_counter() - t1
        ph = ctx.get_timing()
        ctx.set_timing(False)
        out = {"solver": solver, "precision": prec, "dtype": dtype, "impdiff": impdiff,
               # --steps 0 (eager-only profiling runs): no timed graph steps
               "mcell_steps_per_s": round(cells * args.steps / el / 1e6, 2) if args.steps > 0 else None,
               "ms_per_step": round(el * 1e3 / args.steps, 4) if args.steps > 0 else None,
               "poisson_ms_per_step": round(sum(ph[k][0] for k in ("r2r_fwd", "tridiag", "r2r_bwd")) / ne, 4),
               "phase_ms_per_step": {k: round(v[0] / ne, 4) for k, v in ph.items() if v[1] > 0},
               "phase_timing": f"eager launches with per-phase HIP events, {ne} steps, "
                               f"{el_eager * 1e3 / ne:.4f} ms/step",
               "divtot": divtot, "divmax": divmax, "dt": dt}
        if solver == "mg":
            out.update({"hypre_tol": args.tol, "cg_iterations_last_solve": cg_iters,
                        "cg_max_relres_last_solve": cg_relres, "solves_timed": solves,
                        "solves_resumed_after_graph": continued, "solves_at_maxiter": unconverged})
        if solver == "mg" and impdiff:
            out["helmholtz_cg_iterations_last_solve"] = {n: ctx.impdiff_stats(n)[0] for n in "uvw"}
        if learned is not None:
            learned["fallback_solves"] = ctx.learned_stats()[1]
            out["learned"] = learned
        # rk_mom: FP64 or FP32 right-hand side
        f32 = prec != "fp64"
        rk_ms, rk_n = ph["rk_mom"]
        rk_avg = rk_ms / max(rk_n, 1)
        rkv = tune_choices[-1] if f32 else tune_choices[-2]
        out["roofline_stencil"] = {
            "bound": "hbm", "achieved": round(RK_BYTES[f32] * cells / (rk_avg * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(RK_BYTES[f32] * cells / (rk_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": traffic_for("rk_mom", args.workload, args.size, name),
            "kernel": f"rk_mom ({'FP32' if f32 else 'FP64'} right-hand side, variant {rkv}; "
                      f"{RK_BYTES[f32]} B/cell)", "avg_launch_ms": round(rk_avg, 4)}
        if solver == "mg" and ph["mg_tail"][1] > 0:
            out["mg_tail_avg_ms"] = round(ph["mg_tail"][0] / ph["mg_tail"][1], 4)
        if solver == "mg":
            # the MG vectors span the bounding box of the union of blocks
            unk = box_unknowns(args.workload, cells)
            up_ms, up_n = ph["mg_update"]
            if up_n > 0:
                avg = up_ms / up_n
                bytes_ = UPDATE_BYTES[prec] * unk
                out["roofline_update"] = {
                    "bound": "hbm", "achieved": round(bytes_ / (avg * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(bytes_ / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "traffic": traffic_for("k_update_apply_dot", args.workload, args.size, name),
                    "kernel": f"k_update_apply_dot, the CG update pass ({UPDATE_BYTES[prec]} B per unknown)",
                    "avg_launch_ms": round(avg, 4), "launches": up_n, "ms_per_step": round(up_ms / ne, 4)}
            l0_ms, l0_n = ph["mg_line0"]
            if l0_n > 0:
                esz = 4 if f32 else 8
                avg = l0_ms / l0_n
                bytes_ = 1.5 * unk * esz
                out["roofline_line0"] = {
                    "bound": "hbm", "achieved": round(bytes_ / (avg * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(bytes_ / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "traffic": traffic_for("k_line", args.workload, args.size, name),
                    "kernel": f"k_line, finest level, one zebra colour ({'FP32' if f32 else 'FP64'}; "
                              f"1.5 x {esz} B per unknown)", "avg_launch_ms": round(avg, 4),
                    "launches": l0_n, "ms_per_step": round(l0_ms / ne, 4)}
            # the dominant MG kernel: the one with more time per step
            cand = [k for k in ("roofline_update", "roofline_line0") if k in out]
            if cand:
                out["roofline_dominant"] = max(cand, key=lambda k: out[k]["ms_per_step"])
        out["autotune"] = {"choices": tune_choices, "best_ms": tune_ms}
        if solver == "mg":
            l64, l32 = ctx.mg_line_choices()
            out["mg_line_variants"] = {"fp64": l64, "fp32": l32}
        return out
    finally:
        ctx.close()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the time step runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding the blocks over GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import snac_amd

    bfs = args.workload == "bfs"
    forms = (args.forms or ("plain,mixed" if bfs else "plain,mixed,direct")).split(",")
    for f in forms:
        if f not in FORMS:
            sys.exit(f"bench.py: unknown form {f!r} (forms: {', '.join(FORMS)})")
        if bfs and FORMS[f][0] == "tensor":
            sys.exit(f"bench.py: form {f!r} needs a box-shaped case; the backward-facing step is not one")
    size = args.size
    results = {}
    with tempfile.TemporaryDirectory() as tmp:
        d, b = case_setup(tmp, args.workload, size, args.tol)
        probe = snac_amd.Context(d, b, 3, device=-1)  # host-only: the case's cell count
        cells = sum(int(np.prod(np.array(probe.shape(q, "u")) - 2)) for q in range(probe.nblocks))
        probe.close()
        copy_gbs = None
        dt_ref = None
        for name in forms:
            print(f"[bench] {args.workload} {'' if bfs else size}: form {name}", file=sys.stderr, flush=True)
            results[name] = run_form(name, d, b, args, cells, dt_ref)
            dt_ref = results[name]["dt"]
        if not bfs:
            # the plain copy kernel: the practical HBM ceiling on this part
            ctx = snac_amd.Context(d, b, 3, device=0)
            copy_gbs = round(ctx.stream_copy(1 << 27, reps=5), 1)
            ctx.close()
        cpu = None
        if args.cpu_steps > 0 and not bfs:
            thr = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
            print(f"[bench] CPU baseline: {args.cpu_steps} oracle steps on {thr} threads", file=sys.stderr,
                  flush=True)
            cpu = cpu_baseline(d, b, size, args.cpu_steps, thr)

    head = next((f for f in ("learned", "mixed") if f in results), forms[0])
    H = results[head]
    plain = results.get("plain")
    roof = H[H["roofline_dominant"]] if "roofline_dominant" in H else H["roofline_stencil"]
    out = {
        "metric": METRIC, "value": H["mcell_steps_per_s"], "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": H["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": H["dtype"],
        "data": "synthetic (laminar Poiseuille + seeded +-10% uniform noise, seed 20251202)",
        "config": ({"workload": f"FFT-accelerated periodic channel {size[0]}x{size[1]}x{size[2]}, form '{head}'"
                                + (" against the plain FP64 form" if plain and head != "plain" else ""),
                    "grid": list(size), "stretching": "y tanh two-end, factor 1.8", "fft_axis": "z",
                    "Re_b": 5600, "forcing": "bforce(1)=12/5600", "hypre_tol": args.tol,
                    "parallelism": "1 GPU"} if not bfs else
                   {"workload": f"backward-facing step, 3 blocks on one GPU, form '{head}'"
                                + (" against the plain FP64 form" if plain and head != "plain" else ""),
                    "blocks": [[192, 64, 128], [320, 64, 128], [320, 96, 128]], "cells": cells,
                    "stretching": "y tanh two-end per block, factor 1.2", "fft_axis": "z (128 modes)",
                    "Re": 1000, "inflow": "Poiseuille (inflow_type 1)", "outflow": "x+ velocity N, pressure D",
                    "hypre_tol": args.tol, "parallelism": "1 GPU"}),
        "poisson_ms_per_step": H["poisson_ms_per_step"],
        "vs_plain_fp64": (round(H["mcell_steps_per_s"] / plain["mcell_steps_per_s"], 3)
                          if plain and head != "plain" and args.steps > 0 else None),
        "poisson_vs_plain_fp64": (round(plain["poisson_ms_per_step"] / H["poisson_ms_per_step"], 3)
                                  if plain and head != "plain" else None),
        "divtot": H["divtot"], "divmax": H["divmax"],
        "roofline": dict(roof, stream_copy_GBs=copy_gbs),
        "roofline_stencil": H["roofline_stencil"],
        "roofline_update": H.get("roofline_update"),
        "roofline_line0": H.get("roofline_line0"),
        "cpu_baseline": cpu,
        "forms": results,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

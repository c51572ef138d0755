"""The learned preconditioner on the multi-block backward-facing step (the
learned-preconditioner configuration's geometry: inlet, lower and upper
blocks, z periodic, the per-mode problems on an L-shaped union), at a size
the oracle checks in seconds: inlet 24x16, lower 40x16, upper 40x24, 16
modes.

* pairs are recorded from the plain FP64 MG-PCG solves of 8 steps (384
  pairs), the network is trained on the GPU, then 10 steps with the network
  preconditioning the flexible CG (V-cycle fallback on stall) run from the
  same fields as the oracle (direct banded solves per mode): every field
  within 1e-10 after 10 steps, every solve stopped at hypre_tol = 1e-13;
* the CG iterations and solve times of plain FP64, mixed and learned are
  printed side by side (the comparison with the method without the
  network); the network alone (no fallback) must reach a lower residual or
  need fewer iterations once trained than untrained -- what training gains.
"""
import os
import time

import numpy as np
import pytest

import cases
import nmlparse
import oracle_py
from test_gpu_parity import FIELDS3, assert_fields, run_both

pytestmark = pytest.mark.gpu


def bfs(tmp_path):
    d, b = cases.write_case(str(tmp_path), cases.bfs_blocks((24, 16), (40, 16, 24), 16), 1000.0, cfl=0.5)
    return d, b


def set_ic(ctx, oc=None, seed=11):
    rng = np.random.default_rng(seed)
    for q in range(ctx.nblocks):
        yc = ctx.metric(q, 8)
        for f in FIELDS3:
            a = np.zeros(ctx.shape(q, f))
            if f == "u" and yc[1:-1].max() <= 1.0:
                a[1:-1, 1:-1, 1:-1] = (6.0 * yc[1:-1] * (1.0 - yc[1:-1]))[None, :, None]
            a[1:-1, 1:-1, 1:-1] += 0.1 * rng.uniform(-1, 1, tuple(s - 2 for s in a.shape))
            ctx.set(q, f, a)
            if oc is not None:
                oc.set(q, f, a)


def timed_solves(ctx, dt, nsteps):
    """nsteps eager steps: (CG iterations of the last solve, ms per step)"""
    ctx.set_graph(False)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        ctx.step(dt)
    ctx.sync()
    return ctx.solver_stats()[0], (time.perf_counter() - t0) * 1e3 / nsteps


def test_learned_bfs_against_oracle(tmp_path):
    import snac_amd
    d, b = bfs(tmp_path)
    print("learned: backward-facing step 24x16 + 40x16 + 40x24, 16 modes", flush=True)
    ctx = snac_amd.Context(d, b, 3)
    ctx.set_solver("mg")
    set_ic(ctx)
    ctx.init_bounds()
    dt = 0.5 * ctx.chkdt()
    ctx.learned_collect(16 * 3 * 8)
    ctx.steps(dt, 8)
    assert ctx.learned_stats()[0] == 16 * 3 * 8
    hist = ctx.learned_train(epochs=30, batch=16, lr=2e-3)
    print("training loss per epoch", np.round(hist, 4), flush=True)
    assert np.all(np.isfinite(hist)) and hist[-1] < 0.5 * hist[0]
    # the learned solver from fresh fields, against the oracle
    oc = oracle_py.OracleCase(nmlparse.load_case(d, b, 3))
    oc.set_solver("band")
    set_ic(ctx, oc)
    for f in ("p", "pp", "dudtrko", "dvdtrko", "dwdtrko"):
        for q in range(ctx.nblocks):
            ctx.set(q, f, np.zeros(ctx.shape(q, f)))
    ctx.set_precision("learned")
    run_both(ctx, oc, 10)
    it, rel = ctx.solver_stats()
    assert rel <= 1e-13 * 1.0001
    assert_fields(ctx, oc, FIELDS3, 1e-10, interior=True)
    assert_fields(ctx, oc, ("p",), 1e-10, interior=True)
    # side by side from the same fields: plain FP64, mixed, learned
    snap = {(q, f): ctx.get(q, f) for q in range(ctx.nblocks)
            for f in ("u", "v", "w", "p", "pp", "dudtrko", "dvdtrko", "dwdtrko")}
    res = {}
    for prec in ("fp64", "mixed", "learned"):
        for (q, f), a in snap.items():
            ctx.set(q, f, a)
        ctx.set_precision(prec)
        res[prec] = timed_solves(ctx, dt, 3)
    ctx.learned_options(fallback=False)
    for (q, f), a in snap.items():
        ctx.set(q, f, a)
    ctx.step(dt)
    res["network only"] = ctx.solver_stats()
    for k, v in res.items():
        print(f"{k:>13}: {v}", flush=True)
    fallbacks = ctx.learned_stats()[1]
    print(f"solves that fell back to the V-cycle: {fallbacks}", flush=True)
    trained = res["network only"]
    ctx.close()
    # the same network untrained
    ctx2 = snac_amd.Context(d, b, 3)
    ctx2.set_solver("mg")
    set_ic(ctx2)
    ctx2.init_bounds()
    ctx2.learned_collect(16)
    ctx2.steps(dt, 1)
    ctx2.learned_train(epochs=1, batch=16, lr=0.0)
    for (q, f), a in snap.items():
        ctx2.set(q, f, a)
    ctx2.set_precision("learned")
    ctx2.learned_options(fallback=False)
    ctx2.set_graph(False)
    ctx2.step(dt)
    untrained = ctx2.solver_stats()
    print(f"network only, untrained: {untrained} (trained: {trained})", flush=True)
    # a preconditioner that breaks the CG down leaves a NaN residual: the worst outcome
    un = (untrained[0], np.inf if np.isnan(untrained[1]) else untrained[1])
    assert not np.isnan(trained[1])
    assert (trained[0], trained[1]) < un
    ctx2.close()

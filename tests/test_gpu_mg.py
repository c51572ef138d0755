"""Parity of the batched MG-preconditioned CG Poisson path (unions of blocks
that are not boxes, and any case when forced) with the CPU oracle's direct
banded solver, through the C-ABI.

The oracle solves every per-mode system exactly; the GPU path iterates each
mode to ||r|| <= 1e-13 ||b|| (hypre_tol of the test inputs), so the pressure
correction agrees to ~1e-11 and the fields after 10 steps within the 1e-10
gate.  p / pp of singular (all-Neumann or periodic) problems are compared
modulo their mean.

Both precisions run every case.  A "mixed" solve (FP32 V-cycle and search
directions under a flexible CG with FP64 residuals, operator and dot
products) stops at the same FP64 residual tolerance, so its single solves
are held to the same bounds; its 10-step runs also carry the FP32 momentum
right-hand side and are held to the mixed form's 1e-6.
"""
import os

import numpy as np
import pytest

import cases
import nmlparse
import oracle_py
from test_gpu_parity import FIELDS3, GOLD, assert_fields, run_both, seed_fields

pytestmark = pytest.mark.gpu


def pair(tmp_path, geom, prec="fp64"):
    import snac_amd
    fa, singular = 3, False
    if geom == "lchannel":
        blocks, visci = cases.lchannel_blocks(8, 16), 1000.0
    elif geom == "bfs":
        blocks, visci = cases.bfs_blocks((12, 8), (20, 8, 12), 8), 1000.0
    elif geom == "channel":
        blocks, visci, singular = [cases.channel_block((32, 24, 16))], 5600.0, True
    elif geom in ("channel_y160", "channel_y256", "channel_y128"):
        # long wall-normal lines, and more than 32 columns, so that the finest
        # levels are relaxed by the launched line kernels, not the fused tail
        shape = {"channel_y160": (72, 160, 8), "channel_y256": (64, 256, 4), "channel_y128": (96, 128, 6)}[geom]
        blocks, visci, singular = [cases.channel_block(shape)], 5600.0, True
    elif geom == "cavity3d":
        blocks, visci, fa, singular = [cases.cavity_block((12, 10, 16), gr=(0.0, 1.5, 0.0))], 1000.0, 0, True
    elif geom in ("ldc3d", "ldc_fft"):
        fa = 0 if geom == "ldc3d" else 1
        # the reference's LDC inputs with a tight &hypre tolerance appended
        # (the file itself leaves the defaults, 1e-4 / 50 iterations)
        bn = os.path.join(GOLD, "ldc", "blocks.nml")
        dn = str(tmp_path / "dns.nml")
        with open(os.path.join(GOLD, "ldc", "dns.nml")) as f:
            txt = f.read()
        with open(dn, "w") as f:
            f.write(txt + "&hypre\nhypre_solver_i = 2, hypre_tol = 1.e-13, hypre_maxiter = 1000\n/\n")
        print(f"case {geom}: fft_axis={fa}", flush=True)
        ctx = snac_amd.Context(dn, bn, fa)
        ctx.set_solver("mg")
        ctx.set_precision(prec)
        oc = oracle_py.OracleCase(nmlparse.load_case(dn, bn, fa))
        oc.set_solver("band")
        return ctx, oc, True
    else:
        raise ValueError(geom)
    d, b = cases.write_case(str(tmp_path), blocks, visci)
    print(f"case {geom}: fft_axis={fa} blocks={[B['ng'] for B in blocks]}", flush=True)
    ctx = snac_amd.Context(d, b, fa)
    ctx.set_solver("mg")
    ctx.set_precision(prec)
    oc = oracle_py.OracleCase(nmlparse.load_case(d, b, fa))
    oc.set_solver("band")
    return ctx, oc, singular


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("geom", ["lchannel", "bfs", "channel", "cavity3d", "ldc3d", "ldc_fft"])
def test_mg_solve_pressure(tmp_path, geom, prec):
    ctx, oc, singular = pair(tmp_path, geom, prec)
    seed_fields(ctx, oc)
    ctx.bounduvw(False); oc.bounduvw(False)
    ctx.fillps(1e-3); oc.fillps(1e-3)
    ctx.solve_pressure()
    oc.solve_pressure()
    iters, relres = ctx.solver_stats()
    print(f"{geom} {prec}: {iters} CG iterations, max relative residual {relres:.2e}", flush=True)
    assert 0 < iters < 100
    assert relres <= 1e-13
    assert_fields(ctx, oc, ("pp",), 1e-10, interior=True, modmean=("pp",) if singular else ())


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_zero_modes_after_nonzero_solve(tmp_path, prec):
    """A solve whose right-hand side has no z variation (every Fourier mode
    but the first is zero, so those modes are converged before the first CG
    iteration) right after a solve with all modes nonzero, and then a solve
    of a zero right-hand side (every mode done before the first iteration):
    the first CG update pass, which writes x without reading it, sets the
    modes that are already done to zero rather than leaving the previous
    solve's values; the solves against the oracle."""
    ctx, oc, singular = pair(tmp_path, "channel", prec)
    rng = np.random.default_rng(8)
    for rhs in ("3d", "2d", "zero"):
        if rhs == "zero":
            for f in ("u", "v", "w"):
                ctx.set(0, f, np.zeros(ctx.shape(0, f)))
            ctx.bounduvw(False)
            ctx.fillps(1e-3)
            ctx.solve_pressure()
            assert not ctx.get(0, "pp")[1:-1, 1:-1, 1:-1].any()
            break
        for f in ("u", "v", "w"):
            a = rng.uniform(-1, 1, ctx.shape(0, f))
            if rhs == "2d":
                a[:, :, :] = a[:, :, :1]  # the same plane at every z
            ctx.set(0, f, a)
            oc.set(0, f, a)
        ctx.bounduvw(False); oc.bounduvw(False)
        ctx.fillps(1e-3); oc.fillps(1e-3)
        ctx.solve_pressure()
        oc.solve_pressure()
        iters, relres = ctx.solver_stats()
        print(f"{prec} {rhs}: {iters} CG iterations, max relative residual {relres:.2e}", flush=True)
        assert relres <= 1e-13
        assert_fields(ctx, oc, ("pp",), 1e-10, interior=True, modmean=("pp",) if singular else ())
        if rhs == "2d":  # no z variation left in the correction
            pp = ctx.get(0, "pp")[1:-1, 1:-1, 1:-1]
            assert np.abs(pp - pp[:, :, :1]).max() <= 1e-12 * np.abs(pp).max()


def test_mg_matches_tensor_solver(tmp_path):
    """On a box both GPU solvers apply: the iterative one reproduces the
    direct one (stretched wall-normal axis, periodic coarsened axis,
    singular zero mode)."""
    import snac_amd
    d, b = cases.write_case(str(tmp_path), [cases.channel_block((32, 24, 16))], 5600.0)
    a = snac_amd.Context(d, b, 3)
    m = snac_amd.Context(d, b, 3)
    m.set_solver("mg")
    rng = np.random.default_rng(3)
    for f in ("u", "v", "w"):
        x = rng.uniform(-1, 1, a.shape(0, f))
        a.set(0, f, x)
        m.set(0, f, x)
    for ctx in (a, m):
        ctx.bounduvw(False)
        ctx.fillps(1e-3)
        ctx.solve_pressure()
    ga, gm = a.get(0, "pp")[1:-1, 1:-1, 1:-1], m.get(0, "pp")[1:-1, 1:-1, 1:-1]
    ga, gm = ga - ga.mean(), gm - gm.mean()
    assert np.linalg.norm(ga - gm) / np.linalg.norm(ga) < 1e-10


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("geom", ["lchannel", "bfs", "cavity3d"])
def test_mg_10_steps(tmp_path, geom, prec):
    """FP64: the plain form's 1e-10; mixed (FP32 momentum right-hand side,
    FP32 V-cycle and search directions): the mixed form's 1e-6 gate.  The
    cavity has no FFT axis: its solves run the coupled 3D MG-PCG path."""
    ctx, oc, singular = pair(tmp_path, geom, prec)
    tol = 1e-10 if prec == "fp64" else 1e-6
    run_both(ctx, oc, 10, dt_rel=1e-12 if prec == "fp64" else 1e-6)
    if prec == "fp64":
        assert_fields(ctx, oc, FIELDS3, tol, interior=True)
    else:
        # the FP32 rounding is absolute, about 1e-7 of |u|: each component's
        # error against the size of the whole velocity field (w of these
        # cases is a small cross-flow)
        inner = (slice(1, -1),) * 3
        for b in range(ctx.nblocks):
            scale = np.sqrt(sum(np.linalg.norm(oc.get(b, f)[inner]) ** 2 for f in FIELDS3))
            for f in FIELDS3:
                err = np.linalg.norm(ctx.get(b, f)[inner] - oc.get(b, f)[inner])
                assert err <= tol * scale, (b, f, err / scale)
    assert_fields(ctx, oc, ("p",), tol, interior=True, modmean=("p",) if singular else ())
    _, divmax = ctx.chkdiv()
    _, odivmax = oc.chkdiv()
    print(f"{geom} {prec}: divmax GPU {divmax:.2e} oracle {odivmax:.2e}", flush=True)
    assert divmax < 1e-9


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_warm_start_steps(tmp_path, prec):
    """The warm start (each pressure solve from the previous solution, as the
    reference's solves start from po, solver.f90:866-876; the oracle does
    the same): 10 steps on the step, every solve at hypre_tol, against the
    oracle at the plain form's 1e-10 (FP64; the mixed form's FP32 momentum
    right-hand side is checked by the 10-step tests above), and graph steps
    equal to eager steps bit for bit."""
    ctx, oc, _ = pair(tmp_path, "bfs", prec)
    ctx.set_mg_warm_start(True)
    run_both(ctx, oc, 10, dt_rel=1e-12 if prec == "fp64" else 1e-6)
    if prec == "fp64":
        assert_fields(ctx, oc, FIELDS3, 1e-10, interior=True)
        assert_fields(ctx, oc, ("p",), 1e-10, interior=True)
    it, relres = ctx.solver_stats()
    print(f"warm start {prec}: last solve {it} CG iterations, relres {relres:.2e}", flush=True)
    assert 0 < it and relres <= 1e-13 * 1.0001
    ctx.close()
    import snac_amd
    d, b = cases.write_case(str(tmp_path / "g"), cases.bfs_blocks((12, 8), (20, 8, 12), 8), 1000.0)
    eager, graph = snac_amd.Context(d, b, 3), snac_amd.Context(d, b, 3)
    eager.set_graph(False)
    rng = np.random.default_rng(9)
    for c in (eager, graph):
        c.set_precision(prec)
        c.set_mg_warm_start(True)
    for q in range(eager.nblocks):
        for f in FIELDS3:
            a = 0.05 * rng.uniform(-1, 1, eager.shape(q, f))
            eager.set(q, f, a)
            graph.set(q, f, a)
    for c in (eager, graph):
        c.init_bounds()
        c.steps(2e-3, 4)
    for q in range(eager.nblocks):
        for f in FIELDS3 + ("p",):
            np.testing.assert_array_equal(graph.get(q, f), eager.get(q, f), err_msg=f"block {q} {f}")
    eager.close()
    graph.close()


def test_precision_switch_keeps_results(tmp_path):
    """Switching precision on a context re-measures the graph's CG count;
    a mixed solve and an FP64 solve of the same rhs agree to the tolerance."""
    ctx, oc, _ = pair(tmp_path, "bfs", "fp64")
    seed_fields(ctx, oc)
    ctx.bounduvw(False)
    ctx.fillps(1e-3)
    rhs = [ctx.get(b, "pp") for b in range(ctx.nblocks)]
    ctx.solve_pressure()
    ref = [ctx.get(b, "pp")[1:-1, 1:-1, 1:-1] for b in range(ctx.nblocks)]
    ctx.set_precision("mixed")
    for b in range(ctx.nblocks):
        ctx.set(b, "pp", rhs[b])
    ctx.solve_pressure()
    for b in range(ctx.nblocks):
        g = ctx.get(b, "pp")[1:-1, 1:-1, 1:-1]
        assert np.linalg.norm(g - ref[b]) / np.linalg.norm(ref[b]) < 1e-10


@pytest.mark.parametrize("captured", [0, 2])
def test_graph_steps_meet_tolerance(tmp_path, captured):
    """Graph steps on the MG-PCG path end every solve at hypre_tol.  With the
    captured CG count measured (0) and with it forced far below what the
    solves need (2), every solve of every step is resumed by the host until
    it converges, no solve stops above the tolerance, and the fields agree
    bit for bit with eager steps; the right-hand side is then made harder (a
    rough velocity field) between replays and the solves still converge."""
    import snac_amd
    d, b = cases.write_case(str(tmp_path), cases.bfs_blocks((12, 8), (20, 8, 12), 8), 1000.0)
    print(f"case bfs graph captured={captured}", flush=True)
    eager, graph = snac_amd.Context(d, b, 3), snac_amd.Context(d, b, 3)
    eager.set_graph(False)
    graph.set_graph_cg_iters(captured)
    rng = np.random.default_rng(5)
    for q in range(eager.nblocks):
        for f in FIELDS3:
            a = 0.05 * rng.uniform(-1, 1, eager.shape(q, f))
            eager.set(q, f, a)
            graph.set(q, f, a)
    for ctx in (eager, graph):
        ctx.init_bounds()
        ctx.solver_counters()
    dt = 2e-3
    for rep in range(4):
        if rep == 2:  # a much rougher field between replays
            for q in range(eager.nblocks):
                for f in FIELDS3:
                    a = eager.get(q, f) + rng.uniform(-1, 1, eager.shape(q, f))
                    eager.set(q, f, a)
                    graph.set(q, f, a)
        eager.step(dt)
        graph.step(dt)
        solves, resumed, unconverged = graph.solver_counters()
        it, relres = graph.solver_stats()
        print(f"step {rep}: {solves} solves, {resumed} resumed, {unconverged} unconverged, last {it} its "
              f"relres {relres:.2e}", flush=True)
        assert solves == 3 and unconverged == 0 and relres <= 1e-13
        if captured == 2 and rep > 0:
            assert resumed == 3
        assert eager.solver_counters()[2] == 0
        # the count a graph solve reports is the one it needed (the iteration
        # at which every mode met the tolerance), not the captured count
        assert it == eager.solver_stats()[0], (it, eager.solver_stats())
        for q in range(eager.nblocks):
            for f in FIELDS3 + ("p",):
                np.testing.assert_array_equal(graph.get(q, f), eager.get(q, f), err_msg=f"step {rep} block {q} {f}")
    eager.close()
    graph.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("geom", ["bfs", "cavity3d"])
def test_line_variants_agree(tmp_path, geom, prec):
    """The line relaxation kernels -- each lane reading its rows from memory,
    the rows staged through LDS, and the lines of a workgroup over several
    modes sharing staged coefficients -- compute the same values: the
    pressure solves agree bit for bit (FFT path, and the coupled 3D path,
    where the third falls back to the second) when the mode-batched kernels
    divide by IEEE quotients as the others do."""
    out = []
    try:
        for v in (0, 1, 2):
            ctx, oc, _ = pair(tmp_path / f"v{v}", geom, prec)
            ctx.set_mg_fast_rcp(False)  # process-wide
            ctx.set_mg_line_variant(v)
            seed_fields(ctx, oc)
            ctx.bounduvw(False)
            ctx.fillps(1e-3)
            ctx.solve_pressure()
            out.append([ctx.get(b, "pp") for b in range(ctx.nblocks)])
            ctx.set_mg_line_variant(-1)
            ctx.close()
    finally:
        restore_fast_rcp(tmp_path)
    for v in (1, 2):
        for a, b in zip(out[0], out[v]):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("geom", ["channel_y160", "channel_y256"])
def test_line_variants_long_lines(tmp_path, geom, prec):
    """Lines of 160 and 256 rows, where every line kernel runs as itself:
    the 32-lanes-per-line kernels (0..2) agree bit for bit, and the
    16- and 8-lanes-per-line mode-batched kernels (3, 4: the same systems
    in another elimination order) solve to the same tolerance; every
    variant's pressure correction is checked against the oracle's direct
    banded solve (the bit-for-bit comparison with IEEE quotients in every
    kernel)."""
    out = {}
    try:
        run_long_line_variants(tmp_path, geom, prec, out)
    finally:
        restore_fast_rcp(tmp_path)
    for v in (1, 2):
        np.testing.assert_array_equal(out[0], out[v])
    for v in (3, 4):
        print(f"variant {v}: max |pp - pp(variant 0)| = {np.abs(out[v] - out[0]).max():.2e}", flush=True)


def restore_fast_rcp(tmp_path):
    """the process-wide refined reciprocals of the mode-batched line kernels
    back to their default (on)"""
    import snac_amd
    c = snac_amd.Context(*cases.write_case(str(tmp_path / "rcp"), [cases.channel_block((4, 4, 4))], 5600.0), 3)
    c.set_mg_fast_rcp(True)
    c.close()


def run_long_line_variants(tmp_path, geom, prec, out):
    for v in (0, 1, 2, 3, 4):
        ctx, oc, singular = pair(tmp_path / f"v{v}", geom, prec)
        ctx.set_mg_fast_rcp(False)  # process-wide
        ctx.set_mg_line_variant(v)
        seed_fields(ctx, oc)
        ctx.bounduvw(False); oc.bounduvw(False)
        ctx.fillps(1e-3); oc.fillps(1e-3)
        ctx.solve_pressure()
        oc.solve_pressure()
        iters, relres = ctx.solver_stats()
        print(f"{geom} {prec} line variant {v}: {iters} CG iterations, max relative residual {relres:.2e}",
              flush=True)
        assert 0 < iters < 100
        assert relres <= 1e-13
        assert_fields(ctx, oc, ("pp",), 1e-10, interior=True, modmean=("pp",) if singular else ())
        out[v] = ctx.get(0, "pp")
        ctx.set_mg_line_variant(-1)
        ctx.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("geom", ["channel_y128", "channel_y256", "bfs", "lchannel"])
def test_cg_lds_agrees(tmp_path, geom, prec):
    """The CG passes with a workgroup's search directions staged in LDS --
    whole lines with row coefficients (lines of 128 and 256 rows; x
    periodic, so the first and last lines take their wrapped neighbours),
    and chunks that start inside a line with per-cell coefficients (the
    step's 160-row lines, the L-channel's union) -- compute what the forms
    reading every neighbour through the caches compute, bit for bit: 3
    steps each, against the oracle and against each other."""
    out = []
    try:
        for on in (False, True):
            ctx, oc, singular = pair(tmp_path / f"c{int(on)}", geom, prec)
            ctx.set_mg_cg_lds(on)
            seed_fields(ctx, oc, names=("u", "v", "w"))
            run_both(ctx, oc, 3, dt_rel=1e-12 if prec == "fp64" else 1e-6)
            if prec == "fp64":
                assert_fields(ctx, oc, FIELDS3, 1e-10, interior=True)
            it, relres = ctx.solver_stats()
            print(f"{geom} {prec} cg lds {on}: {it} CG iterations, relres {relres:.2e}", flush=True)
            assert relres <= 1e-13 * 1.0001
            out.append({(q, f): ctx.get(q, f) for q in range(ctx.nblocks) for f in FIELDS3 + ("p",)})
            ctx.close()
    finally:
        import snac_amd
        c = snac_amd.Context(*cases.write_case(str(tmp_path / "r"), [cases.channel_block((4, 4, 4))], 5600.0), 3)
        c.set_mg_cg_lds(True)  # process-wide: back to the default
        c.close()
    for f in out[0]:
        np.testing.assert_array_equal(out[1][f], out[0][f], err_msg=f)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("geom", ["bfs", "channel_y256"])
def test_tail_lds_agrees(tmp_path, geom, prec):
    """The fused coarse end with its vectors resident in LDS computes what
    the global-memory form computes, bit for bit (where the vectors do not
    fit, both runs take the global form)."""
    out = []
    for lds in (False, True):
        ctx, oc, singular = pair(tmp_path / f"t{int(lds)}", geom, prec)
        ctx.set_mg_tail_lds(lds)
        seed_fields(ctx, oc)
        ctx.bounduvw(False)
        ctx.fillps(1e-3)
        ctx.solve_pressure()
        iters, relres = ctx.solver_stats()
        print(f"{geom} {prec} tail in LDS {lds}: {iters} CG iterations, max relative residual {relres:.2e}",
              flush=True)
        out.append([ctx.get(b, "pp") for b in range(ctx.nblocks)])
        ctx.set_mg_tail_lds(True)
        ctx.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)

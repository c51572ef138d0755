"""Parity of the scalar-transport path (rk_scal with bulk forcing, scalar
boundp, the Boussinesq term of rk_mom) with the CPU oracle, through the
C-ABI, plus the reference's differentially heated cavity (tests/
differentially_heated_cavity, inputs in tests/golden/dhc)."""
import os

import numpy as np
import pytest

import cases
import nmlparse
import oracle_py
from test_gpu_parity import GOLD, assert_fields, run_both

pytestmark = pytest.mark.gpu

SCAL_FIELDS = ("u", "v", "w", "p", "s1", "dsdtrko1", "s2", "dsdtrko2")


def scalar_pair(tmp_path, forced=False, fa=0, gacc=(0.3, -0.2, -1.0), beta=0.7, buoyancy=True):
    import snac_amd
    blk = cases.cavity_block((12, 10, 8), gr=(0.0, 1.4, 0.0))
    blk["cbcscal"] = [[["D", "N"], ["N", "D"], ["N", "N"]], [["N", "N"], ["D", "D"], ["N", "N"]]]
    blk["bcscal"] = [[[-0.5, 0.1], [0.0, 0.25], [0.0, 0.0]], [[0.0, 0.0], [1.0, 2.0], [0.0, 0.0]]]
    d, b = cases.write_case(str(tmp_path), [blk], 500.0, nscal=2, gacc=gacc,
                            scalar=dict(alphai=[700.0, 300.0], beta=beta, ssource=[0.3, -0.1],
                                        is_sforced=[forced, False], scalf=[0.4, 0.0]))
    print(f"case scalars forced={forced} fa={fa}", flush=True)
    ctx = snac_amd.Context(d, b, fa)
    oc = oracle_py.OracleCase(nmlparse.load_case(d, b, fa))
    oc.set_solver("band")
    if buoyancy:  # the reference's BOUSSINESQ_BUOYANCY=1 build
        ctx.set_buoyancy(True)
        oc.set_buoyancy(True)
    return ctx, oc


def seed(ctx, oc, seed=3, names=SCAL_FIELDS):
    rng = np.random.default_rng(seed)
    for f in names:
        a = rng.uniform(-1.0, 1.0, ctx.shape(0, f))
        ctx.set(0, f, a)
        oc.set(0, f, a)


def interior(a):
    return a[1:-1, 1:-1, 1:-1] if a.ndim == 3 and a.shape[0] > 2 else a


@pytest.mark.parametrize("forced", [False, True])
def test_rk_scal(tmp_path, forced):
    ctx, oc = scalar_pair(tmp_path, forced=forced)
    for irk in (1, 2, 3):
        seed(ctx, oc, seed=irk)
        ctx.rk_scal(irk, 2e-3)
        fo = oc.rk_scal(irk, 2e-3)
        for f in ("s1", "dsdtrko1", "s2", "dsdtrko2"):
            g, o = ctx.get(0, f), oc.get(0, f)
            if g.shape[0] == ctx.shape(0, "s1")[0]:
                g, o = interior(g), interior(o)  # new buffer: ghosts come from boundp
            r = np.linalg.norm(g - o) / np.linalg.norm(o)
            assert r <= 1e-13, f"irk {irk} {f}: {r:.3e}"
        np.testing.assert_allclose(ctx.scalar_forcing(), fo, rtol=1e-12, atol=1e-15)


def test_scalar_boundp(tmp_path):
    ctx, oc = scalar_pair(tmp_path)
    seed(ctx, oc)
    for f in ("s1", "s2"):
        ctx.boundp(f)
        oc.boundp(f)
        np.testing.assert_array_equal(ctx.get(0, f), oc.get(0, f))


@pytest.mark.parametrize("variant", [0, 2, 8])
def test_rk_mom_buoyancy(tmp_path, variant):
    ctx, oc = scalar_pair(tmp_path)
    ctx.set_rk_variant(variant)
    for irk in (1, 2, 3):
        seed(ctx, oc, seed=10 + irk, names=("u", "v", "w", "p", "s1", "dudtrko", "dvdtrko", "dwdtrko"))
        ctx.rk_mom(irk, 1.5e-3)
        oc.rk_mom(irk, 1.5e-3)
        assert_fields(ctx, oc, ("u", "v", "w", "dudtrko", "dvdtrko", "dwdtrko"), 1e-13, interior=True)


def test_buoyancy_off_by_default(tmp_path):
    """Without the switch (a default reference build) gacc and beta leave
    rk_mom alone, as rk.f90:104-108 compiles the term only under
    _BOUSSINESQ_BUOYANCY; on, the velocities change."""
    off, oc = scalar_pair(tmp_path, buoyancy=False)
    on, _ = scalar_pair(tmp_path / "on")
    seed(off, oc, seed=5, names=("u", "v", "w", "p", "s1"))
    seed(on, oc, seed=5, names=("u", "v", "w", "p", "s1"))
    off.rk_mom(1, 1.5e-3)
    on.rk_mom(1, 1.5e-3)
    oc.rk_mom(1, 1.5e-3)
    assert_fields(off, oc, ("u", "v", "w"), 1e-13, interior=True)
    assert np.abs(on.get(0, "w") - off.get(0, "w")).max() > 1e-6


@pytest.mark.parametrize("forced", [False, True])
def test_scalars_10_steps(tmp_path, forced):
    ctx, oc = scalar_pair(tmp_path, forced=forced)
    seed(ctx, oc, names=("u", "v", "w", "s1", "s2"))
    run_both(ctx, oc, 10)
    assert_fields(ctx, oc, ("u", "v", "w", "s1", "s2"), 1e-10, interior=True)
    assert_fields(ctx, oc, ("p",), 1e-10, interior=True, modmean=("p",))


def test_dhc_reference_case():
    """The reference's DHC test through the HIP path: 20 steps of dt_f =
    1e-3 with FFT along y, Boussinesq on; GPU fields against the oracle after
    10 steps (1e-10), then the checks of test.py after 20."""
    import snac_amd
    dn, bn = os.path.join(GOLD, "dhc", "dns.nml"), os.path.join(GOLD, "dhc", "blocks.nml")
    ctx = snac_amd.Context(dn, bn, 2)
    oc = oracle_py.OracleCase(nmlparse.load_case(dn, bn, 2))
    oc.set_solver("band")
    ctx.set_buoyancy(True)  # testit.sh builds with BOUSSINESQ_BUOYANCY=1
    oc.set_buoyancy(True)
    xc = ctx.metric(0, 6)
    s = np.zeros(ctx.shape(0, "s1"))
    s[:, :, :] = ((1.0 - xc) * -0.5 + xc * 0.5)[:, None, None]  # iniscal 'dhc', initflow.f90:247-256
    ctx.set(0, "s1", s)
    oc.set(0, "s1", s)
    run_both(ctx, oc, 10)
    assert_fields(ctx, oc, ("u", "v", "w", "s1"), 1e-10, interior=True)
    ctx.steps(1e-3, 10)
    sc = ctx.get(0, "s1")[1:-1, 1:-1, 1:-1]
    u, w = ctx.get(0, "u")[1:-1, 1:-1, 1:-1], ctx.get(0, "w")[1:-1, 1:-1, 1:-1]
    assert np.all(np.isfinite(sc)) and np.all(np.isfinite(u)) and np.all(np.isfinite(w))
    assert sc[0].mean() < sc[-1].mean()
    _, divmax = ctx.chkdiv()
    assert divmax < 1e-10


def test_scalar_graph_replays_match_eager(tmp_path):
    """A step flips the scalar buffers three times: the graph path keeps one
    capture per starting parity; every replay matches eager launches."""
    eager, oc = scalar_pair(tmp_path, fa=0)
    graph, _ = scalar_pair(tmp_path / "g", fa=0)
    eager.set_graph(False)
    rng = np.random.default_rng(4)
    for f in ("u", "v", "w", "s1", "s2"):
        a = rng.uniform(-1, 1, eager.shape(0, f))
        eager.set(0, f, a)
        graph.set(0, f, a)
    for c in (eager, graph):
        c.init_bounds()
    dt = 1e-3
    for rep in range(4):
        eager.step(dt)
        graph.step(dt)
        for f in ("u", "v", "w", "p", "s1", "s2"):
            np.testing.assert_array_equal(graph.get(0, f), eager.get(0, f), err_msg=f"replay {rep} field {f}")
    graph.steps(dt, 3)
    eager.steps(dt, 3)
    for f in ("u", "s1", "s2"):
        np.testing.assert_array_equal(graph.get(0, f), eager.get(0, f))

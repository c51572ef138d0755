"""The oracle's scalar transport (rk_scal rk.f90:130-204, scal.f90:89-199,
the Boussinesq term rk.f90:104-108, scalar boundp) pinned by closed forms and
by the reference's differentially-heated-cavity test
(tests/differentially_heated_cavity: inputs copied to tests/golden/dhc, the
checks of its test.py).
"""
import os

import numpy as np
import pytest

import cases
import nmlparse
import oracle_py

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def scalar_box(tmp_path, n=(8, 6, 10), nscal=1, alphai=100.0, gacc=(0.0, 0.0, 0.0), beta=0.0, ssource=0.0,
               forced=False, scalf=0.0, periodic=True):
    blk = cases.triperiodic_block(n, (1.0, 1.0, 1.0)) if periodic else cases.cavity_block(n)
    d, b = cases.write_case(str(tmp_path), [blk], 100.0, nscal=nscal, gacc=gacc,
                            scalar=dict(alphai=[alphai] * nscal, beta=beta, ssource=[ssource] * nscal,
                                        is_sforced=[forced] * nscal, scalf=[scalf] * nscal))
    oc = oracle_py.OracleCase(nmlparse.load_case(d, b, 0 if not periodic else 3))
    return oc


def test_linear_scalar_without_flow_is_steady(tmp_path):
    """scal_d of a linear field is zero on a uniform grid and scal_a vanishes
    without flow: rk_scal only adds f12*ssource per substage, dt per step."""
    oc = scalar_box(tmp_path, periodic=False, ssource=0.7)
    x = oc.metric(0, 6)
    s = np.broadcast_to((0.3 + 2.0 * x)[:, None, None], oc.shape(0, "s1")).copy()
    dt, acc = 1e-3, 0.0
    for irk in (1, 2, 3):
        oc.set(0, "s1", s + acc)  # ghosts follow the uniform shift (still linear)
        oc.rk_scal(irk, dt)
        acc += sum(oracle_py.RKCOEFF[irk - 1]) * dt * 0.7
        out = oc.get(0, "s1")[1:-1, 1:-1, 1:-1]
        np.testing.assert_allclose(out, s[1:-1, 1:-1, 1:-1] + acc, rtol=0, atol=1e-13)
    assert acc == pytest.approx(0.7 * dt, rel=1e-14)


def test_uniform_advection_conserves_scalar(tmp_path):
    """Periodic box, uniform velocity: the advective fluxes telescope, so the
    volume integral of s is unchanged (diffusion also conserves it)."""
    oc = scalar_box(tmp_path, n=(8, 6, 10))
    rng = np.random.default_rng(5)
    for f, val in (("u", 0.8), ("v", -0.3), ("w", 0.5)):
        oc.set(0, f, np.full(oc.shape(0, f), val))
    s = rng.uniform(-1, 1, oc.shape(0, "s1"))
    oc.set(0, "s1", s)
    oc.boundp("s1")
    s0 = oc.get(0, "s1")[1:-1, 1:-1, 1:-1].sum()
    oc.rk_scal(1, 1e-3)
    s1 = oc.get(0, "s1")[1:-1, 1:-1, 1:-1]
    assert abs(s1.sum() - s0) < 1e-12 * np.abs(s1).sum()
    assert np.abs(s1 - s[1:-1, 1:-1, 1:-1]).max() > 1e-6  # it did change


def test_bulk_forcing_sets_the_mean(tmp_path):
    oc = scalar_box(tmp_path, forced=True, scalf=2.5)
    rng = np.random.default_rng(6)
    oc.set(0, "s1", rng.uniform(-1, 1, oc.shape(0, "s1")))
    oc.boundp("s1")
    f = oc.rk_scal(1, 1e-3)
    s = oc.get(0, "s1")[1:-1, 1:-1, 1:-1]
    assert s.mean() == pytest.approx(2.5, abs=1e-13)  # uniform grid: volume mean = plain mean
    assert f[0] != 0.0


def test_boussinesq_term(tmp_path):
    """Periodic box at rest, constant s = c, p = 0: each substage adds
    -f12*gacc*beta*c (the uniform velocity it creates has no advection or
    diffusion), so three substages give -dt*gacc*beta*c (no projection)."""
    oc = scalar_box(tmp_path, gacc=(0.0, 0.0, -2.0), beta=0.5)
    oc.set_buoyancy(True)
    oc.set(0, "s1", np.full(oc.shape(0, "s1"), 3.0))
    dt = 1e-2
    for irk in (1, 2, 3):
        oc.rk_mom(irk, dt)
        oc.bounduvw(False)
    w = oc.get(0, "w")[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(w, -dt * (-2.0) * 0.5 * 3.0, rtol=1e-14)
    assert np.abs(oc.get(0, "u")[1:-1, 1:-1, 1:-1]).max() == 0.0


def test_scalar_boundp_rules(tmp_path):
    """Centered D: ghost = 2b - f ; centered N: ghost = f -+ dr*b."""
    blk = cases.cavity_block((6, 5, 4))
    blk["cbcscal"] = [[["D", "N"], ["N", "N"], ["N", "D"]]]
    blk["bcscal"] = [[[1.5, 0.25], [0.0, 0.0], [0.0, -2.0]]]
    d, b = cases.write_case(str(tmp_path), [blk], 100.0, nscal=1, scalar=dict(alphai=[50.0]))
    oc = oracle_py.OracleCase(nmlparse.load_case(d, b, 0))
    s = np.random.default_rng(1).uniform(-1, 1, oc.shape(0, "s1"))
    oc.set(0, "s1", s)
    oc.boundp("s1")
    g = oc.get(0, "s1")
    dxc = oc.metric(0, 0)
    np.testing.assert_allclose(g[0, 1:-1, 1:-1], 2 * 1.5 - s[1, 1:-1, 1:-1])
    np.testing.assert_allclose(g[-1, 1:-1, 1:-1], s[-2, 1:-1, 1:-1] + dxc[-2] * 0.25)
    np.testing.assert_allclose(g[1:-1, 1:-1, -1], 2 * -2.0 - s[1:-1, 1:-1, -2])


@pytest.mark.slow
def test_dhc_reference_checks():
    """The reference's differentially heated cavity (FFT along y, Boussinesq,
    20 steps of dt_f = 1e-3, initial scalar linear in x between the wall
    values, initflow.f90:247-256): fields finite and the mean scalar next to
    the cold wall below that next to the hot wall (test.py:11-25)."""
    dn, bn = os.path.join(GOLD, "dhc", "dns.nml"), os.path.join(GOLD, "dhc", "blocks.nml")
    c = nmlparse.load_case(dn, bn, 2)
    oc = oracle_py.OracleCase(c)
    oc.set_solver("band")
    oc.set_buoyancy(True)  # testit.sh builds with BOUSSINESQ_BUOYANCY=1
    xc = oc.metric(0, 6)
    s = np.zeros(oc.shape(0, "s1"))
    s[:, :, :] = ((1.0 - xc) * -0.5 + xc * 0.5)[:, None, None]
    oc.set(0, "s1", s)
    oc.init_bounds()
    for _ in range(20):
        oc.step(1e-3)
    sc = oc.get(0, "s1")[1:-1, 1:-1, 1:-1]
    u, w = oc.get(0, "u")[1:-1, 1:-1, 1:-1], oc.get(0, "w")[1:-1, 1:-1, 1:-1]
    assert np.all(np.isfinite(sc)) and np.all(np.isfinite(u)) and np.all(np.isfinite(w))
    assert sc[0].mean() < sc[-1].mean()
    assert np.abs(w).max() > 0.0  # buoyancy drives a flow
    _, divmax = oc.chkdiv()
    assert divmax < 1e-10


def test_buoyancy_switch(tmp_path):
    """The _BOUSSINESQ_BUOYANCY switch: off by default (the Boussinesq term
    of test_boussinesq_term is absent), and refused without scalars, as
    param.f90:166-170 stops such a build."""
    import snac_amd
    oc = scalar_box(tmp_path, gacc=(0.0, 0.0, -2.0), beta=0.5)
    oc.set(0, "s1", np.full(oc.shape(0, "s1"), 3.0))
    oc.rk_mom(1, 1e-2)
    assert np.abs(oc.get(0, "w")).max() == 0.0
    d, b = cases.write_case(str(tmp_path / "ns"), [cases.cavity_block((4, 4, 4))], 100.0)
    oc0 = oracle_py.OracleCase(nmlparse.load_case(d, b, 0))
    with pytest.raises(RuntimeError, match="requires nscal > 0"):
        oc0.set_buoyancy(True)
    ctx = snac_amd.Context(d, b, 0, device=-1)
    with pytest.raises(snac_amd.SnacError, match="requires nscal > 0"):
        ctx.set_buoyancy(True)
    ctx.close()

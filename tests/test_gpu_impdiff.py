"""Implicit diffusion (the reference's _IMPDIFF build) on the HIP path,
through the C-ABI, against the CPU oracle and a closed form.

* Helmholtz solves of u, v, w (solve_impdiff_field) from seeded fields on
  every geometry of impdiff_cases: staggered periodic and Dirichlet axes
  (transformed, along the FFT axis with n-1 points, and along the line axis
  with the boundary face excluded), inhomogeneous Dirichlet values (the
  cavity lid) and an inflow plane (duct).  The oracle solves the assembled
  systems directly (banded Cholesky); gate: relative L2 1e-12.
* updt_pressure with the Laplacian term, chkdt without the diffusive limit.
* whole steps (hipGraph replay) against the oracle's steps, relative L2
  1e-10 (p modulo its mean).
* a shear wave on a periodic box: each substage multiplies it by
  (1 - a)/(1 + a), a = visc dtrk kappa^2 / 2.
"""
import numpy as np
import pytest

import cases
import impdiff_cases
import nmlparse
import oracle_py

pytestmark = pytest.mark.gpu


def make_pair(tmp_path, geom, fft_axis, visci=100.0):
    import snac_amd
    blocks, dns = impdiff_cases.GEOMS[geom]()
    d, b = cases.write_case(str(tmp_path), blocks, visci, **dns)
    print(f"impdiff case {geom}: fft_axis={fft_axis}", flush=True)
    ctx = snac_amd.Context(d, b, fft_axis)
    ctx.set_impdiff(True)
    c = nmlparse.load_case(d, b, fft_axis)
    c["impdiff"] = True
    oc = oracle_py.OracleCase(c)
    oc.set_solver("band")
    return ctx, oc


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def unknowns(ctx, comp):
    """interior slice of component comp's unknowns (block 0): the bounded
    upper face along its own axis is excluded"""
    info = ctx.block_info(0)
    sl = [slice(1, -1)] * 3
    if info["nb"][comp][1] < 0:
        sl[comp] = slice(1, -2)
    return tuple(sl)


CASES = [("channel", 3), ("channel", 1), ("cavity", 0), ("cavity", 3), ("duct", 3), ("box", 1)]


@pytest.mark.parametrize("geom,fft_axis", CASES)
def test_helmholtz_solves(tmp_path, geom, fft_axis):
    ctx, oc = make_pair(tmp_path, geom, fft_axis)
    rng = np.random.default_rng(11)
    for comp, name in enumerate("uvw"):
        for alpha in (-2.0e-3, -0.35):
            a = rng.uniform(-1, 1, ctx.shape(0, name))
            ctx.set(0, name, a)
            oc.set(0, name, a)
            ctx.solve_impdiff(name, alpha)
            oc.solve_impdiff(name, alpha)
            sl = unknowns(ctx, comp)
            r = rel(ctx.get(0, name)[sl], oc.get(0, name)[sl])
            assert r <= 1e-12, f"{name} alpha {alpha}: rel L2 {r:.3e}"
    ctx.close()


def test_updt_pressure_and_chkdt(tmp_path):
    ctx, oc = make_pair(tmp_path, "channel", 3)
    rng = np.random.default_rng(5)
    for f in ("p", "pp", "u", "v", "w"):
        a = rng.uniform(-1, 1, ctx.shape(0, f))
        ctx.set(0, f, a)
        oc.set(0, f, a)
    ctx.updt_pressure_impdiff(-3.1e-3)
    oc.updt_pressure_impdiff(-3.1e-3)
    assert rel(ctx.get(0, "p")[1:-1, 1:-1, 1:-1], oc.get(0, "p")[1:-1, 1:-1, 1:-1]) <= 1e-14
    assert ctx.chkdt() == pytest.approx(oc.chkdt(), rel=1e-13)
    ctx.close()


@pytest.mark.parametrize("geom,fft_axis", [("channel", 3), ("cavity", 0), ("cavity", 3), ("duct", 3)])
def test_steps(tmp_path, geom, fft_axis):
    ctx, oc = make_pair(tmp_path, geom, fft_axis)
    rng = np.random.default_rng(3)
    for f in ("u", "v", "w"):
        a = np.zeros(ctx.shape(0, f))
        a[1:-1, 1:-1, 1:-1] = 0.2 * rng.uniform(-1, 1, tuple(s - 2 for s in a.shape))
        ctx.set(0, f, a)
        oc.set(0, f, a)
    ctx.init_bounds()
    oc.init_bounds()
    dt = 2.0e-3
    ctx.steps(dt, 4)
    for _ in range(4):
        oc.step(dt)
    for f in ("u", "v", "w", "p"):
        g, o = ctx.get(0, f)[1:-1, 1:-1, 1:-1], oc.get(0, f)[1:-1, 1:-1, 1:-1]
        if f == "p":
            g, o = g - g.mean(), o - o.mean()
        r = rel(g, o)
        assert r <= 1e-10, f"{f}: rel L2 {r:.3e}"
    d1, d2 = ctx.chkdiv()
    assert d2 < 1e-9
    ctx.close()


@pytest.mark.parametrize("geom", ["channel_scalar", "channel_scalar_free"])
def test_scalar_solve_and_steps(tmp_path, geom):
    """the scalar's Helmholtz solve (Dirichlet values 0 and 1 on the walls:
    BC planes) and whole steps with the scalar forced to its bulk value (the
    mean taken before the half diffusion comes off) or free, with buoyancy"""
    ctx, oc = make_pair(tmp_path, geom, 3)
    rng = np.random.default_rng(21)
    a = rng.uniform(-1, 1, ctx.shape(0, "s1"))
    ctx.set(0, "s1", a)
    oc.set(0, "s1", a)
    ctx.solve_impdiff("s1", -0.01)
    oc.solve_impdiff("s1", -0.01)
    in3 = (slice(1, -1),) * 3
    assert rel(ctx.get(0, "s1")[in3], oc.get(0, "s1")[in3]) <= 1e-12
    for f in ("u", "v", "w", "s1"):
        a = np.zeros(ctx.shape(0, f))
        a[in3] = 0.2 * rng.uniform(-1, 1, tuple(q - 2 for q in a.shape))
        ctx.set(0, f, a)
        oc.set(0, f, a)
    ctx.init_bounds()
    oc.init_bounds()
    dt = 2.0e-3
    ctx.steps(dt, 3)
    for _ in range(3):
        oc.step(dt)
    for f in ("u", "v", "w", "p", "s1"):
        g, o = ctx.get(0, f)[in3], oc.get(0, f)[in3]
        if f == "p":
            g, o = g - g.mean(), o - o.mean()
        r = rel(g, o)
        assert r <= 1e-10, f"{f}: rel L2 {r:.3e}"
    ctx.close()


@pytest.mark.parametrize("fft_axis", [2, 0])
def test_dhc_reference_case_steps(tmp_path, fft_axis):
    """the reference's differentially heated cavity (tests/
    differentially_heated_cavity inputs) with implicit diffusion: walls in
    x (scalar -0.5 / 0.5) and z, periodic y, Boussinesq buoyancy"""
    import os
    import snac_amd
    g = os.path.join(os.path.dirname(__file__), "golden", "dhc")
    d, b = os.path.join(g, "dns.nml"), os.path.join(g, "blocks.nml")
    ctx = snac_amd.Context(d, b, fft_axis)
    ctx.set_impdiff(True)
    c = nmlparse.load_case(d, b, fft_axis)
    c["impdiff"] = True
    oc = oracle_py.OracleCase(c)
    oc.set_solver("band")
    ctx.set_buoyancy(True)  # the reference's test builds with BOUSSINESQ_BUOYANCY=1
    oc.set_buoyancy(True)
    rng = np.random.default_rng(4)
    in3 = (slice(1, -1),) * 3
    for f in ("u", "v", "w", "s1"):
        a = np.zeros(ctx.shape(0, f))
        a[in3] = 0.1 * rng.uniform(-1, 1, tuple(q - 2 for q in a.shape))
        ctx.set(0, f, a)
        oc.set(0, f, a)
    ctx.init_bounds()
    oc.init_bounds()
    dt = 1.0e-3
    ctx.steps(dt, 3)
    for _ in range(3):
        oc.step(dt)
    for f in ("u", "v", "w", "p", "s1"):
        gg, o = ctx.get(0, f)[in3], oc.get(0, f)[in3]
        if f == "p":
            gg, o = gg - gg.mean(), o - o.mean()
        r = rel(gg, o)
        assert r <= 1e-10, f"{f}: rel L2 {r:.3e}"
    ctx.close()


def test_shear_wave_closed_form(tmp_path):
    """u = A sin(y) on a periodic box (every axis transformed): advection and
    pressure vanish, each substage applies (1 - a)/(1 + a)"""
    import snac_amd
    n = (8, 32, 8)
    L = (2 * np.pi,) * 3
    visci = 2.0
    d, b = cases.write_case(str(tmp_path), [cases.triperiodic_block(n, lengths=L)], visci)
    ctx = snac_amd.Context(d, b, 1)
    ctx.set_impdiff(True)
    yc = ctx.metric(0, 8)
    u = np.zeros(ctx.shape(0, "u"))
    u[:] = 0.8 * np.sin(yc)[None, :, None]
    ctx.set(0, "u", u)
    ctx.init_bounds()
    dt, nsteps = 0.04, 3
    ctx.steps(dt, nsteps)
    h = L[1] / n[1]
    k2 = 4.0 / h ** 2 * np.sin(h / 2) ** 2
    g = 1.0
    for a1, a2 in oracle_py.RKCOEFF:
        a = (a1 + a2) * dt * k2 / visci / 2
        g *= (1 - a) / (1 + a)
    in3 = (slice(1, -1),) * 3
    np.testing.assert_allclose(ctx.get(0, "u")[in3], g ** nsteps * u[in3], rtol=0, atol=1e-13)
    for f in ("v", "w"):
        assert np.abs(ctx.get(0, f)[in3]).max() < 1e-13
    ctx.close()

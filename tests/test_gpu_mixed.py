"""Mixed precision against the FP64 oracle.

The mixed mode evaluates the momentum right-hand side of rk_mom
(mom.f90:112-327) in FP32, on neighbour offsets from each cell's FP64 values,
and keeps the old tendencies in FP32, while the RK3 accumulation
(rk.f90:68-85), the pressure
gradient and the body force (rk.f90:86-111) stay FP64; on the MG-PCG path
the V-cycle also runs in FP32 under a flexible CG with FP64 residuals.  The
oracle is the FP64 restatement of the same routines, so these gates bound
the error the FP32 arithmetic adds:

* one substage from seeded fields: tendencies within 1e-6 (FP32 rounding of
  the right-hand side, about 1e-7 measured), velocities within 1e-8 (the
  right-hand side enters scaled by dt);
* ten whole steps of a stretched channel: every field within 1e-6 and the
  divergence below 1e-10, the gates the mixed-precision configuration sets;
* 1000 steps of a small perturbed channel: the per-field errors of the
  mixed and of the FP64 run against the oracle, stated side by side (the
  flow amplifies differences by about 1e9 over the run, so the mixed run's
  fluctuations decorrelate; its mean profile and fluctuation energies are
  gated instead).
"""
import numpy as np
import pytest

import cases
from test_gpu_parity import FIELDS3, assert_fields, make_pair, rel, seed_fields

pytestmark = pytest.mark.gpu

DTO = ("dudtrko", "dvdtrko", "dwdtrko")


def channel(tmp_path, n, **kw):
    return make_pair(tmp_path, [cases.channel_block(n)], 3, bforce=(12 / 5600, 0, 0), **kw)


@pytest.mark.parametrize("n", [(70, 13, 37), (64, 24, 160)])
@pytest.mark.parametrize("variant", range(7, 17))
def test_rk_mom_f32(tmp_path, n, variant):
    """Every FP32 kernel variant, on a ragged grid (partial tiles along i, j
    and k) and on a 160-plane grid (the 64-plane walk hands over twice)."""
    ctx, oc = channel(tmp_path, n)
    ctx.set_precision("mixed")
    ctx.set_rk32_variant(variant)
    for irk in (1, 2, 3):
        seed_fields(ctx, oc, seed=20 + irk)
        ctx.rk_mom(irk, 1.3e-3)
        oc.rk_mom(irk, 1.3e-3)
        errs = {f: rel(ctx.get(0, f), oc.get(0, f)) for f in FIELDS3 + DTO}
        print(f"substage {irk}: " + " ".join(f"{f} {e:.2e}" for f, e in errs.items()), flush=True)
        for f in FIELDS3:
            assert errs[f] <= 1e-8, (f, errs[f])
        for f in DTO:
            assert 0.0 < errs[f] <= 1e-6, (f, errs[f])  # FP32 rounding is visible, and small


def test_rk_mom_f32_uses_fp32_tendencies(tmp_path):
    """The old tendencies are held in FP32: an uploaded FP64 value comes back
    rounded to FP32, and switching back to FP64 keeps the FP32 values."""
    ctx, _ = channel(tmp_path, (16, 8, 8))
    a = np.random.default_rng(3).uniform(-1, 1, ctx.shape(0, "dvdtrko"))
    ctx.set_precision("mixed")
    ctx.set(0, "dvdtrko", a)
    back = ctx.get(0, "dvdtrko")
    assert np.array_equal(back, a.astype(np.float32).astype(np.float64))
    ctx.set_precision("fp64")
    assert np.array_equal(ctx.get(0, "dvdtrko"), back)
    assert not ctx.get(0, "dudtrko").any()


def run_mixed(ctx, oc, nsteps):
    """nsteps whole steps; the oracle takes the GPU's dt (the CFL check reads
    velocities that differ in the last digits)."""
    ctx.init_bounds()
    oc.init_bounds()
    for _ in range(nsteps):
        dt = ctx.new_dt()
        dto = min(ctx.cfl * oc.chkdt(), ctx.dtmax)
        assert dt == pytest.approx(dto, rel=1e-6)
        ctx.step(dt)
        oc.step(dt)


def field_errors(ctx, oc):
    out = {}
    for f in FIELDS3 + ("p",):
        g, o = ctx.get(0, f)[1:-1, 1:-1, 1:-1], oc.get(0, f)[1:-1, 1:-1, 1:-1]
        if f == "p":
            g, o = g - g.mean(), o - o.mean()
        out[f] = rel(g, o)
    return out


@pytest.mark.parametrize("solver", ["tensor", "mg"])
def test_mixed_channel_10_steps(tmp_path, solver):
    """Ten steps of a stretched periodic channel in mixed precision against
    the FP64 oracle: per-field relative L2 <= 1e-6, max|div| <= 1e-10.  On
    the MG-PCG path the V-cycle runs in FP32 too."""
    ctx, oc = channel(tmp_path, (64, 48, 32))
    if solver == "mg":
        ctx.set_solver("mg")
    ctx.set_precision("mixed")
    u, v, w = cases.channel_ic(ctx.shape(0, "u"), ctx.metric(0, 8))
    for f, a in zip(FIELDS3, (u, v, w)):
        ctx.set(0, f, a)
        oc.set(0, f, a)
    run_mixed(ctx, oc, 10)
    errs = field_errors(ctx, oc)
    _, divmax = ctx.chkdiv()
    print(f"{solver}: " + " ".join(f"{f} {e:.2e}" for f, e in errs.items()) + f" divmax {divmax:.2e}", flush=True)
    for f, e in errs.items():
        assert e <= 1e-6, (f, e)
    assert divmax <= 1e-10
    if solver == "mg":
        it, relres = ctx.solver_stats()
        assert relres <= 1e-13


def test_precision_switch_mid_run(tmp_path):
    """FP64 steps, then mixed, then FP64 again on one context: the old
    tendencies follow each switch (FP64 -> FP32 -> FP64) and the run stays
    within the mixed gate of the FP64 oracle."""
    ctx, oc = channel(tmp_path, (32, 24, 16))
    u, v, w = cases.channel_ic(ctx.shape(0, "u"), ctx.metric(0, 8))
    for f, a in zip(FIELDS3, (u, v, w)):
        ctx.set(0, f, a)
        oc.set(0, f, a)
    ctx.init_bounds()
    oc.init_bounds()
    for mode in ("fp64", "mixed", "fp64"):
        ctx.set_precision(mode)
        for _ in range(3):
            dt = ctx.new_dt()
            ctx.step(dt)
            oc.step(dt)
    errs = field_errors(ctx, oc)
    for f, e in errs.items():
        assert e <= 1e-6, (f, e)
    assert_fields(ctx, oc, DTO, 1e-6)


def plane_stats(get):
    """the mean streamwise profile <u>(y) over x and z, and the mean kinetic
    energies of u - <u>, v and w over the interior"""
    inner = (slice(1, -1),) * 3
    u, v, w = (get(f)[inner] for f in FIELDS3)
    um = u.mean(axis=(0, 2))
    return um, np.mean((u - um[None, :, None]) ** 2), np.mean(v ** 2), np.mean(w ** 2)


def test_mixed_channel_1000_steps(tmp_path):
    """1000 steps of the stretched periodic channel (32x24x16, fixed dt,
    pressure on the MG-PCG path) in mixed precision and in FP64, each
    against the FP64 oracle.  The per-field errors after 1000 steps are
    stated side by side: this perturbed channel amplifies differences by
    about 1e9 over the 1000 steps (the FP64 run, whose differences with the
    FP64 oracle start at rounding level, ends near 1e-7), so FP32 rounding
    decorrelates the fluctuations by then.  Gates: the FP64 run's fields
    within 1e-5; the mixed run's statistics -- the mean streamwise profile
    and the fluctuation energies of u, v and w -- against the oracle's;
    max|div| <= 1e-10 for both."""
    res = {}
    for prec in ("fp64", "mixed"):
        ctx, oc = channel(tmp_path / prec, (32, 24, 16))
        ctx.set_solver("mg")
        ctx.set_precision(prec)
        u, v, w = cases.channel_ic(ctx.shape(0, "u"), ctx.metric(0, 8))
        for f, a in zip(FIELDS3, (u, v, w)):
            ctx.set(0, f, a)
            oc.set(0, f, a)
        ctx.init_bounds()
        oc.init_bounds()
        dt = 0.5 * ctx.chkdt()
        for _ in range(10):
            ctx.steps(dt, 100)
            for _ in range(100):
                oc.step(dt)
        errs = field_errors(ctx, oc)
        _, divmax = ctx.chkdiv()
        g = plane_stats(lambda f: ctx.get(0, f))
        o = plane_stats(lambda f: oc.get(0, f))
        prof = rel(g[0], o[0])
        en = [abs(a / b - 1.0) for a, b in zip(g[1:], o[1:])]
        print(f"{prec} after 1000 steps: " + " ".join(f"{f} {e:.2e}" for f, e in errs.items()) +
              f" divmax {divmax:.2e}; mean profile {prof:.2e}, fluctuation energies (u, v, w) relative to the "
              f"oracle's {en[0]:.2e} {en[1]:.2e} {en[2]:.2e}", flush=True)
        res[prec] = (errs, prof, en)
        assert divmax <= 1e-10
        ctx.close()
    for f in FIELDS3 + ("p",):
        assert res["fp64"][0][f] <= 1e-5, (f, res["fp64"][0][f])
    # measured: mean profile 1.3e-2, fluctuation energies 6-12 %
    assert res["mixed"][1] <= 5e-2, res["mixed"][1]
    assert max(res["mixed"][2]) <= 0.3, res["mixed"][2]

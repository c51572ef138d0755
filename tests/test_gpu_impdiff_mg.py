"""Implicit diffusion with the pressure on the batched MG-PCG path: each
field's Helmholtz problem (solve_impdiff_field, solve_helmholtz.f90:30-100)
is solved by its own MG plan -- the component's staggered unknowns and
boundary conditions, the per-mode eigenvalues shifted by 1/alpha -- against
the CPU oracle's direct banded solves of the assembled systems.

* single Helmholtz solves of u, v, w (and a scalar) on a channel forced onto
  the MG-PCG path and on the L-shaped channel (walls along the FFT axis:
  sine transforms for u and v, w over its n-1 interior faces), from a shift
  that makes the finest level's line sweeps the whole preconditioner to one
  that needs the V-cycle: every solve stops at hypre_tol = 1e-13, gate 1e-10;
* whole steps on the backward-facing step and the L-shaped channel (unions
  of blocks that are not boxes, so no tensor-product plan exists) and on the
  forced channel, gate 1e-10 on every field (p modulo its mean where the
  pressure is singular); after the first step the Helmholtz solves run as
  graph segments with the host resuming any solve left above hypre_tol;
* the reference's differentially heated cavity with the scalar and
  Boussinesq buoyancy, along the periodic and along the walled FFT axis.
"""
import numpy as np
import pytest

import cases
import impdiff_cases
import nmlparse
import oracle_py

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make_pair(tmp_path, blocks, visci, fft_axis=3, **dns):
    import snac_amd
    d, b = cases.write_case(str(tmp_path), blocks, visci, **dns)
    ctx = snac_amd.Context(d, b, fft_axis)
    ctx.set_solver("mg")
    ctx.set_impdiff(True)
    c = nmlparse.load_case(d, b, fft_axis)
    c["impdiff"] = True
    oc = oracle_py.OracleCase(c)
    oc.set_solver("band")
    return ctx, oc


def unknowns(ctx, q, comp):
    """interior slice of component comp's unknowns in block q: the bounded
    upper face along its own axis is excluded"""
    info = ctx.block_info(q)
    sl = [slice(1, -1)] * 3
    if comp < 3 and info["nb"][comp][1] < 0:
        sl[comp] = slice(1, -2)
    return tuple(sl)


def geometry(geom):
    """(blocks, visci, dns, singular pressure)"""
    if geom == "bfs":
        return cases.bfs_blocks((12, 8), (20, 8, 12), 8), 1000.0, {}, False
    if geom == "lchannel":
        return cases.lchannel_blocks(8, 16), 100.0, {}, False
    blocks, dns = impdiff_cases.GEOMS["channel"]()
    return blocks, 100.0, dns, True


@pytest.mark.parametrize("geom,prec", [("channel", "fp64"), ("channel", "mixed"), ("lchannel", "fp64"),
                                       ("lchannel", "mixed")])
def test_helmholtz_solves_mg(tmp_path, geom, prec):
    blocks, visci, dns, _ = geometry(geom)
    ctx, oc = make_pair(tmp_path, blocks, visci, **dns)
    ctx.set_precision(prec)
    rng = np.random.default_rng(11)
    for comp, name in enumerate("uvw"):
        # -1e-6: the shift outweighs the couplings across lines by far, the
        # finest level's zebra sweeps alone precondition the CG
        for alpha in (-1.0e-6, -2.0e-3, -0.35):
            for q in range(ctx.nblocks):
                a = rng.uniform(-1, 1, ctx.shape(q, name))
                ctx.set(q, name, a)
                oc.set(q, name, a)
            ctx.solve_impdiff(name, alpha)
            oc.solve_impdiff(name, alpha)
            iters, relres = ctx.impdiff_stats(name)
            assert 0 < iters and relres <= 1e-13 * 1.0001, (iters, relres)
            for q in range(ctx.nblocks):
                sl = unknowns(ctx, q, comp)
                r = rel(ctx.get(q, name)[sl], oc.get(q, name)[sl])
                print(f"{geom} {prec} {name} alpha {alpha} block {q}: {iters} CG iterations, relres {relres:.2e}, "
                      f"rel L2 {r:.2e}", flush=True)
                assert r <= 1e-10, f"{name} alpha {alpha} block {q}: rel L2 {r:.3e}"
    ctx.close()


def test_scalar_helmholtz_mg(tmp_path):
    blocks, dns = impdiff_cases.GEOMS["channel_scalar"]()
    ctx, oc = make_pair(tmp_path, blocks, 100.0, **dns)
    rng = np.random.default_rng(21)
    a = rng.uniform(-1, 1, ctx.shape(0, "s1"))
    ctx.set(0, "s1", a)
    oc.set(0, "s1", a)
    ctx.solve_impdiff("s1", -0.01)
    oc.solve_impdiff("s1", -0.01)
    in3 = (slice(1, -1),) * 3
    r = rel(ctx.get(0, "s1")[in3], oc.get(0, "s1")[in3])
    assert r <= 1e-10, r
    ctx.close()


def seed(ctx, oc, rng, amp=0.1):
    for q in range(ctx.nblocks):
        yc = ctx.metric(q, 8)
        for f in ("u", "v", "w"):
            a = np.zeros(ctx.shape(q, f))
            if f == "u" and yc[1:-1].max() <= 1.0:
                a[1:-1, 1:-1, 1:-1] = (6.0 * yc[1:-1] * (1.0 - yc[1:-1]))[None, :, None]
            a[1:-1, 1:-1, 1:-1] += amp * rng.uniform(-1, 1, tuple(s - 2 for s in a.shape))
            ctx.set(q, f, a)
            oc.set(q, f, a)


@pytest.mark.parametrize("geom,prec", [("bfs", "fp64"), ("lchannel", "fp64"), ("channel", "fp64"),
                                       ("bfs", "mixed")])
def test_steps_mg(tmp_path, geom, prec):
    """prec mixed: the FP32 momentum right-hand side and FP32 V-cycles of
    the pressure and Helmholtz solves (FP64 CG residuals): every field within
    the mixed form's 1e-6 of the FP64 oracle"""
    blocks, visci, dns, singular = geometry(geom)
    ctx, oc = make_pair(tmp_path, blocks, visci, **dns)
    ctx.set_precision(prec)
    tol = 1e-10 if prec == "fp64" else 1e-6
    print(f"impdiff steps on the MG-PCG path: {geom} {prec}", flush=True)
    seed(ctx, oc, np.random.default_rng(3))
    ctx.init_bounds()
    oc.init_bounds()
    dt = 2.0e-3
    ctx.steps(dt, 4)
    for _ in range(4):
        oc.step(dt)
    for q in range(ctx.nblocks):
        for f in ("u", "v", "w", "p"):
            g, o = ctx.get(q, f)[1:-1, 1:-1, 1:-1], oc.get(q, f)[1:-1, 1:-1, 1:-1]
            if f == "p" and singular:
                g, o = g - g.mean(), o - o.mean()
            r = rel(g, o)
            assert r <= tol, f"block {q} {f}: rel L2 {r:.3e}"
    _, divmax = ctx.chkdiv()
    assert divmax < 1e-9
    # the first step measures each solve's CG count eagerly, the others
    # replay graphs with that many iterations captured and the host
    # resuming a solve left above hypre_tol: every field's last Helmholtz
    # solve met it
    for name in "uvw":
        iters, relres = ctx.impdiff_stats(name)
        print(f"{geom} {name}: last Helmholtz solve {iters} CG iterations, relres {relres:.2e}", flush=True)
        assert iters > 0 and relres <= 1e-13 * 1.0001, (name, iters, relres)
    ctx.close()


@pytest.mark.parametrize("fft_axis", [2, 3])
def test_dhc_steps_mg(tmp_path, fft_axis):
    """the reference's differentially heated cavity inputs (walls in x with
    the scalar at -0.5 / 0.5, walls in z, periodic y, Boussinesq buoyancy)
    with implicit diffusion of u, v, w and the scalar on the MG-PCG path:
    along the periodic y (fft_axis 2), and along the walled z (fft_axis 3:
    sine transforms for u and v, w over its n-1 interior faces, a cosine
    transform for the scalar and the pressure).  3 steps against the
    oracle, gate 1e-10 on every field (p modulo its mean)."""
    import os
    import snac_amd
    g = os.path.join(os.path.dirname(__file__), "golden", "dhc")
    b = os.path.join(g, "blocks.nml")
    # the reference's inputs with a tight &hypre tolerance (the file's 1e-4 /
    # 50 iterations would stop the iterative solves far from the oracle's
    # direct ones)
    with open(os.path.join(g, "dns.nml")) as f:
        txt = f.read()
    assert "hypre_tol = 1.e-4, hypre_maxiter = 50" in txt
    d = str(tmp_path / "dns.nml")
    with open(d, "w") as f:
        f.write(txt.replace("hypre_tol = 1.e-4, hypre_maxiter = 50", "hypre_tol = 1.e-13, hypre_maxiter = 1000"))
    print(f"dhc impdiff on the MG-PCG path: fft_axis={fft_axis}", flush=True)
    ctx = snac_amd.Context(d, b, fft_axis)
    ctx.set_solver("mg")
    ctx.set_impdiff(True)
    c = nmlparse.load_case(d, b, fft_axis)
    c["impdiff"] = True
    oc = oracle_py.OracleCase(c)
    oc.set_solver("band")
    ctx.set_buoyancy(True)
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
    for name in ("u", "v", "w", "s1"):
        it, relres = ctx.impdiff_stats(name)
        assert it > 0 and relres <= 1e-13 * 1.0001, (name, it, relres)
    ctx.close()

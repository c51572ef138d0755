"""Pins of the oracle's building blocks against what the mathematics or an
independent library fixes, so that a dropped term, wrong sign/index or
transposed operand fails here.

* r2r kinds: FFTW's definitions (fft.f90:96-149 selects them) equal
  scipy.fft.dct/dst with norm=None for types II/III/IV, and R2HC is numpy's
  rfft in halfcomplex order; forward+backward = 2n (or n) times identity.
* eigenvalues (fft.f90:151-189): the spectrum of the 1D second-difference
  matrix with the same boundary conditions (numpy eigvalsh).
* Poisson operators (solver.f90 init_matrix_2d/3d): symmetric, constant null
  vector in singular cases; the band, CG and tensor solvers agree and leave
  a residual at rounding level.
* momentum RHS: a linear velocity field has zero diffusion; on a uniform
  periodic box the diffusion of a sine wave is the discrete Laplacian's
  eigenvalue times the wave, and the vv-advection of a sine by a uniform
  stream is its central difference.
* projection: fillps -> solve -> correc leaves a divergence-free field.
"""
import os

import numpy as np
import pytest
import scipy.fft as sf

import cases
import nmlparse
import oracle_py


KINDS = {"REDFT10": ("dct", 2), "REDFT01": ("dct", 3), "RODFT10": ("dst", 2), "RODFT01": ("dst", 3),
         "REDFT11": ("dct", 4), "RODFT11": ("dst", 4), "RODFT00": ("dst", 1)}


@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 8, 16, 64])
@pytest.mark.parametrize("kind", sorted(KINDS))
def test_r2r_vs_scipy(kind, n):
    x = np.random.default_rng(n).uniform(-1, 1, n)
    fam, typ = KINDS[kind]
    ref = (sf.dct if fam == "dct" else sf.dst)(x, type=typ, norm=None)
    np.testing.assert_allclose(oracle_py.r2r(kind, x, direct=True), ref, rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(oracle_py.r2r(kind, x), ref, rtol=0, atol=1e-12 * n)


@pytest.mark.parametrize("n", [1, 2, 4, 5, 8, 64])
def test_r2hc_hc2r(n):
    x = np.random.default_rng(n).uniform(-1, 1, n)
    X = np.fft.rfft(x)
    hc = np.zeros(n)
    hc[: n // 2 + 1] = X.real
    for k in range(1, (n + 1) // 2):
        hc[n - k] = X[k].imag
    for direct in (True, False):
        np.testing.assert_allclose(oracle_py.r2r("R2HC", x, direct=direct), hc, atol=1e-12 * n)
        np.testing.assert_allclose(oracle_py.r2r("HC2R", hc, direct=direct), n * x, atol=1e-12 * n)


@pytest.mark.parametrize("pair", [("REDFT10", "REDFT01"), ("RODFT10", "RODFT01"), ("REDFT11", "REDFT11"),
                                  ("RODFT11", "RODFT11")])
def test_r2r_round_trip(pair):
    n = 16
    x = np.random.default_rng(3).uniform(-1, 1, n)
    y = oracle_py.r2r(pair[1], oracle_py.r2r(pair[0], x))
    np.testing.assert_allclose(y / (2 * n), x, atol=1e-14)


def second_difference(n, bc):
    """1D operator with unit spacing and the ghost rules of set_bc
    (centered): N -> ghost = interior, D -> ghost = -interior, F -> periodic."""
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = -2.0
        if i > 0:
            A[i, i - 1] = 1.0
        if i < n - 1:
            A[i, i + 1] = 1.0
    if bc == "FF":
        A[0, n - 1] += 1.0
        A[n - 1, 0] += 1.0
    else:
        A[0, 0] += 1.0 if bc[0] == "N" else -1.0
        A[n - 1, n - 1] += 1.0 if bc[1] == "N" else -1.0
    return A


@pytest.mark.parametrize("bc", ["FF", "NN", "DD", "ND", "DN"])
@pytest.mark.parametrize("n", [1, 2, 5, 8, 13])
def test_eigenvalues(bc, n):
    lam = np.sort(oracle_py.eigenvalues(n, bc))
    ref = np.sort(np.linalg.eigvalsh(second_difference(n, bc)))
    np.testing.assert_allclose(lam, ref, atol=1e-12)


@pytest.mark.parametrize("bc", ["FF", "NN", "DD", "ND", "DN"])
def test_transform_diagonalises(bc):
    """the forward kind of find_fft maps the 1D operator to lambda_k times
    the transformed vector, slot by slot (the ordering the per-mode solves
    rely on)."""
    n = 8
    kinds = {"FF": ("R2HC", "HC2R"), "NN": ("REDFT10", "REDFT01"), "DD": ("RODFT10", "RODFT01"),
             "ND": ("REDFT11", "REDFT11"), "DN": ("RODFT11", "RODFT11")}[bc]
    A = second_difference(n, bc)
    lam = oracle_py.eigenvalues(n, bc)
    x = np.random.default_rng(5).uniform(-1, 1, n)
    np.testing.assert_allclose(oracle_py.r2r(kinds[0], A @ x), lam * oracle_py.r2r(kinds[0], x), atol=1e-12)


def small_case(tmp_path, blocks, fft_axis, visci=100.0):
    d, b = cases.write_case(str(tmp_path), blocks, visci)
    return oracle_py.OracleCase(nmlparse.load_case(d, b, fft_axis))


def dense_operator(oc, mode):
    n = oc.nunk()
    A = np.zeros((n, n))
    for q in range(n):
        e = np.zeros(n); e[q] = 1.0
        A[:, q] = oc.apply_operator(mode, e)
    return A


@pytest.mark.parametrize("geom", ["channel", "cavity", "ldc"])
def test_operator_and_solvers(tmp_path, geom):
    import scipy.linalg as sl
    if geom == "channel":
        oc = small_case(tmp_path, [cases.channel_block((8, 6, 4))], 3)
        modes = range(4)
    elif geom == "cavity":
        oc = small_case(tmp_path, [cases.cavity_block((5, 4, 6))], 0)
        modes = [0]
    else:
        g = os.path.join(os.path.dirname(__file__), "golden", "ldc")
        oc = oracle_py.OracleCase(nmlparse.load_case(os.path.join(g, "dns.nml"), os.path.join(g, "blocks.nml"), 1))
        modes = [0, 1]
    rng = np.random.default_rng(0)
    for m in modes:
        A = dense_operator(oc, m)
        np.testing.assert_allclose(A, A.T, atol=1e-12 * np.abs(A).max())
        w = np.linalg.eigvalsh(A)
        assert w.max() <= 1e-9 * np.abs(w).min() + 1e-9 * np.abs(w).max()  # negative semi-definite
        singular = abs(w).min() < 1e-10 * abs(w).max()
        b = rng.uniform(-1, 1, oc.nunk())
        if singular:
            np.testing.assert_allclose(A @ np.ones(oc.nunk()), 0.0, atol=1e-10 * np.abs(A).max())
            b -= b.mean()
        sols = {}
        for method in ("band", "cg", "tensor"):
            if method == "tensor" and geom == "ldc":
                continue
            oc.set_solver(method, tol=1e-14)
            x = oc.solve_operator(m, b)
            r = A @ x - b
            assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b), (method, m)
            sols[method] = x - x.mean() if singular else x
        if not singular:
            xd = sl.solve(A, b)
            for method, x in sols.items():
                np.testing.assert_allclose(x, xd, atol=1e-10 * np.abs(xd).max(), err_msg=method)
        for method, x in sols.items():
            np.testing.assert_allclose(x, sols["band"], atol=1e-9 * np.abs(sols["band"]).max(), err_msg=method)


def test_linear_field_has_no_diffusion(tmp_path):
    """momx/y/z_d of u = a*y + b*z on a uniform triply periodic-free box is
    zero in the interior (second differences of a linear field)."""
    blk = cases.cavity_block((6, 6, 6))
    oc = small_case(tmp_path, [blk], 0, visci=1.0)
    shp = oc.shape(0, "u")
    y = oc.metric(0, 8)  # yc
    z = oc.metric(0, 10)  # zc
    u = 0.3 * y[None, :, None] + 0.7 * z[None, None, :] + np.zeros(shp)
    oc.set(0, "u", u)
    for f in ("v", "w", "p", "dudtrko", "dvdtrko", "dwdtrko"):
        oc.set(0, f, np.zeros(oc.shape(0, f)))
    # advection of u alone (v = w = 0) is -d(uu)/dx = 0 (no x dependence)
    oc.rk_mom(1, 1.0)
    np.testing.assert_allclose(oc.get(0, "dudtrko"), 0.0, atol=1e-12)


def periodic_box(tmp_path, n=(8, 8, 8), visci=1.0):
    blk = cases.channel_block(n, lengths=(1.0, 1.0, 1.0), gr_y=0.0)
    for c in range(3):
        blk["cbcvel"][c] = [["F", "F"], ["F", "F"], ["F", "F"]]
        blk["bcvel"][c] = [[1.0, 1.0], [1.0, 1.0], [1.0, 1.0]]
    blk["cbcpre"] = [["F", "F"], ["F", "F"], ["F", "F"]]
    blk["bcpre"] = [[1.0, 1.0], [1.0, 1.0], [1.0, 1.0]]
    return small_case(tmp_path, [blk], 3, visci=visci)


def zero_rest(oc, names):
    for f in names:
        oc.set(0, f, np.zeros(oc.shape(0, f)))


def test_diffusion_closed_form(tmp_path):
    """u = sin(2 pi y) on a uniform periodic box: momx_d gives exactly
    -visc*4/dy^2*sin^2(pi dy)*u (the discrete Laplacian's eigenvalue)."""
    n = 8
    oc = periodic_box(tmp_path, (n, n, n), visci=2.0)
    yc = oc.metric(0, 8)
    u = np.sin(2 * np.pi * yc)[None, :, None] + np.zeros(oc.shape(0, "u"))
    oc.set(0, "u", u)
    zero_rest(oc, ("v", "w", "p", "dudtrko", "dvdtrko", "dwdtrko"))
    oc.bounduvw(False)
    oc.rk_mom(1, 1.0)
    dy = 1.0 / n
    expect = -0.5 * 4.0 / dy ** 2 * np.sin(np.pi * dy) ** 2 * u[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(oc.get(0, "dudtrko"), expect, atol=1e-12 * np.abs(expect).max())


def test_advection_closed_form(tmp_path):
    """uniform u = U0 advecting v = sin(2 pi x): momy_a_vv is the central
    difference -U0 (v(i+1)-v(i-1)) / (2 dx), i.e. -U0 sin(2 pi dx)/dx cos(2 pi x)."""
    n = 16
    U0 = 0.7
    oc = periodic_box(tmp_path, (n, 8, 8), visci=1e300)
    xc = oc.metric(0, 6)
    v = np.sin(2 * np.pi * xc)[:, None, None] + np.zeros(oc.shape(0, "v"))
    oc.set(0, "u", np.full(oc.shape(0, "u"), U0))
    oc.set(0, "v", v)
    zero_rest(oc, ("w", "p", "dudtrko", "dvdtrko", "dwdtrko"))
    oc.bounduvw(False)
    oc.rk_mom(1, 1.0)
    dx = 1.0 / n
    expect = -U0 * np.sin(2 * np.pi * dx) / dx * np.cos(2 * np.pi * xc[1:-1])[:, None, None] + np.zeros((n, 8, 8))
    np.testing.assert_allclose(oc.get(0, "dvdtrko"), expect, atol=1e-11)
    np.testing.assert_allclose(oc.get(0, "dudtrko"), 0.0, atol=1e-11)


def test_projection_makes_divergence_free(tmp_path):
    """fillps -> solve -> correc leaves a discretely divergence-free field
    (the solve is exact to rounding)."""
    oc = small_case(tmp_path, [cases.channel_block((8, 6, 8))], 3)
    rng = np.random.default_rng(4)
    for f in ("u", "v", "w"):
        oc.set(0, f, rng.uniform(-1, 1, oc.shape(0, f)))
    oc.bounduvw(False)
    oc.fillps(1.0)
    oc.solve_pressure()
    oc.boundp("pp")
    oc.correc(1.0)
    oc.bounduvw(True)
    tot, mx = oc.chkdiv()
    assert mx < 1e-12


def test_divtot_closed_form(tmp_path):
    """chkdiv (chkdiv.f90:31-43) on u = x^2, v = y^2, w = z^2 at the faces of
    a box stretched along every axis: the divergence of cell (i,j,k) is
    (xf_i^2 - xf_{i-1}^2)/dxf_i + ..., so the volume-weighted sum telescopes
    to divtot = (Lx^2 Ly Lz + Lx Ly^2 Lz + Lx Ly Lz^2) / vol_all = Lx+Ly+Lz,
    and divmax = (xf_n + xf_{n-1}) + (yf_n + yf_{n-1}) + (zf_n + zf_{n-1}).
    A wrong weight (dxc for dxf, a missing factor or 1/vol_all) fails."""
    L = (2.0, 1.0, 1.5)
    oc = small_case(tmp_path, [cases.quadratic_box((12, 10, 8), L)], 0)
    xf, yf, zf = oc.metric(0, 7), oc.metric(0, 9), oc.metric(0, 11)
    for f, a in zip(("u", "v", "w"), cases.quadratic_fields(oc.shape(0, "u"), xf, yf, zf)):
        oc.set(0, f, a)
    tot, mx = oc.chkdiv()
    assert tot == pytest.approx(sum(L), rel=1e-13)
    n = (12, 10, 8)
    expect = sum(f[q] + f[q - 1] for f, q in zip((xf, yf, zf), n))
    assert mx == pytest.approx(expect, rel=1e-13)

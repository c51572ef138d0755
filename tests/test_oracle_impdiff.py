"""Pins of the oracle's implicit diffusion (the reference's _IMPDIFF build:
rk.f90:29-127, solve_helmholtz.f90:30-136, solver.f90:44-372,
updt_pressure.f90:30-45, chkdt.f90:51-55, main.f90:769-1031) against what the
mathematics fixes:

* Helmholtz solves (I + alpha L) x = phi of every component on discrete
  eigenmodes of its own staggered operator: x = phi / (1 - alpha kappa^2),
  with kappa^2 the closed-form eigenvalue (periodic, centred Dirichlet,
  staggered Dirichlet with the boundary face excluded, staggered Neumann);
  along the FFT axis and on the non-FFT axes, with and without FFT.
* inhomogeneous boundary values: a field that the discrete Laplacian with
  those boundary conditions maps to zero (constants under Dirichlet values,
  linear profiles under Neumann gradients) is returned unchanged.
* one RK3 step of a shear wave on a periodic box: advection and pressure
  vanish, each substage multiplies the wave by (1 - a)/(1 + a) with
  a = nu dtrk kappa^2 / 2; the same for a scalar wave.
* updt_pressure: p += pp - alpha kappa^2 pp on a Laplacian eigenmode.
* chkdt: without the diffusive limit.
"""
import numpy as np
import pytest

import cases
import nmlparse
import oracle_py

TWO_PI = 2.0 * np.pi


def make_case(tmp_path, blocks, fft_axis, visci=10.0, **dns):
    d, b = cases.write_case(str(tmp_path), blocks, visci, **dns)
    c = nmlparse.load_case(d, b, fft_axis)
    c["impdiff"] = True
    return oracle_py.OracleCase(c)


def coords(oc, b=0):
    """cell centres and faces per axis (n+2 entries, index 0 = lo-1)"""
    return [oc.metric(b, 6 + 2 * a) for a in range(3)], [oc.metric(b, 7 + 2 * a) for a in range(3)]


def box_block(ng, L, vel, pre, gr=(0.0, 0.0, 0.0), vals=None):
    return cases._block(ng, (0, 0, 0), L, vel=vel, pre=pre, gr=gr, vals=vals)


F1 = ("F", 1)


def sep_mode(oc, comp, kinds):
    """separable mode of component comp (0..2 = u,v,w): kinds[a] = (fn, k)
    with fn(pos) along axis a, staggered (faces) along the component's axis.
    Returns the ghosted field and kappa^2 given by kinds[a][2]."""
    xc, xf = coords(oc)
    pos = [xf[a] if a == comp else xc[a] for a in range(3)]
    f = [kinds[a][0](pos[a]) for a in range(3)]
    fld = f[0][:, None, None] * f[1][None, :, None] * f[2][None, None, :]
    return fld, sum(k[1] for k in kinds)


def per_mode(L, n, m):
    h = L / n
    k = TWO_PI * m / L
    return (lambda x: np.cos(k * x + 0.3)), 4.0 / h ** 2 * np.sin(k * h / 2) ** 2


def dir_mode(L, n, m):
    """sin(m pi x / L): the centred-D (ghost = -interior) and the staggered-D
    (boundary faces zero) eigenvector"""
    h = L / n
    return (lambda x: np.sin(m * np.pi * x / L)), 4.0 / h ** 2 * np.sin(m * np.pi * h / (2 * L)) ** 2


def stag_neu_mode(L, n, m):
    """staggered NN on the faces 1..n-1 (ghost faces copy their neighbour):
    cos(m pi (i - 1/2) / (n - 1)) at face i"""
    h = L / n
    return (lambda x: np.cos(m * np.pi * (x / h - 0.5) / (n - 1))), \
        4.0 / h ** 2 * np.sin(m * np.pi / (2 * (n - 1))) ** 2


def check_solve(oc, comp, phi, kap2, alpha, excl_axis=None):
    name = "uvw"[comp]
    oc.set(0, name, phi)
    oc.solve_impdiff(name, alpha)
    x = oc.get(0, name)
    sl = [slice(1, -1)] * 3
    if excl_axis is not None:  # the bounded hi face is not an unknown
        sl[excl_axis] = slice(1, -2)
    sl = tuple(sl)
    ref = phi[sl] / (1.0 - alpha * kap2)
    np.testing.assert_allclose(x[sl], ref, rtol=0, atol=1e-12 * np.abs(phi).max())


@pytest.mark.parametrize("fft_axis", [0, 1, 3])
@pytest.mark.parametrize("comp", [0, 1, 2])
def test_periodic_eigenmode(tmp_path, fft_axis, comp):
    n, L = (8, 6, 4), (TWO_PI, 3.0, 2.0)
    oc = make_case(tmp_path, [box_block(n, L, [[F1, F1]] * 3, [[F1, F1]] * 3)], fft_axis)
    kinds = [per_mode(L[a], n[a], m) for a, m in zip(range(3), (1, 2, 1))]
    phi, k2 = sep_mode(oc, comp, kinds)
    check_solve(oc, comp, phi, k2, alpha=-0.05)


@pytest.mark.parametrize("fft_axis", [0, 3])
@pytest.mark.parametrize("comp", [0, 1, 2])
def test_wall_eigenmode(tmp_path, fft_axis, comp):
    """Dirichlet walls on every face: centred along the tangential axes,
    staggered (boundary face excluded) along the component's own axis;
    with FFT_AXIS = 3 the z transform is RODFT10 (u, v) or RODFT00 over
    n-1 points (w)."""
    n, L = (6, 5, 8), (1.0, 1.5, 2.0)
    oc = make_case(tmp_path, [box_block(n, L, [["D", "D"]] * 3, [["N", "N"]] * 3)], fft_axis)
    kinds = [dir_mode(L[a], n[a], m) for a, m in zip(range(3), (2, 1, 3))]
    phi, k2 = sep_mode(oc, comp, kinds)
    check_solve(oc, comp, phi, k2, alpha=-0.02, excl_axis=comp)


def test_staggered_neumann_eigenmode(tmp_path):
    """u with Neumann walls in x (staggered NN, faces 1..n-1), periodic y, z"""
    n, L = (7, 4, 4), (1.4, 1.0, 1.0)
    blk = box_block(n, L, [["N", "N"], [F1, F1], [F1, F1]], [["N", "N"], [F1, F1], [F1, F1]])
    oc = make_case(tmp_path, [blk], 0)
    kinds = [stag_neu_mode(L[0], n[0], 2), per_mode(L[1], n[1], 1), per_mode(L[2], n[2], 1)]
    phi, k2 = sep_mode(oc, 0, kinds)
    check_solve(oc, 0, phi, k2, alpha=-0.03, excl_axis=0)


def test_staggered_neumann_fft_axis_rejected(tmp_path):
    """fft.f90's eigenvalues do not match its n-1 point transforms for a
    staggered NN pair: rejected"""
    n, L = (8, 4, 4), (1.0, 1.0, 1.0)
    blk = box_block(n, L, [["N", "N"], [F1, F1], [F1, F1]], [["N", "N"], [F1, F1], [F1, F1]])
    d, b = cases.write_case(str(tmp_path), [blk], 10.0)
    c = oracle_py.OracleCase(nmlparse.load_case(d, b, 1))
    with pytest.raises(ValueError, match="staggered"):
        c.set_impdiff(True)


@pytest.mark.parametrize("fft_axis", [0, 3])
def test_dirichlet_values_keep_constants(tmp_path, fft_axis):
    """u = U0 with U0 on every wall (v = V0, w = W0 alike): L maps the
    constant to zero, so the solve returns it; pins the -2 bc (centred) and
    -bc (staggered) factors of init_bc_rhs"""
    n, L = (6, 5, 8), (1.0, 1.5, 2.0)
    vals = {(c, d, s): (0.7, -0.4, 1.3)[c] for c in range(3) for d in range(3) for s in range(2)}
    oc = make_case(tmp_path, [box_block(n, L, [["D", "D"]] * 3, [["N", "N"]] * 3, vals=vals)], fft_axis)
    for comp in range(3):
        U0 = (0.7, -0.4, 1.3)[comp]
        phi = np.full(oc.shape(0, "u"), U0)
        sl = [slice(1, -1)] * 3
        sl[comp] = slice(1, -2)
        oc.set(0, "uvw"[comp], phi)
        oc.solve_impdiff("uvw"[comp], -0.05)
        np.testing.assert_allclose(oc.get(0, "uvw"[comp])[tuple(sl)], U0, rtol=1e-12)


def test_neumann_values_keep_linear_profiles(tmp_path):
    """v = g z (and w = g y on the stretched axis excluded): a Neumann
    gradient g on the z walls (centred for v) leaves the linear profile
    unchanged; u = g x with staggered Neumann walls in x likewise"""
    n, L = (6, 4, 8), (1.2, 1.0, 2.0)
    g = 0.8
    vel = [["N", "N"], [F1, F1], ["N", "N"]]
    vals = {(1, 2, 0): g, (1, 2, 1): g, (0, 0, 0): g, (0, 0, 1): g}
    blk = box_block(n, L, vel, [["N", "N"], [F1, F1], ["N", "N"]], vals=vals)
    oc = make_case(tmp_path, [blk], 0)
    xc, xf = coords(oc)
    phi = np.zeros(oc.shape(0, "v"))
    phi[:] = g * xc[2][None, None, :]
    oc.set(0, "v", phi)
    oc.solve_impdiff("v", -0.05)
    np.testing.assert_allclose(oc.get(0, "v")[1:-1, 1:-1, 1:-1], phi[1:-1, 1:-1, 1:-1], rtol=0, atol=1e-12)
    phi = np.zeros(oc.shape(0, "u"))
    phi[:] = g * xf[0][:, None, None] + 0.25
    oc.set(0, "u", phi)
    oc.solve_impdiff("u", -0.05)
    np.testing.assert_allclose(oc.get(0, "u")[1:-2, 1:-1, 1:-1], phi[1:-2, 1:-1, 1:-1], rtol=0, atol=1e-12)


@pytest.mark.parametrize("fft_axis", [0, 2])
def test_step_shear_wave(tmp_path, fft_axis):
    """u = A sin(y) and a scalar s = B cos(2 pi z / Lz) on a periodic box at rest
    otherwise: each substage applies (1 - a)/(1 + a), a = nu dtrk kappa^2/2
    (rk_mom / rk_scal keep half the diffusion, the Helmholtz solve takes
    the rest); v, w, p stay zero."""
    n = (8, 16, 4)
    L = (TWO_PI, TWO_PI, 1.0)
    visci, alphai = 4.0, 2.5
    blk = box_block(n, L, [[F1, F1]] * 3, [[F1, F1]] * 3)
    blk["cbcscal"] = [[["F", "F"]] * 3]
    blk["bcscal"] = [[[1.0, 1.0]] * 3]
    d, b = cases.write_case(str(tmp_path), [blk], visci, nscal=1,
                            scalar=dict(alphai=[alphai], beta=0.0, iniscal=["zer"], ssource=[0.0],
                                        is_sforced=[False], scalf=[0.0]))
    c = nmlparse.load_case(d, b, fft_axis)
    c["impdiff"] = True
    oc = oracle_py.OracleCase(c)
    xc, xf = coords(oc)
    u = np.zeros(oc.shape(0, "u"))
    u[:] = 0.9 * np.sin(xc[1])[None, :, None]
    s = np.zeros(oc.shape(0, "s1"))
    s[:] = 0.6 * np.cos(TWO_PI * xc[2] / L[2])[None, None, :]
    oc.set(0, "u", u)
    oc.set(0, "s1", s)
    oc.init_bounds()
    dt = 0.05
    oc.step(dt)
    ky2 = 4.0 / (L[1] / n[1]) ** 2 * np.sin((L[1] / n[1]) / 2) ** 2
    hz = L[2] / n[2]
    kz2 = 4.0 / hz ** 2 * np.sin(np.pi * hz / L[2]) ** 2
    gu = gs = 1.0
    for a1, a2 in oracle_py.RKCOEFF:
        dtrk = (a1 + a2) * dt
        au = dtrk * ky2 / visci / 2
        as_ = dtrk * kz2 / alphai / 2
        gu *= (1 - au) / (1 + au)
        gs *= (1 - as_) / (1 + as_)
    in3 = (slice(1, -1),) * 3
    np.testing.assert_allclose(oc.get(0, "u")[in3], gu * u[in3], rtol=0, atol=1e-13)
    np.testing.assert_allclose(oc.get(0, "s1")[in3], gs * s[in3], rtol=0, atol=1e-13)
    for f in ("v", "w", "p"):
        assert np.abs(oc.get(0, f)[in3]).max() < 1e-13


def test_updt_pressure_laplacian(tmp_path):
    n, L = (8, 6, 4), (TWO_PI, 3.0, 2.0)
    oc = make_case(tmp_path, [box_block(n, L, [[F1, F1]] * 3, [[F1, F1]] * 3)], 0)
    xc, _ = coords(oc)
    kinds = [per_mode(L[a], n[a], m) for a, m in zip(range(3), (1, 1, 1))]
    pp = kinds[0][0](xc[0])[:, None, None] * kinds[1][0](xc[1])[None, :, None] * kinds[2][0](xc[2])[None, None, :]
    k2 = sum(k[1] for k in kinds)
    p0 = np.random.default_rng(1).uniform(-1, 1, pp.shape)
    oc.set(0, "pp", pp)
    oc.set(0, "p", p0)
    alpha = -0.07
    oc.updt_pressure_impdiff(alpha)
    in3 = (slice(1, -1),) * 3
    np.testing.assert_allclose(oc.get(0, "p")[in3], p0[in3] + pp[in3] * (1 - alpha * k2), rtol=0, atol=1e-12)


def test_chkdt_without_diffusive_limit(tmp_path):
    n, L = (8, 6, 4), (1.0, 1.0, 1.0)
    oc = make_case(tmp_path, [box_block(n, L, [[F1, F1]] * 3, [[F1, F1]] * 3)], 0, visci=0.01)
    U = 2.0
    oc.set(0, "u", np.full(oc.shape(0, "u"), U))
    h = [L[a] / n[a] for a in range(3)]
    dti = max(U / h[0], U / h[0], U / h[0])
    assert oc.chkdt() == pytest.approx(np.sqrt(3.0) / dti, rel=1e-14)
    oc.set_impdiff(False)
    assert oc.chkdt() < np.sqrt(3.0) / dti  # the diffusive limit of visc = 100


def test_inflow_plane_scaling(tmp_path):
    """An inflow face (init_bc_rhs with the inflow plane, main.f90:600-636):
    v = uref = 1.7 entering at y = 0 with periodic x, z (a uniform profile)
    and zero-gradient outflow at y = Ly is a constant the Laplacian maps to
    zero, so the solve returns it; a dropped or inverted bc_array / bc factor
    would not"""
    uref = 1.7
    blk = cases._block((6, 8, 4), (0, 0, 0), (1.0, 2.0, 1.0), vel=[[F1, F1], ["D", "N"], [F1, F1]],
                       pre=[[F1, F1], ["N", "D"], [F1, F1]], vals={(1, 1, 0): uref}, inflow={(1, 0): 1})
    for fft_axis in (0, 1):
        oc = make_case(tmp_path / str(fft_axis), [blk], fft_axis)
        oc.set(0, "v", np.full(oc.shape(0, "v"), uref))
        oc.solve_impdiff("v", -0.04)
        np.testing.assert_allclose(oc.get(0, "v")[1:-1, 1:-2, 1:-1], uref, rtol=1e-13)


@pytest.mark.parametrize("geom", ["bfs", "lchannel"])
@pytest.mark.parametrize("comp", [0, 1, 2])
def test_union_helmholtz_self_adjoint(tmp_path, comp, geom):
    """On the backward-facing step and the L-shaped channel (unions of blocks
    that are not boxes) the Helmholtz solve S of every component is self-adjoint in the weighted
    inner product: S = (A + W/alpha)^-1 W/alpha with A the symmetric
    assembled operator, so <phi2, W S phi1> = <phi1, W S phi2> for the
    linear part of S (S(0), from the inflow plane, subtracted).  The CG
    solver applies the assembled links as they are (the banded Cholesky
    reads one triangle only, so it is self-adjoint by construction) and must
    agree with the banded solve.  v meets the one link that leaves the
    struct grid: the lower block's top face (an interface) next to the
    inlet's top wall face, which is not an unknown; indexing it like an
    unknown of the inlet would link two cells that are not neighbours, one
    way only.  On the L-shaped channel u meets the same situation (the upper
    left block's x+ face is an interface, the lower block's a wall) and the
    FFT axis is bounded: sine transforms, w over n-1 faces."""
    blocks = cases.bfs_blocks((6, 4), (8, 4, 6), 4) if geom == "bfs" else cases.lchannel_blocks(4, 6)
    oc = make_case(tmp_path, blocks, 3, visci=100.0)
    name = "uvw"[comp]
    nb = oc.nblocks

    def sl(q):
        s = [slice(1, -1)] * 3
        if oc.block_info(q)["nb"][comp][1] < 0:
            s[comp] = slice(1, -2)
        return tuple(s)

    def weight(q):
        """dx2 of the non-FFT axes (z, the FFT axis, is uniform)"""
        mx, my = (oc.metric(q, 2 * a if a == comp else 2 * a + 1)[1:-1] for a in range(2))
        return (mx[:, None] * my[None, :])[..., None]

    def solve(fields):
        for q in range(nb):
            oc.set(q, name, fields[q])
        oc.solve_impdiff(name, -0.05)
        return [oc.get(q, name) for q in range(nb)]

    def dot(a, b):
        tot = 0.0
        for q in range(nb):
            u, v = a[q][sl(q)], b[q][sl(q)]
            tot += np.sum(weight(q)[:u.shape[0], :u.shape[1]] * u * v)
        return tot

    rng = np.random.default_rng(5 + comp)
    phi = []
    for _ in range(2):
        f = []
        for q in range(nb):
            a = np.zeros(oc.shape(q, name))
            a[sl(q)] = rng.uniform(-1, 1, a[sl(q)].shape)
            f.append(a)
        phi.append(f)
    zero = [np.zeros(oc.shape(q, name)) for q in range(nb)]
    res = {}
    for method in ("cg", "band"):
        oc.set_solver(method, tol=1e-15, maxiter=5000)
        s0 = solve(zero)
        res[method] = [[s - z for s, z in zip(solve(p), s0)] for p in phi]
    x = res["cg"]
    d12, d21 = dot(phi[1], x[0]), dot(phi[0], x[1])
    assert abs(d12 - d21) <= 1e-10 * max(abs(d12), abs(d21)), (d12, d21)
    for i in range(2):
        for q in range(nb):
            np.testing.assert_allclose(x[i][q][sl(q)], res["band"][i][q][sl(q)], rtol=0, atol=1e-10)

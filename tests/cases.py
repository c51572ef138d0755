"""Case definitions for the tests and the benchmark, written as the
reference's own input files (dns.nml / blocks.nml, param.f90:88-106), so
that the product reads exactly what the reference would.

channel(): the periodic channel of the benchmark configuration (x periodic,
y wall-normal tanh-stretched, z periodic and the FFT axis), any size.
cavity(): the 3D lid-driven cavity (all walls, lid u=1 on z+, no FFT).
"""
import os

import numpy as np


def _fmt(v):
    if isinstance(v, bool):
        return "T" if v else "F"
    if isinstance(v, str):
        return "'" + v + "'"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


def write_dns(path, visci, nstep=10, dt_f=-1.0, cfl=0.5, dtmax=1e5, icheck=10, bforce=(0, 0, 0), nscal=0,
              gacc=(0, 0, 0), scalar=None):
    """scalar: the &scalar group as {name: value or list} (alphai, beta,
    iniscal, ssource, is_sforced, scalf), written when nscal > 0."""
    sc = ""
    if nscal > 0:
        sc = "&scalar\n"
        for k, v in (scalar or {}).items():
            if isinstance(v, (list, tuple)):
                sc += f"{k}(1:{len(v)}) = {', '.join(_fmt(x) for x in v)}\n"
            else:
                sc += f"{k} = {_fmt(v)}\n"
        sc += "/\n"
    txt = ("&dns\n"
           f"cfl = {_fmt(cfl)}, dtmax = {_fmt(dtmax)}, dt_f = {_fmt(dt_f)}\n"
           f"visci = {_fmt(visci)}\n"
           f"nstep = {nstep}, time_max = 100., tw_max = 0.1\n"
           "stop_type(1:3) = T, F, F\n"
           "restart = F, is_overwrite_save = T, nsaves_max = 0\n"
           f"icheck = {icheck}, iout0d = 0, iout1d = 0, iout2d = 0, iout3d = 0, isave = 0\n"
           f"bforce(1:3) = {', '.join(_fmt(b) for b in bforce)}\n"
           f"gacc(1:3) = {', '.join(_fmt(g) for g in gacc)}\n"
           f"nscal = {nscal}\n"
           "/\n&hypre\nhypre_solver_i = 2, hypre_tol = 1.e-13, hypre_maxiter = 1000\n/\n" + sc)
    open(path, "w").write(txt)


def write_blocks(path, blocks):
    """blocks: list of dicts with ng, lmin, lmax, gt, gr, cbcvel[comp][dir][side],
    bcvel, cbcpre[dir][side], bcpre, inflow_type (optional)."""
    out = [f"&blocks\nnblocks = {len(blocks)}\n"]
    for b, B in enumerate(blocks, start=1):
        out.append(f"block_dims(1:3,{b}) = 1, 1, 1\n")
        for key in ("ng", "lmin", "lmax", "gt", "gr"):
            out.append(f"block_{key}(1:3,{b}) = {', '.join(_fmt(v) for v in B[key])}\n")
        for c in range(3):
            out.append(f"block_cbcvel(0:1,1:3,{c + 1},{b}) = " +
                       ", ".join(_fmt(B["cbcvel"][c][d][s]) for d in range(3) for s in range(2)) + "\n")
            out.append(f"block_bcvel(0:1,1:3,{c + 1},{b}) = " +
                       ", ".join(_fmt(B["bcvel"][c][d][s]) for d in range(3) for s in range(2)) + "\n")
        out.append(f"block_cbcpre(0:1,1:3,{b}) = " +
                   ", ".join(_fmt(B["cbcpre"][d][s]) for d in range(3) for s in range(2)) + "\n")
        out.append(f"block_bcpre(0:1,1:3,{b}) = " +
                   ", ".join(_fmt(B["bcpre"][d][s]) for d in range(3) for s in range(2)) + "\n")
        for i, (cbs, bs) in enumerate(zip(B.get("cbcscal", []), B.get("bcscal", [])), start=1):
            out.append(f"block_cbcscal(0:1,1:3,{i},{b}) = " +
                       ", ".join(_fmt(cbs[d][s]) for d in range(3) for s in range(2)) + "\n")
            out.append(f"block_bcscal(0:1,1:3,{i},{b}) = " +
                       ", ".join(_fmt(bs[d][s]) for d in range(3) for s in range(2)) + "\n")
        it = B.get("inflow_type", [[0, 0]] * 3)
        out.append(f"block_inflow_type(0:1,1:3,{b}) = " +
                   ", ".join(_fmt(it[d][s]) for d in range(3) for s in range(2)) + "\n")
        out.append(f"block_inivel({b}) = 'zer'\n\n")
    out.append("/\n")
    open(path, "w").write("".join(out))


def _vel(pattern, values):
    """pattern: [dir][side] BC chars for all three components, with the
    wall-normal component sharing them."""
    return [[[pattern[d][s] for s in range(2)] for d in range(3)] for _ in range(3)], values


def channel_block(ng, lengths=(6.0, 1.0, 3.0), gr_y=1.8, friend=1):
    F = float(friend)
    cbcv = [[["F", "F"], ["D", "D"], ["F", "F"]] for _ in range(3)]
    bcv = [[[F, F], [0.0, 0.0], [F, F]] for _ in range(3)]
    return dict(ng=list(ng), lmin=[0.0, 0.0, 0.0], lmax=list(lengths), gt=[0, 0, 0], gr=[0.0, gr_y, 0.0],
                cbcvel=cbcv, bcvel=bcv, cbcpre=[["F", "F"], ["N", "N"], ["F", "F"]],
                bcpre=[[F, F], [0.0, 0.0], [F, F]])


def cavity_block(ng, lid=1.0, gr=(0.0, 0.0, 0.0)):
    cbcv = [[["D", "D"], ["D", "D"], ["D", "D"]] for _ in range(3)]
    bcv = [[[0.0, 0.0], [0.0, 0.0], [0.0, 0.0]] for _ in range(3)]
    bcv[0][2][1] = lid  # u = lid on z+
    return dict(ng=list(ng), lmin=[0.0, 0.0, 0.0], lmax=[1.0, 1.0, 1.0], gt=[0, 0, 0], gr=list(gr),
                cbcvel=cbcv, bcvel=bcv, cbcpre=[["N", "N"], ["N", "N"], ["N", "N"]],
                bcpre=[[0.0, 0.0], [0.0, 0.0], [0.0, 0.0]])


def quadratic_box(ng, lengths=(2.0, 1.0, 1.5), gr=(1.3, 0.8, 1.1)):
    """A walled box stretched along all three axes (tanh two-end), for the
    chkdiv closed form: u = x^2, v = y^2, w = z^2 at the faces make every
    cell's flux differences telescope, so divtot = Lx + Ly + Lz."""
    B = cavity_block(ng, lid=0.0, gr=gr)
    B["lmax"] = list(lengths)
    return B


def quadratic_fields(shape, xf, yf, zf):
    """u = xf^2, v = yf^2, w = zf^2 on the ghosted arrays (ghosts included)."""
    u = np.zeros(shape) + (xf ** 2)[:, None, None]
    v = np.zeros(shape) + (yf ** 2)[None, :, None]
    w = np.zeros(shape) + (zf ** 2)[None, None, :]
    return u, v, w


def _block(ng, lmin, lmax, vel, pre, gr=(0.0, 0.0, 0.0), vals=None, inflow=None):
    """vel / pre: [dir][side] BC characters (shared by the three velocity
    components); vals: {(comp, dir, side): value} for velocity BC values and
    friend numbers are taken from `friends` entries ('F', b)."""
    cbcv = [[[vel[d][s] if isinstance(vel[d][s], str) else vel[d][s][0] for s in range(2)] for d in range(3)]
            for _ in range(3)]
    fr = [[0.0 if isinstance(vel[d][s], str) else float(vel[d][s][1]) for s in range(2)] for d in range(3)]
    bcv = [[[fr[d][s] for s in range(2)] for d in range(3)] for _ in range(3)]
    for (c, d, s), v in (vals or {}).items():
        bcv[c][d][s] = v
    cbcp = [[pre[d][s] if isinstance(pre[d][s], str) else pre[d][s][0] for s in range(2)] for d in range(3)]
    bcp = [[fr[d][s] for s in range(2)] for d in range(3)]
    B = dict(ng=list(ng), lmin=list(lmin), lmax=list(lmax), gt=[0, 0, 0], gr=list(gr),
             cbcvel=cbcv, bcvel=bcv, cbcpre=cbcp, bcpre=bcp)
    if inflow:
        it = [[0, 0], [0, 0], [0, 0]]
        for (d, s), t in inflow.items():
            it[d][s] = t
        B["inflow_type"] = it
    return B


def triperiodic_block(ng, lengths=(6.283185307179586,) * 3):
    """A periodic box along all three axes (examples/triperiodic)."""
    F = ("F", 1)
    return _block(ng, (0, 0, 0), lengths, vel=[[F, F]] * 3, pre=[[F, F]] * 3)


def lchannel_blocks(n=8, nz=16):
    """The reference's multi-block L-shaped channel
    (examples/__MULTI_BLOCK_GEOMETRY/L_channel/blocks.nml) at n x n cells per
    unit block: v = 1 enters at y = 0 under block 1, turns into block 3 and
    leaves at x = 2 (velocity N, pressure D); walls at z = 0, 1 (the FFT
    axis, a cosine transform)."""
    F = lambda b: ("F", b)
    b1 = _block((n, n, nz), (0, 0, 0), (0.5, 0.5, 1.0),
                vel=[["D", "D"], ["D", F(2)], ["D", "D"]], pre=[["N", "N"], ["N", F(2)], ["N", "N"]],
                vals={(1, 1, 0): 1.0})
    b2 = _block((n, n, nz), (0, 0.5, 0), (0.5, 1.0, 1.0),
                vel=[["D", F(3)], [F(1), "D"], ["D", "D"]], pre=[["N", F(3)], [F(1), "N"], ["N", "N"]])
    b3 = _block((3 * n, n, nz), (0.5, 0.5, 0), (2.0, 1.0, 1.0),
                vel=[[F(2), "N"], ["D", "D"], ["D", "D"]], pre=[[F(2), "D"], ["N", "N"], ["N", "N"]])
    return [b1, b2, b3]


def bfs_blocks(n_in=(192, 64), n_out=(320, 64, 96), nz=128, lengths=(3.0, 5.0, 1.0, 1.5, 2.0), gr=(1.2, 1.2)):
    """Backward-facing step of three blocks: inlet (n_in[0] x n_in[1]) with a
    Poiseuille inflow at x- (inflow_type 1), lower block (n_out[0] x n_out[1])
    downstream of it sharing its y-grid, upper block (n_out[0] x n_out[2]) on
    top of the lower one; outflow at x+ (velocity N, pressure D), walls in y
    (tanh stretching gr per wall-normal block) and at the step face, z
    periodic (the FFT axis).  lengths = (L_inlet, L_out, h_low, h_up, L_z)."""
    F = lambda b: ("F", b)
    lin, lout, h1, h2, lz = lengths
    inlet = _block((n_in[0], n_in[1], nz), (0, 0, 0), (lin, h1, lz),
                   vel=[["D", F(2)], ["D", "D"], [F(1), F(1)]], pre=[["N", F(2)], ["N", "N"], [F(1), F(1)]],
                   gr=(0.0, gr[0], 0.0), vals={(0, 0, 0): 1.0}, inflow={(0, 0): 1})
    lower = _block((n_out[0], n_out[1], nz), (lin, 0, 0), (lin + lout, h1, lz),
                   vel=[[F(1), "N"], ["D", F(3)], [F(2), F(2)]], pre=[[F(1), "D"], ["N", F(3)], [F(2), F(2)]],
                   gr=(0.0, gr[0], 0.0))
    upper = _block((n_out[0], n_out[2], nz), (lin, h1, 0), (lin + lout, h1 + h2, lz),
                   vel=[["D", "N"], [F(2), "D"], [F(3), F(3)]], pre=[["N", "D"], [F(2), "N"], [F(3), F(3)]],
                   gr=(0.0, gr[1], 0.0))
    return [inlet, lower, upper]


def write_case(dirpath, blocks, visci, **dns):
    os.makedirs(dirpath, exist_ok=True)
    d = os.path.join(dirpath, "dns.nml")
    b = os.path.join(dirpath, "blocks.nml")
    write_dns(d, visci, **dns)
    write_blocks(b, blocks)
    return d, b


def channel_ic(shape_u, yc, seed=20251202, amp=0.1):
    """Laminar Poiseuille u = 6 y (1-y) at the cell centres plus seeded uniform
    +-amp noise on u, v, w (interior), p = 0; ghosts are filled by the boundary
    routines afterwards."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = (s - 2 for s in shape_u)
    u = np.zeros(shape_u); v = np.zeros(shape_u); w = np.zeros(shape_u)
    prof = 6.0 * yc[1:-1] * (1.0 - yc[1:-1])
    u[1:-1, 1:-1, 1:-1] = prof[None, :, None]
    u[1:-1, 1:-1, 1:-1] += amp * rng.uniform(-1, 1, (nx, ny, nz))
    v[1:-1, 1:-1, 1:-1] = amp * rng.uniform(-1, 1, (nx, ny, nz))
    w[1:-1, 1:-1, 1:-1] = amp * rng.uniform(-1, 1, (nx, ny, nz))
    return u, v, w

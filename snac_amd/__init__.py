"""snac_amd -- MI355X (gfx950) time-step hot path of SNaC behind a C-ABI.

The compute lives in ``libsnac_amd.so`` (HIP kernels + C++ host set-up), built
in-tree by ``snac_amd/csrc/Makefile``.  This module is the Python mirror of the
reference's operator interface for the path: a :class:`Context` is created from
the same ``dns.nml``/``blocks.nml`` inputs and exposes the reference routines
by name (``rk_mom``, ``bounduvw``, ``boundp``, ``inflow``, ``fillps``,
``solve_pressure``, ``correc``, ``updt_pressure``, ``chkdiv``, ``chkdt``) plus a
whole RK3 ``step``.  There is no CPU fallback: if the library is missing the
import raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libsnac_amd.so")

FIELDS = {"u": 0, "v": 1, "w": 2, "p": 3, "pp": 4, "po": 5, "dudtrko": 6, "dvdtrko": 7, "dwdtrko": 8}
SCAL0 = 16
for _i in range(1, 9):  # scalars s1.. (scalars(i)%val) and their dsdtrko1..
    FIELDS[f"s{_i}"] = SCAL0 + 2 * (_i - 1)
    FIELDS[f"dsdtrko{_i}"] = SCAL0 + 2 * (_i - 1) + 1


def _ghosted(fid):
    f = FIELDS[fid]
    return f < 6 or (f >= SCAL0 and (f - SCAL0) % 2 == 0)
RKCOEFF = ((32.0 / 60.0, 0.0), (25.0 / 60.0, -17.0 / 60.0), (45.0 / 60.0, -25.0 / 60.0))
PHASES = ("rk_mom", "bound", "fillps", "r2r_fwd", "tridiag", "r2r_bwd", "correc", "rk_scal", "helmholtz",
          "mg_line0", "mg_tail", "mg_update")
R2R_KINDS = {"R2HC": 0, "HC2R": 1, "REDFT00": 3, "REDFT01": 4, "REDFT10": 5, "REDFT11": 6,
             "RODFT00": 7, "RODFT01": 8, "RODFT10": 9, "RODFT11": 10}


class SnacError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} is missing: build it with `make -C snac_amd/csrc` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(LIBPATH)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    vp = C.c_void_p
    L.snac_last_error.restype = C.c_char_p
    L.snac_ctx_create_from_nml.argtypes = [C.POINTER(vp), C.c_char_p, C.c_char_p, C.c_int, C.c_int]
    L.snac_ctx_destroy.argtypes = [vp]
    L.snac_ctx_num_blocks.argtypes = [vp]
    L.snac_ctx_block_info.argtypes = [vp, C.c_int, ip, ip, ip, ip]
    L.snac_ctx_get_metric.argtypes = [vp, C.c_int, C.c_int, dp]
    L.snac_ctx_get_params.argtypes = [vp, dp, ip]
    L.snac_ctx_set_solver.argtypes = [vp, C.c_int]
    L.snac_set_precision.argtypes = [vp, C.c_int]
    L.snac_solver_stats.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.snac_solver_counters.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.snac_set_graph_cg_iters.argtypes = [vp, C.c_int]
    L.snac_upload_field.argtypes = [vp, C.c_int, C.c_int, dp]
    L.snac_download_field.argtypes = [vp, C.c_int, C.c_int, dp]
    i3 = np.ctypeslib.ndpointer(dtype=np.int32, shape=(3,), flags="C_CONTIGUOUS")
    L.snac_load_one.argtypes = [vp, C.c_char, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                C.POINTER(C.c_int)]
    L.snac_restart_io.argtypes = [vp, C.c_char, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.snac_write_visu_3d.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, i3, i3,
                                     i3, C.c_double, C.c_int]
    L.snac_save_grid.argtypes = [vp, C.c_char_p]
    L.snac_rk_mom.argtypes = [vp, C.c_int, C.c_double]
    L.snac_rk_scal.argtypes = [vp, C.c_int, C.c_double]
    L.snac_scalar_forcing.argtypes = [vp, dp]
    L.snac_bounduvw.argtypes = [vp, C.c_int]
    L.snac_inflow.argtypes = [vp, C.c_int]
    L.snac_boundp.argtypes = [vp, C.c_int]
    L.snac_fillps.argtypes = [vp, C.c_double]
    L.snac_solve_pressure.argtypes = [vp]
    L.snac_correc.argtypes = [vp, C.c_double]
    L.snac_updt_pressure.argtypes = [vp]
    L.snac_set_impdiff.argtypes = [vp, C.c_int]
    L.snac_solve_impdiff.argtypes = [vp, C.c_int, C.c_double]
    L.snac_impdiff_stats.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.snac_updt_pressure_impdiff.argtypes = [vp, C.c_double]
    L.snac_chkdiv.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.snac_chkdt.argtypes = [vp, C.POINTER(C.c_double)]
    L.snac_init_bounds.argtypes = [vp]
    L.snac_initflow.argtypes = [vp]
    L.snac_set_buoyancy.argtypes = [vp, C.c_int]
    L.snac_step.argtypes = [vp, C.c_double]
    L.snac_steps.argtypes = [vp, C.c_double, C.c_int]
    L.snac_sync.argtypes = [vp]
    L.snac_set_graph.argtypes = [vp, C.c_int]
    L.snac_r2r_lines.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
    L.snac_stream_copy.argtypes = [vp, C.c_long, C.c_int, C.POINTER(C.c_double)]
    L.snac_get_autotune.argtypes = [vp, ip, dp, C.c_int]
    L.snac_mg_line_choices.argtypes = [vp, ip, ip, C.c_int]
    L.snac_set_kernel_variant.argtypes = [vp, C.c_int, C.c_int]
    L.snac_set_timing.argtypes = [vp, C.c_int]
    L.snac_get_timing.argtypes = [vp, dp, ip]
    fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.snac_learned_collect.argtypes = [vp, C.c_int]
    L.snac_learned_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_uint, fp]
    L.snac_learned_options.argtypes = [vp, C.c_int, C.c_double, C.c_int]
    L.snac_learned_stats.argtypes = [vp, ip, ip]
    L.snac_learned_time.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.snac_net_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_uint, C.c_int]
    L.snac_net_destroy.argtypes = [vp]
    L.snac_net_conv_variant.argtypes = [C.c_int]
    L.snac_net_num_params.argtypes = [vp]
    L.snac_net_num_params.restype = C.c_long
    L.snac_net_get_params.argtypes = [vp, fp]
    L.snac_net_set_params.argtypes = [vp, fp]
    L.snac_net_forward.argtypes = [vp, C.c_int, fp, fp]
    L.snac_net_loss_grad.argtypes = [vp, C.c_int, fp, fp, C.POINTER(C.c_float), fp]
    L.snac_net_train.argtypes = [vp, C.c_int, fp, fp, C.c_int, C.c_int, C.c_float, fp]
    return L


lib = _load()


def _check(rc, what):
    if rc != 0:
        raise SnacError(f"{what}: {lib.snac_last_error().decode()}")


class Context:
    """One case on one GPU; mirrors the reference's hot-path routines."""

    def __init__(self, dns_nml, blocks_nml, fft_axis, device=0):
        h = C.c_void_p()
        _check(lib.snac_ctx_create_from_nml(C.byref(h), str(dns_nml).encode(), str(blocks_nml).encode(),
                                            int(fft_axis), int(device)), "snac_ctx_create_from_nml")
        self.h = h
        self.fft_axis = fft_axis
        self.nblocks = lib.snac_ctx_num_blocks(h)
        dpar = np.zeros(11)
        ipar = np.zeros(8, np.int32)
        _check(lib.snac_ctx_get_params(h, dpar, ipar), "snac_ctx_get_params")
        self.visc, self.cfl, self.dtmax, self.dt_f = dpar[:4]
        self.bforce = dpar[4:7].copy()
        self.vol_all = dpar[7]
        self.nstep, self.icheck = int(ipar[0]), int(ipar[1])
        self.nscal = int(ipar[2])

    def close(self):
        if getattr(self, "h", None):
            lib.snac_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- geometry ----
    def block_info(self, b):
        n = np.zeros(3, np.int32); lo = np.zeros(3, np.int32); nb = np.zeros(6, np.int32); per = np.zeros(3, np.int32)
        _check(lib.snac_ctx_block_info(self.h, b, n, lo, nb, per), "snac_ctx_block_info")
        return dict(n=n, lo_g=lo, nb=nb.reshape(3, 2), periods=per)

    def metric(self, b, which):
        n = self.block_info(b)["n"][(which % 6) // 2]
        out = np.zeros(n + 2)
        _check(lib.snac_ctx_get_metric(self.h, b, which, out), "snac_ctx_get_metric")
        return out

    def shape(self, b, fid):
        n = self.block_info(b)["n"]
        return tuple(int(x) + 2 for x in n) if _ghosted(fid) else tuple(int(x) for x in n)

    # ---- fields (Fortran-ordered numpy arrays, ghosts included) ----
    def get(self, b, fid):
        shp = self.shape(b, fid)
        out = np.zeros(int(np.prod(shp)))
        _check(lib.snac_download_field(self.h, b, FIELDS[fid], out), "snac_download_field")
        return out.reshape(shp, order="F")

    def set(self, b, fid, arr):
        shp = self.shape(b, fid)
        a = np.asarray(arr, dtype=np.float64)
        if a.shape != shp:
            raise ValueError(f"field {fid} of block {b} must have shape {shp}, got {a.shape}")
        _check(lib.snac_upload_field(self.h, b, FIELDS[fid], np.ascontiguousarray(a.reshape(-1, order="F"))),
               "snac_upload_field")

    # ---- files in the reference's formats ----
    def load_one(self, io, filename, b, fid, time=0.0, istep=0):
        """load_one (load.f90:84-187) of block b's field: io 'w' writes the
        interior then [time, istep]; 'r' checks the size, sets the interior
        (zero ghosts) and returns (time, istep)."""
        t, s = C.c_double(time), C.c_int(istep)
        _check(lib.snac_load_one(self.h, io.encode(), str(filename).encode(), b, FIELDS[fid], C.byref(t),
                                 C.byref(s)), "load_one")
        return t.value, s.value

    def restart_io(self, io, datadir, time=0.0, istep=0):
        """The checkpoint of every block, <datadir>fld_b_NNN_{u,v,w,p}.bin
        (+ _s_MMM per scalar), main.f90:543-560 / 1090-1134; after 'r' po = p
        and (time, istep) are returned; call init_bounds next."""
        t, s = C.c_double(time), C.c_int(istep)
        _check(lib.snac_restart_io(self.h, io.encode(), str(datadir).encode(), C.byref(t), C.byref(s)),
               "restart_io")
        return t.value, s.value

    def write_visu_3d(self, datadir, fname_bin, fname_log, varname, b, fid, nmin, nmax, nskip, time, istep):
        """write_visu_3d (output.f90:289-303): out3d + write_log_output."""
        a = [np.ascontiguousarray(x, dtype=np.int32) for x in (nmin, nmax, nskip)]
        _check(lib.snac_write_visu_3d(self.h, str(datadir).encode(), fname_bin.encode(), fname_log.encode(),
                                      varname.encode(), b, FIELDS[fid], a[0], a[1], a[2], time, istep),
               "write_visu_3d")

    def save_grid(self, datadir):
        """grid_{x,y,z}_b_NNN.bin/.out and geometry_b_NNN.out of every block
        (initgrid.f90:128-143, main.f90:429-437)."""
        _check(lib.snac_save_grid(self.h, str(datadir).encode()), "save_grid")

    # ---- reference routines ----
    def rk_mom(self, irk, dt):
        _check(lib.snac_rk_mom(self.h, irk, dt), "rk_mom")

    def bounduvw(self, is_correc):
        _check(lib.snac_bounduvw(self.h, int(bool(is_correc))), "bounduvw")

    def inflow(self, is_correc):
        _check(lib.snac_inflow(self.h, int(bool(is_correc))), "inflow")

    def rk_scal(self, irk, dt):
        _check(lib.snac_rk_scal(self.h, irk, dt), "rk_scal")

    def scalar_forcing(self):
        f = np.zeros(max(self.nscal, 1))
        _check(lib.snac_scalar_forcing(self.h, f), "scalar_forcing")
        return f[:self.nscal]

    def boundp(self, which="p"):
        _check(lib.snac_boundp(self.h, FIELDS[which]), "boundp")

    def fillps(self, dtrk):
        _check(lib.snac_fillps(self.h, dtrk), "fillps")

    def solve_pressure(self):
        _check(lib.snac_solve_pressure(self.h), "solve_pressure")

    def correc(self, dtrk):
        _check(lib.snac_correc(self.h, dtrk), "correc")

    def updt_pressure(self):
        _check(lib.snac_updt_pressure(self.h), "updt_pressure")

    def set_impdiff(self, enable=True):
        """Implicit diffusion (the reference's _IMPDIFF build) on or off."""
        _check(lib.snac_set_impdiff(self.h, int(bool(enable))), "set_impdiff")

    def solve_impdiff(self, which, alpha):
        """solve_impdiff_field of 'u', 'v' or 'w' with alpha = -visc*dtrk/2."""
        _check(lib.snac_solve_impdiff(self.h, FIELDS[which], float(alpha)), "solve_impdiff")

    def impdiff_stats(self, which):
        """(CG iterations, largest relative residual over the modes) of the
        last MG-PCG Helmholtz solve of 'u', 'v', 'w' or a scalar; (0, 0.0)
        when the field was solved on the tensor-product plan."""
        it = C.c_int()
        rel = C.c_double()
        _check(lib.snac_impdiff_stats(self.h, FIELDS[which], C.byref(it), C.byref(rel)), "impdiff_stats")
        return it.value, rel.value

    def updt_pressure_impdiff(self, alpha):
        _check(lib.snac_updt_pressure_impdiff(self.h, float(alpha)), "updt_pressure_impdiff")

    def chkdiv(self):
        a, b = C.c_double(), C.c_double()
        _check(lib.snac_chkdiv(self.h, C.byref(a), C.byref(b)), "chkdiv")
        return a.value, b.value

    def chkdt(self):
        a = C.c_double()
        _check(lib.snac_chkdt(self.h, C.byref(a)), "chkdt")
        return a.value

    def init_bounds(self):
        _check(lib.snac_init_bounds(self.h), "init_bounds")

    def set_buoyancy(self, enable):
        """The reference's _BOUSSINESQ_BUOYANCY build as a switch (off by
        default): rk_mom adds the buoyancy of scalar 1 (rk.f90:104-108)."""
        _check(lib.snac_set_buoyancy(self.h, int(bool(enable))), "set_buoyancy")

    def initflow(self):
        """Initial fields of every block's block_inivel / iniscal
        (initflow.f90:10-262, no noise); ghosts and p zero."""
        _check(lib.snac_initflow(self.h), "initflow")

    def step(self, dt):
        _check(lib.snac_step(self.h, dt), "step")

    def steps(self, dt, n):
        _check(lib.snac_steps(self.h, dt, int(n)), "steps")

    def set_graph(self, enable):
        _check(lib.snac_set_graph(self.h, int(bool(enable))), "set_graph")

    def sync(self):
        _check(lib.snac_sync(self.h), "sync")

    def r2r_lines(self, kind, x, strided=False, tile=0):
        """Transforms every row of x (howmany x n); strided=True hands the
        lines to the device transposed (element stride = howmany).  tile:
        the kernel choice 0..7 the autotuner picks from (5..7 and 3, 4 for
        R2HC/HC2R only; the other kinds treat them as 0)."""
        x = np.atleast_2d(np.asarray(x, dtype=np.float64))
        h, n = x.shape
        src = np.ascontiguousarray(x.T if strided else x)
        out = np.zeros_like(src)
        _check(lib.snac_r2r_lines(self.h, R2R_KINDS[kind], n, h, int(bool(strided)), int(tile), src.reshape(-1),
                                  out.reshape(-1)), "r2r_lines")
        return out.T if strided else out

    def set_timing(self, enable):
        _check(lib.snac_set_timing(self.h, int(bool(enable))), "set_timing")

    def get_timing(self):
        """{phase: (total ms, number of timed intervals)} since the last call."""
        ms = np.zeros(len(PHASES))
        cnt = np.zeros(len(PHASES), np.int32)
        _check(lib.snac_get_timing(self.h, ms, cnt), "get_timing")
        return {p: (float(ms[q]), int(cnt[q])) for q, p in enumerate(PHASES)}

    SOLVERS = {"auto": 0, "tensor": 1, "mg": 2}

    def set_solver(self, name):
        """Poisson solver: 'auto' (tensor-product when the union is a box with
        uniform transform axes, else MG-PCG), 'tensor' or 'mg'."""
        _check(lib.snac_ctx_set_solver(self.h, self.SOLVERS[name]), "set_solver")

    PRECISIONS = {"fp64": 0, "mixed": 1, "learned": 2}

    def set_precision(self, mode):
        """'fp64' (default), 'mixed' or 'learned'.  mixed: the momentum
        right-hand side of rk_mom in FP32 with FP32 old tendencies (the RK3
        accumulation, pressure gradient and forcing stay FP64) and, on the
        MG-PCG path, the V-cycle in FP32 under a flexible CG whose residuals,
        operator, dot products and null-space removal stay FP64 (same
        stopping tolerance).  fillps, correc, the ghost fills and the direct
        tensor-product solver stay FP64.  learned: FP64 stencils, the trained
        network preconditioning the flexible CG of the mixed mode."""
        _check(lib.snac_set_precision(self.h, self.PRECISIONS[mode]), "set_precision")

    def learned_collect(self, npairs):
        """Record the next MG-PCG solves' (rhs, solution) pairs (npairs planes)."""
        _check(lib.snac_learned_collect(self.h, int(npairs)), "learned_collect")

    def learned_train(self, epochs, batch, lr, seed=1):
        hist = np.zeros(epochs, np.float32)
        _check(lib.snac_learned_train(self.h, int(epochs), int(batch), float(lr), int(seed), hist), "learned_train")
        return hist

    def learned_options(self, fallback=True, stall_ratio=0.5, initial_guess=False):
        _check(lib.snac_learned_options(self.h, int(bool(fallback)), float(stall_ratio), int(bool(initial_guess))),
               "learned_options")

    def learned_time(self, reps=10):
        """(ms of one network forward over every mode, its GFLOP)"""
        a, b = C.c_double(), C.c_double()
        _check(lib.snac_learned_time(self.h, int(reps), C.byref(a), C.byref(b)), "learned_time")
        return a.value, b.value

    def learned_stats(self):
        a, b = np.zeros(1, np.int32), np.zeros(1, np.int32)
        _check(lib.snac_learned_stats(self.h, a, b), "learned_stats")
        return int(a[0]), int(b[0])

    def solver_stats(self):
        """(CG iterations, largest relative residual over the modes) of the
        last MG-PCG pressure solve; (0, 0.0) after a tensor-product solve."""
        it = C.c_int()
        rel = C.c_double()
        _check(lib.snac_solver_stats(self.h, C.byref(it), C.byref(rel)), "solver_stats")
        return it.value, rel.value

    def solver_counters(self):
        """(solves, host-resumed solves, solves that stopped above the
        tolerance) of the MG-PCG path since the last call."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib.snac_solver_counters(self.h, C.byref(a), C.byref(b), C.byref(c)), "solver_counters")
        return a.value, b.value, c.value

    def set_graph_cg_iters(self, n):
        """CG iterations of each solve captured in the step graph (0: measured)."""
        _check(lib.snac_set_graph_cg_iters(self.h, int(n)), "set_graph_cg_iters")

    def stream_copy(self, n, reps=10):
        """GB/s of a plain device copy of n doubles (HBM reference)."""
        g = C.c_double()
        _check(lib.snac_stream_copy(self.h, int(n), int(reps), C.byref(g)), "stream_copy")
        return g.value

    def autotune(self):
        """(choices, best ms) of the autotuner: transform tile per call site
        (forward, then inverse), then the tridiagonal variant (both on the
        tensor-product solver path only), then the rk_mom variant, then the
        FP32 rk_mom variant of the mixed-precision modes."""
        ch = np.zeros(16, np.int32)
        ms = np.zeros(16)
        _check(lib.snac_get_autotune(self.h, ch, ms, 16), "get_autotune")
        k = int(np.argmax(ch == -2)) if (ch == -2).any() else 16
        return ch[:k].tolist(), [round(float(x), 4) for x in ms[:k] if x > 0]

    def mg_line_choices(self):
        """Per level of the FP64 and of the FP32 V-cycle, the line relaxation
        kernel it runs (set_mg_line_variant's numbering; -1: a level of the
        fused coarse end, or the MG-PCG path not planned yet)."""
        a = np.zeros(32, np.int32)
        b = np.zeros(32, np.int32)
        _check(lib.snac_mg_line_choices(self.h, a, b, 32), "mg_line_choices")
        trim = lambda v: [int(x) for x in v[: max([i + 1 for i, x in enumerate(v) if x >= 0] or [0])]]
        return trim(a), trim(b)

    def set_rk_variant(self, v):
        """Overrides the autotuned rk_mom kernel variant (0..16)."""
        _check(lib.snac_set_kernel_variant(self.h, 0, int(v)), "set_kernel_variant")

    def set_rk32_variant(self, v):
        """Overrides the autotuned FP32 rk_mom kernel variant (7..16) of the
        mixed-precision modes."""
        _check(lib.snac_set_kernel_variant(self.h, 3, int(v)), "set_kernel_variant")

    def set_mg_line_variant(self, v):
        """MG-PCG line relaxation kernel, process-wide: 0 rows read per lane,
        1 right-hand sides staged through LDS, 2 as 1 with the lines of a
        workgroup over several modes sharing LDS-staged coefficients (FFT
        path), 3 and 4 as 2 with 16 and 8 lanes per line (longer row chunks;
        the same systems in another elimination order); a negative value
        (the default) restores each level's fastest, timed at set-up.  0..2
        compute the same values."""
        _check(lib.snac_set_kernel_variant(self.h, 4, int(v)), "set_kernel_variant")

    def set_mg_tail_lds(self, on):
        """Process-wide: the fused coarse end of the V-cycle with one mode's
        vectors of every coarse level in LDS where they fit (the default),
        or in global memory; both compute the same values."""
        _check(lib.snac_set_kernel_variant(self.h, 6, int(bool(on))), "set_kernel_variant")

    def set_mg_cg_lds(self, on):
        """Process-wide: the CG direction pass with a workgroup's new
        directions staged in LDS where the level allows it (the default), or
        formed again for every neighbour; both compute the same values."""
        _check(lib.snac_set_kernel_variant(self.h, 8, int(bool(on))), "set_kernel_variant")

    def set_mg_warm_start(self, on):
        """This context: start each MG-PCG pressure solve from the previous
        solve's solution (as the reference's solves start from po) or from
        zero (the default); both stop at the same relative residual."""
        _check(lib.snac_set_kernel_variant(self.h, 7, int(bool(on))), "set_kernel_variant")

    def set_mg_fast_rcp(self, on):
        """Process-wide: the mode-batched MG line sweeps divide by refined
        hardware reciprocals (the default) or by IEEE quotients."""
        _check(lib.snac_set_kernel_variant(self.h, 5, int(bool(on))), "set_kernel_variant")

    def set_tridiag_variant(self, v):
        """Overrides the autotuned tridiagonal kernel variant (0..3)."""
        _check(lib.snac_set_kernel_variant(self.h, 1, int(v)), "set_kernel_variant")

    def set_ghost_variant(self, v):
        """Ghost fills: 0 one launch per phase, 1 (default) composed into one
        launch per fill where the fill allows it."""
        _check(lib.snac_set_kernel_variant(self.h, 2, int(v)), "set_kernel_variant")

    def new_dt(self):
        """main.f90:660-661: dt = dt_f if > 0 else min(cfl*dt_cfl, dtmax)."""
        if self.dt_f > 0:
            return self.dt_f
        return min(self.cfl * self.chkdt(), self.dtmax)


class Net:
    """The learned preconditioner's network on its own (snac_net_*)."""

    @staticmethod
    def conv_variant(v):
        """0: per-tap reads from memory, 1: input staged in LDS (default)."""
        _check(lib.snac_net_conv_variant(int(v)), "net_conv_variant")

    def __init__(self, H, W, maxbatch=8, seed=1, device=0):
        h = C.c_void_p()
        _check(lib.snac_net_create(C.byref(h), int(H), int(W), int(maxbatch), int(seed), int(device)), "net_create")
        self.h, self.H, self.W = h, H, W

    def close(self):
        if getattr(self, "h", None):
            lib.snac_net_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def params(self):
        p = np.zeros(lib.snac_net_num_params(self.h), np.float32)
        _check(lib.snac_net_get_params(self.h, p), "net_get_params")
        return p

    def set_params(self, p):
        _check(lib.snac_net_set_params(self.h, np.ascontiguousarray(p, dtype=np.float32)), "net_set_params")

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(x.shape[:3], np.float32)
        _check(lib.snac_net_forward(self.h, x.shape[0], x, y), "net_forward")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        g = np.zeros(lib.snac_net_num_params(self.h), np.float32)
        loss = C.c_float()
        _check(lib.snac_net_loss_grad(self.h, x.shape[0], x, np.ascontiguousarray(t, dtype=np.float32),
                                      C.byref(loss), g), "net_loss_grad")
        return loss.value, g

    def train(self, x, t, batch, epochs, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        hist = np.zeros(epochs, np.float32)
        _check(lib.snac_net_train(self.h, x.shape[0], x, np.ascontiguousarray(t, dtype=np.float32), int(batch),
                                  int(epochs), float(lr), hist), "net_train")
        return hist

"""ctypes front-end of the CPU oracle (oracle/_build/libsnac_oracle.so).

Test infrastructure: the oracle is the checker, never the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "libsnac_oracle.so")

FIELDS = {"u": 0, "v": 1, "w": 2, "p": 3, "pp": 4, "po": 5, "dudtrko": 6, "dvdtrko": 7, "dwdtrko": 8}
for _i in range(1, 9):  # scalars s1.., their dsdtrko1..
    FIELDS[f"s{_i}"] = 9 + 2 * (_i - 1)
    FIELDS[f"dsdtrko{_i}"] = 10 + 2 * (_i - 1)


def ghosted(fid):
    f = FIELDS[fid]
    return f < 6 or (f >= 9 and (f - 9) % 2 == 0)
SOLVERS = {"band": 0, "cg": 1, "tensor": 2, "auto": 3}
R2R = {"R2HC": 0, "HC2R": 1, "REDFT00": 3, "REDFT01": 4, "REDFT10": 5, "REDFT11": 6,
       "RODFT00": 7, "RODFT01": 8, "RODFT10": 9, "RODFT11": 10}
RKCOEFF = [(32.0 / 60.0, 0.0), (25.0 / 60.0, -17.0 / 60.0), (45.0 / 60.0, -25.0 / 60.0)]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
        L = C.CDLL(LIB)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.oc_create.restype = C.c_void_p
        L.oc_create.argtypes = [C.c_int, ip, dp, dp, ip, dp, C.c_char_p, dp, C.c_char_p, dp, ip,
                                C.c_int, C.c_double, dp, C.c_char_p, C.c_int]
        L.oc_destroy.argtypes = [C.c_void_p]
        L.oc_set_solver.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        L.oc_set_buoyancy.argtypes = [C.c_void_p, C.c_int]
        L.oc_set_threads.argtypes = [C.c_int]
        L.oc_nblocks.argtypes = [C.c_void_p]
        L.oc_block_info.argtypes = [C.c_void_p, C.c_int, ip, ip, ip, ip]
        L.oc_vol_all.restype = C.c_double
        L.oc_vol_all.argtypes = [C.c_void_p]
        L.oc_get_metric.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.oc_field_size.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.oc_get_field.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.oc_set_field.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.oc_rk_mom.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.oc_bounduvw.argtypes = [C.c_void_p, C.c_int]
        L.oc_boundp.argtypes = [C.c_void_p, C.c_int]
        L.oc_inflow.argtypes = [C.c_void_p, C.c_int]
        L.oc_fillps.argtypes = [C.c_void_p, C.c_double]
        L.oc_solve_pressure.argtypes = [C.c_void_p]
        L.oc_correc.argtypes = [C.c_void_p, C.c_double]
        L.oc_updt_pressure.argtypes = [C.c_void_p]
        L.oc_chkdiv.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.oc_chkdt.restype = C.c_double
        L.oc_chkdt.argtypes = [C.c_void_p]
        L.oc_step.argtypes = [C.c_void_p, C.c_double]
        L.oc_set_scalars.argtypes = [C.c_void_p, C.c_int, dp, C.c_double, dp, dp, ip, dp, C.c_char_p, dp]
        L.oc_rk_scal.argtypes = [C.c_void_p, C.c_int, C.c_double, dp]
        L.oc_init_bounds.argtypes = [C.c_void_p]
        L.oc_r2r_direct.argtypes = [C.c_int, C.c_int, dp, dp]
        L.oc_r2r.argtypes = [C.c_int, C.c_int, dp, dp]
        L.oc_find_fft.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), dp]
        L.oc_eigenvalues.argtypes = [C.c_int, C.c_char_p, C.c_int, dp]
        L.oc_initgrid.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, dp, dp, dp, dp]
        L.oc_num_unknowns.argtypes = [C.c_void_p]
        L.oc_apply_operator.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.oc_solve_operator.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.oc_get_lambda.argtypes = [C.c_void_p, dp, dp]
        L.oc_set_impdiff.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int]
        L.oc_solve_impdiff.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.oc_updt_pressure_impdiff.argtypes = [C.c_void_p, C.c_double]
        _lib = L
    return _lib


def _chars(a):
    return "".join(np.asarray(a).reshape(-1).tolist()).encode()


class OracleCase:
    """A multi-block case held by the oracle; API mirrors the product's."""

    def __init__(self, case):
        L = lib()
        self.case = case
        err = C.create_string_buffer(256)
        nb = case["nblocks"]
        self.h = L.oc_create(
            nb, np.ascontiguousarray(case["ng"], dtype=np.int32).reshape(-1),
            np.ascontiguousarray(case["lmin"], dtype=float).reshape(-1),
            np.ascontiguousarray(case["lmax"], dtype=float).reshape(-1),
            np.ascontiguousarray(case["gt"], dtype=np.int32).reshape(-1),
            np.ascontiguousarray(case["gr"], dtype=float).reshape(-1),
            _chars(case["cbcvel"]), np.ascontiguousarray(case["bcvel"], dtype=float).reshape(-1),
            _chars(case["cbcpre"]), np.ascontiguousarray(case["bcpre"], dtype=float).reshape(-1),
            np.ascontiguousarray(case["inflow_type"], dtype=np.int32).reshape(-1),
            int(case["fft_axis"]), float(case["visci"]),
            np.ascontiguousarray(case["bforce"], dtype=float), err, 256)
        if not self.h:
            raise ValueError("oracle rejected the case: " + err.value.decode())
        self.nblocks = nb
        self.nscal = int(case.get("nscal", 0))
        if self.nscal > 0:
            L.oc_set_scalars(self.h, self.nscal, np.ascontiguousarray(case["alphai"], dtype=float),
                             float(case["beta"]), np.ascontiguousarray(case["gacc"], dtype=float),
                             np.ascontiguousarray(case["ssource"], dtype=float),
                             np.ascontiguousarray(case["is_sforced"], dtype=np.int32),
                             np.ascontiguousarray(case["scalf"], dtype=float), _chars(case["cbcscal"]),
                             np.ascontiguousarray(case["bcscal"], dtype=float).reshape(-1))

        if case.get("impdiff"):
            self.set_impdiff(True)

    def set_impdiff(self, on=True):
        """Implicit diffusion (the reference's _IMPDIFF build)."""
        err = C.create_string_buffer(256)
        if lib().oc_set_impdiff(self.h, int(on), err, 256):
            raise ValueError("oracle rejected implicit diffusion: " + err.value.decode())

    def solve_impdiff(self, which, alpha):
        """solve_impdiff_field of u/v/w or scalar s<i> with alpha = -nu*dtrk/2."""
        q = {"u": 1, "v": 2, "w": 3}.get(which)
        if q is None:
            q = 3 + int(which[1:])
        rc = lib().oc_solve_impdiff(self.h, q, float(alpha))
        if rc:
            raise RuntimeError(f"oracle Helmholtz solve failed ({rc})")

    def updt_pressure_impdiff(self, alpha):
        lib().oc_updt_pressure_impdiff(self.h, float(alpha))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_destroy(self.h)
            self.h = None

    def set_solver(self, method="auto", tol=1e-13, maxiter=20000):
        lib().oc_set_solver(self.h, SOLVERS[method], tol, maxiter)

    def set_buoyancy(self, enable):
        """The _BOUSSINESQ_BUOYANCY build (rk.f90:104-108); off by default."""
        if lib().oc_set_buoyancy(self.h, int(enable)) != 0:
            raise RuntimeError("_BOUSSINESQ_BUOYANCY requires nscal > 0.")

    def block_info(self, b):
        n = np.zeros(3, np.int32); lo = np.zeros(3, np.int32); nbr = np.zeros(6, np.int32); per = np.zeros(3, np.int32)
        lib().oc_block_info(self.h, b, n, lo, nbr, per)
        return dict(n=n, lo_g=lo, nb=nbr.reshape(3, 2), periods=per)

    def metric(self, b, which):
        n = self.block_info(b)["n"][(which % 6) // 2]
        out = np.zeros(n + 2)
        lib().oc_get_metric(self.h, b, which, out)
        return out

    def shape(self, b, fid):
        n = self.block_info(b)["n"]
        if ghosted(fid):
            return (n[0] + 2, n[1] + 2, n[2] + 2)
        return tuple(n)

    def get(self, b, fid):
        shp = self.shape(b, fid)
        out = np.zeros(int(np.prod(shp)))
        lib().oc_get_field(self.h, b, FIELDS[fid], out)
        return out.reshape(shp, order="F")

    def set(self, b, fid, arr):
        shp = self.shape(b, fid)
        a = np.asarray(arr, dtype=float)
        assert a.shape == shp, (a.shape, shp)
        lib().oc_set_field(self.h, b, FIELDS[fid], np.ascontiguousarray(a.reshape(-1, order="F")))

    def rk_mom(self, irk, dt):
        lib().oc_rk_mom(self.h, irk, dt)

    def rk_scal(self, irk, dt):
        """rk_scal + bulk forcing of every scalar; returns the forcings f."""
        f = np.zeros(max(self.nscal, 1))
        lib().oc_rk_scal(self.h, irk, dt, f)
        return f[:self.nscal]

    def bounduvw(self, is_correc):
        lib().oc_bounduvw(self.h, int(is_correc))

    def boundp(self, which="p"):
        lib().oc_boundp(self.h, FIELDS[which])

    def inflow(self, is_correc):
        lib().oc_inflow(self.h, int(is_correc))

    def fillps(self, dtrk):
        lib().oc_fillps(self.h, dtrk)

    def solve_pressure(self):
        rc = lib().oc_solve_pressure(self.h)
        if rc:
            raise RuntimeError(f"oracle pressure solve failed ({rc})")

    def correc(self, dtrk):
        lib().oc_correc(self.h, dtrk)

    def updt_pressure(self):
        lib().oc_updt_pressure(self.h)

    def chkdiv(self):
        a, b = C.c_double(), C.c_double()
        lib().oc_chkdiv(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def chkdt(self):
        return lib().oc_chkdt(self.h)

    def init_bounds(self):
        lib().oc_init_bounds(self.h)

    def step(self, dt):
        rc = lib().oc_step(self.h, dt)
        if rc:
            raise RuntimeError(f"oracle step failed ({rc})")

    def nunk(self):
        return lib().oc_num_unknowns(self.h)

    def apply_operator(self, mode, x):
        y = np.zeros_like(x)
        lib().oc_apply_operator(self.h, mode, np.ascontiguousarray(x), y)
        return y

    def solve_operator(self, mode, b, x0=None):
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=float)
        rc = lib().oc_solve_operator(self.h, mode, np.ascontiguousarray(b), x)
        if rc:
            raise RuntimeError(f"oracle solve failed ({rc})")
        return x

    def lambdas(self):
        lam = np.zeros(4096)
        nf = np.zeros(1)
        n = lib().oc_get_lambda(self.h, lam, nf)
        return lam[:n].copy(), float(nf[0])


def r2r(kind, x, direct=False):
    x = np.ascontiguousarray(x, dtype=float)
    y = np.zeros_like(x)
    f = lib().oc_r2r_direct if direct else lib().oc_r2r
    rc = f(R2R[kind], len(x), x, y)
    assert rc == 0
    return y


def eigenvalues(n, bc, centered=True):
    lam = np.zeros(n)
    rc = lib().oc_eigenvalues(n, bc.encode(), int(centered), lam)
    assert rc == 0
    return lam


def initgrid(n, gt, gr, lmin, lmax):
    a = [np.zeros(n + 2) for _ in range(4)]
    lib().oc_initgrid(n, gt, gr, lmin, lmax, *a)
    return dict(drc=a[0], drf=a[1], rc=a[2], rf=a[3])


def set_threads(n):
    return lib().oc_set_threads(n)

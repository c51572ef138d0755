"""snac_chkdiv_noalloc: chkdiv over caller device arrays with no allocation,
checked against snac_chkdiv of a one-block context on the same fields
(chkdiv.f90:8-44).  Both run the same kernels and the same summation order,
so divtot and divmax agree bit for bit.  Two calls in a row, the second on a
smaller block, share the library's one scratch array; a block that needs
more partials than the scratch holds is refused before any launch.

Device memory for the fields is allocated here through the same HIP runtime
the library uses (the test's own buffers; the entry point allocates none).
"""
import ctypes as C

import numpy as np
import pytest

import cases
from test_gpu_parity import make_pair

pytestmark = pytest.mark.gpu


def hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevArray:
    """A device copy of a float64 array (Fortran order), freed on close."""

    def __init__(self, h, a):
        self.h = h
        self.n = a.size * 8
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), self.n) == 0
        buf = np.asfortranarray(a, dtype=np.float64)
        assert h.hipMemcpy(self.p, buf.ctypes.data, self.n, H2D) == 0

    def get(self, shape):
        out = np.empty(shape, np.float64)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, self.n, D2H) == 0
        return out

    def close(self):
        if self.p:
            self.h.hipFree(self.p)
            self.p = C.c_void_p()


def lib():
    import snac_amd
    L = snac_amd.lib
    L.snac_chkdiv_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_long, C.c_long] + \
        [C.c_void_p] * 6 + [C.c_double, C.c_void_p, C.c_void_p]
    L.snac_chkdiv_noalloc.restype = C.c_int
    return L


def noalloc_vs_ctx(tmp_path, n, seed):
    ctx, _ = make_pair(tmp_path, [cases.channel_block(n)], 3, bforce=(12 / 5600, 0, 0))
    rng = np.random.default_rng(seed)
    fields = []
    for f in ("u", "v", "w"):
        a = rng.uniform(-1, 1, ctx.shape(0, f))
        ctx.set(0, f, a)
        fields.append(a)
    ref = ctx.chkdiv()
    h = hip()
    dev = [DevArray(h, a) for a in fields]
    met = [DevArray(h, ctx.metric(0, q)) for q in (1, 3, 5)]  # dxf, dyf, dzf
    out = DevArray(h, np.zeros(2))
    try:
        nx, ny, nz = n
        rc = lib().snac_chkdiv_noalloc(0, nx, ny, nz, nx + 2, ny + 2, *(d.p for d in dev), *(m.p for m in met),
                                       ctx.vol_all, out.p, None)
        assert rc == 0, lib().snac_last_error()
        assert h.hipDeviceSynchronize() == 0
        got = out.get(2)
    finally:
        for d in dev + met + [out]:
            d.close()
        ctx.close()
    return ref, got


def test_noalloc_chkdiv_matches_context(tmp_path):
    ref1, got1 = noalloc_vs_ctx(tmp_path / "a", (96, 40, 24), seed=5)
    print(f"call 1: ctx {ref1} noalloc {tuple(got1)}", flush=True)
    assert got1[0] == ref1[0] and got1[1] == ref1[1]
    ref2, got2 = noalloc_vs_ctx(tmp_path / "b", (40, 12, 8), seed=6)  # smaller: fewer partials than call 1 left
    print(f"call 2: ctx {ref2} noalloc {tuple(got2)}", flush=True)
    assert got2[0] == ref2[0] and got2[1] == ref2[1]


def test_noalloc_chkdiv_refuses_oversized_block():
    L = lib()
    # ceil(4096/64) * ceil(4096/4) * 8 partials: far beyond the scratch; refused before any launch
    dummy = C.c_void_p(16)
    rc = L.snac_chkdiv_noalloc(0, 4096, 4096, 8, 4098, 4098, dummy, dummy, dummy, dummy, dummy, dummy, 1.0, dummy,
                               None)
    assert rc != 0
    assert b"partial sums" in L.snac_last_error()

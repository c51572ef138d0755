"""The Kovasznay inflow (inflow_type 2, initflow.f90:407-416) of the
oracle against the closed-form Kovasznay flow (L. I. G. Kovasznay, 1948):
u = 1 - e^(lambda x) cos(2 pi z), w = lambda/(2 pi) e^(lambda x) sin(2 pi z),
lambda = Re/2 - sqrt(Re^2/4 + 4 pi^2), taken at the inflow face x = -0.5 of
the reference's own example case (examples/__MULTI_BLOCK_GEOMETRY/
kovasznay_flow, fixture in tests/golden/multiblock/kovasznay_flow): the
normal component is the face value u(lo-1) at z cell centres, the second
tangential one enters the ghost w(lo-1) = 2 w_in - w(lo) at z faces.
"""
import os

import numpy as np

import nmlparse
import oracle_py

KOV = os.path.join(os.path.dirname(__file__), "golden", "multiblock", "kovasznay_flow")


def test_kovasznay_inflow_plane_closed_form():
    case = nmlparse.load_case(os.path.join(KOV, "dns.nml"), os.path.join(KOV, "blocks.nml"), 3)
    oc = oracle_py.OracleCase(case)
    oc.inflow(False)  # fields start at zero: the ghosts take the inflow values
    u, v, w = (oc.get(0, f) for f in ("u", "v", "w"))
    nz = 64
    zf = -0.5 + np.arange(0, nz + 2) * (1.0 / nz)           # faces zf(lo-1 .. hi+1), zf(lo-1) = lmin
    zc = zf - 0.5 / nz                                     # centres
    re = 1.0 * 1.0 / (1.0 / 0.025)                         # uref * lref / visc, visci = 0.025
    lam = re / 2 - np.sqrt(re ** 2 / 4 + 4 * np.pi ** 2)
    un = 1.0 - np.exp(lam * -0.5) * np.cos(2 * np.pi * zc)
    wt = lam / (2 * np.pi) * np.exp(lam * -0.5) * np.sin(2 * np.pi * zf)
    np.testing.assert_allclose(u[0], np.broadcast_to(un[None, :], u[0].shape), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(w[0], np.broadcast_to(2 * wt[None, :], w[0].shape), rtol=1e-14, atol=1e-15)
    assert not v[0].any()
    assert not u[1:].any() and not w[1:].any()

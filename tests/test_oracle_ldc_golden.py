"""The oracle against the reference's own known-answer test.

tests/lid_driven_cavity of the reference (test.py:5-16, testit.sh:28-57): four
2x32x32 blocks, Re=1000, 1500 steps, v-profile at the y-index 32 face of the
stitched (y,z) plane, compared with data_ldc_re1000.txt at rtol=1e-7,
atol=5e-8, for both FFT_AXIS=0 (3D Poisson) and FFT_AXIS=1.  The fixture files
under tests/golden/ldc are copies of the reference's data files.
"""
import os

import numpy as np
import pytest

import nmlparse
import oracle_py

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ldc")


def run_ldc(fft_axis, nsteps=None):
    case = nmlparse.load_case(os.path.join(GOLD, "dns.nml"), os.path.join(GOLD, "blocks.nml"), fft_axis)
    oc = oracle_py.OracleCase(case)
    oc.init_bounds()

    def new_dt():
        if case["dt_f"] > 0:
            return case["dt_f"]
        return min(case["cfl"] * oc.chkdt(), case["dtmax"])

    dt = new_dt()
    for istep in range(1, (nsteps or case["nstep"]) + 1):
        oc.step(dt)
        if istep % case["icheck"] == 0:
            dt = new_dt()
            divtot, divmax = oc.chkdiv()
            assert divtot == divtot and divmax < 1e-10
    return oc


def ldc_profile(get_v):
    v = [get_v(b)[1, 1:-1, 1:-1] for b in range(4)]
    data = np.block([[v[0], v[2]], [v[1], v[3]]])
    return data[data.shape[0] // 2, :]


@pytest.mark.slow
@pytest.mark.parametrize("fft_axis", [1, 0])
def test_oracle_ldc_golden(fft_axis):
    oracle_py.set_threads(min(8, os.cpu_count() or 1))
    oc = run_ldc(fft_axis)
    prof = ldc_profile(lambda b: oc.get(b, "v"))
    ref = np.loadtxt(os.path.join(GOLD, "data_ldc_re1000.txt"))
    np.testing.assert_allclose(prof, ref[:, 1], rtol=1e-7, atol=5e-8)

"""The Fortran binding (include/snac_amd.f90), compiled with amdflang by
build() together with two Fortran hosts (tests/fortran/): the drop-in the
reference's Fortran main program would use, in the idiom of its fftw.f90
bind(C) interfaces.

* host-only (no GPU): ldc_info creates a context of the reference's
  lid-driven-cavity case without a device and prints every block's sizes,
  placement, neighbours and periods, which must equal what the Python
  binding reads from the same context; a missing namelist file must come
  back as a failing status with the library's message.
* GPU: ldc_steps runs the reference's time loop (main.f90:893-1069) with one
  C-ABI call per reference routine for the 1500 steps of the reference's
  known-answer test (tests/lid_driven_cavity: test.py:5-16, testit.sh:28-57);
  its v profile must match data_ldc_re1000.txt at rtol 1e-7, atol 5e-8.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FBUILD = os.path.join(ROOT, "snac_amd", "_fbuild")
GOLD = os.path.join(ROOT, "tests", "golden", "ldc")


def program(name):
    p = os.path.join(FBUILD, name)
    if not os.path.exists(p):
        pytest.fail(f"{p} is missing: build() compiles the Fortran binding and its hosts with amdflang")
    return p


def ldc_dir(tmp_path):
    for f in ("dns.nml", "blocks.nml"):
        shutil.copy(os.path.join(GOLD, f), tmp_path / f)
    return tmp_path


def test_fortran_host_only_block_info(tmp_path):
    import snac_amd
    d = ldc_dir(tmp_path)
    r = subprocess.run([program("ldc_info")], cwd=d, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    ctx = snac_amd.Context(str(d / "dns.nml"), str(d / "blocks.nml"), 1, device=-1)
    assert lines[0].split() == ["blocks", str(ctx.nblocks), "nstep", str(ctx.nstep), "icheck", str(ctx.icheck)]
    for b in range(ctx.nblocks):
        info = ctx.block_info(b)
        want = [b] + list(info["n"]) + list(info["lo_g"]) + list(info["nb"].ravel()) + list(info["periods"])
        assert [int(x) for x in lines[1 + b].split()[1:]] == want
    ctx.close()
    assert "missing file: status -1 message: cannot open no_such_dns.nml" in r.stdout


@pytest.mark.gpu
def test_fortran_host_ldc_reference_procedure(tmp_path):
    d = ldc_dir(tmp_path)
    print("fortran host: ldc, FFT_AXIS=1, 1500 steps", flush=True)
    r = subprocess.run([program("ldc_steps"), "1"], cwd=d, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    print(r.stdout.strip(), flush=True)
    v = []
    for b in range(1, 5):
        a = np.fromfile(d / f"v_b_{b:03d}.bin", dtype="float64").reshape((4, 34, 34), order="F")
        v.append(a[1, 1:-1, 1:-1])
    full = np.block([[v[0], v[2]], [v[1], v[3]]])
    ref = np.loadtxt(os.path.join(GOLD, "data_ldc_re1000.txt"))
    np.testing.assert_allclose(full[full.shape[0] // 2, :], ref[:, 1], rtol=1e-7, atol=5e-8)

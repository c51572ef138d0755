"""Checkpoint, visualisation and grid files in the reference's formats
(load.f90, output.f90, initgrid.f90 save_grid/load_grid, main.f90 restart),
checked against oracle/fileformats.py.  Grid files work on host-only
contexts (CPU); field files go through the device (gpu)."""
import os

import numpy as np
import pytest

import cases
import oracle_py
from oracle import fileformats as ff


def test_fortran_e_worked_examples():
    """Ew.d of the Fortran standard: 0.d1..dd E+ee, rounded to nearest."""
    assert ff.fortran_e(0.5) == "  0.5000000E+00"
    assert ff.fortran_e(1.0) == "  0.1000000E+01"
    assert ff.fortran_e(1234.5678) == "  0.1234568E+04"
    assert ff.fortran_e(-0.000123456789) == " -0.1234568E-03"
    assert ff.fortran_e(9.99999999) == "  0.1000000E+02"  # carry into the exponent
    assert ff.fortran_e(0.0) == "  0.0000000E+00"
    assert ff.fortran_e(6.02e123) == "  0.6020000+124"  # three-digit exponent drops the E
    assert ff.visu_log_line("vex.bin", "Velocity_X", [1, 1, 1], [8, 9, 10], [1, 1, 1], 0.25, 42) == \
        "vex.bin Velocity_X    1    1    1    8    9   10    1    1    1  0.2500000E+00        42"


def _host_ctx(tmp_path, blocks):
    import snac_amd
    d, b = cases.write_case(str(tmp_path), blocks, 1000.0)
    return snac_amd.Context(d, b, 3, device=-1), blocks


def test_save_grid_matches_initgrid(tmp_path):
    """grid_{x,y,z}_b_NNN.bin/.out of every block hold initgrid's arrays
    (before bound_grid) in save_grid's layout; geometry_b_NNN.out holds
    lo_g, hi_g, lmin, lmax."""
    ctx, blocks = _host_ctx(tmp_path / "case", cases.bfs_blocks((12, 8), (20, 8, 12), 8))
    out = tmp_path / "data"
    out.mkdir()
    ctx.save_grid(str(out) + "/")
    for b, B in enumerate(blocks):
        info = ctx.block_info(b)
        for q, ax in enumerate("xyz"):
            g = oracle_py.initgrid(B["ng"][q], B["gt"][q], B["gr"][q], B["lmin"][q], B["lmax"][q])
            base = out / f"grid_{ax}_b_{b + 1:03d}"
            assert (base.with_suffix(".bin")).read_bytes() == ff.grid_bin_bytes(g["rf"], g["rc"], g["drf"], g["drc"])
            lines = (base.with_suffix(".out")).read_text().splitlines()
            assert lines == ff.grid_out_lines(g["rf"], g["rc"], g["drf"], g["drc"])
        geo = np.loadtxt(out / f"geometry_b_{b + 1:03d}.out")
        np.testing.assert_array_equal(geo[0], info["lo_g"])
        np.testing.assert_array_equal(geo[1], info["lo_g"] + np.array(B["ng"]) - 1)
        np.testing.assert_array_equal(geo[2], B["lmin"])
        np.testing.assert_array_equal(geo[3], B["lmax"])
    ctx.close()


def test_load_grid_replaces_generated_grid(tmp_path, monkeypatch):
    """A grid/grid_y_b_001.bin in the working directory replaces the
    generated y grid: interior values from the file, ghosts rebuilt as
    load_grid does, then bound_grid's cell-centre spacing at the ends."""
    ny = 10
    rf = np.zeros(ny + 2); rc = np.zeros(ny + 2); drf = np.zeros(ny + 2); drc = np.zeros(ny + 2)
    faces = np.concatenate([[0.0], np.cumsum(np.linspace(1.0, 2.0, ny))])
    faces /= faces[-1]
    rf[1:-1] = faces[1:]
    drf[1:-1] = np.diff(faces)
    rc[1:-1] = faces[1:] - 0.5 * drf[1:-1]
    drc[1:-2] = rc[2:-1] - rc[1:-2]
    drc[-2] = drf[-2]
    wd = tmp_path / "run"
    (wd / "grid").mkdir(parents=True)
    (wd / "grid" / "grid_y_b_001.bin").write_bytes(ff.grid_bin_bytes(rf, rc, drf, drc))
    monkeypatch.chdir(wd)
    ctx, _ = _host_ctx(tmp_path / "case", [cases.channel_block((8, ny, 6), gr_y=0.0)])
    erf, erc, edrf, edrc = ff.load_grid_ghosts(rf, rc, drf, drc)
    # bound_grid (initgrid.f90:91-127) on a block without neighbours
    edrc[0] = (edrf[1] + edrf[0]) / 2
    edrc[ny] = (edrf[ny + 1] + edrf[ny]) / 2
    edrc[ny + 1] = edrc[ny]
    np.testing.assert_array_equal(ctx.metric(0, 9), erf)   # yf
    np.testing.assert_array_equal(ctx.metric(0, 8), erc)   # yc
    np.testing.assert_array_equal(ctx.metric(0, 3), edrf)  # dyf
    np.testing.assert_array_equal(ctx.metric(0, 2), edrc)  # dyc
    ctx.close()


def test_save_then_load_grid_is_identity(tmp_path, monkeypatch):
    """Grid files written by save_grid, read back from grid/, give the same
    metrics as the generated grid."""
    blocks = [cases.channel_block((8, 12, 6))]
    ref, _ = _host_ctx(tmp_path / "a", blocks)
    wd = tmp_path / "run"
    (wd / "grid").mkdir(parents=True)
    ref.save_grid(str(wd / "grid") + "/")
    monkeypatch.chdir(wd)
    ctx, _ = _host_ctx(tmp_path / "b", blocks)
    for which in range(12):
        np.testing.assert_allclose(ctx.metric(0, which), ref.metric(0, which), rtol=0, atol=1e-15)
    ref.close()
    ctx.close()


# ------------------------------------------------------------------ fields
def _gpu_channel(tmp_path, n=(16, 12, 8), nscal=0):
    import snac_amd
    kw = dict(nscal=nscal, scalar=dict(alphai=[700.0], beta=0.0, ssource=[0.0], is_sforced=[False],
                                       scalf=[0.0])) if nscal else {}
    blk = cases.channel_block(n)
    if nscal:
        blk["cbcscal"] = [[["F", "F"], ["D", "D"], ["F", "F"]]]
        blk["bcscal"] = [[[0.0, 0.0], [0.0, 1.0], [0.0, 0.0]]]
    d, b = cases.write_case(str(tmp_path), [blk], 5600.0, cfl=0.5, bforce=(12 / 5600, 0, 0), **kw)
    print(f"case channel {n} nscal={nscal}", flush=True)
    return snac_amd.Context(d, b, 3)


@pytest.mark.gpu
def test_load_one_write_read(tmp_path):
    ctx = _gpu_channel(tmp_path / "c")
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, ctx.shape(0, "u"))
    ctx.set(0, "u", u)
    path = tmp_path / "fld_b_001_u.bin"
    ctx.load_one("w", path, 0, "u", time=1.25, istep=77)
    assert path.read_bytes() == ff.field_file_bytes(u[1:-1, 1:-1, 1:-1], 1.25, 77)
    ctx.set(0, "p", rng.uniform(-1, 1, ctx.shape(0, "p")))
    t, s = ctx.load_one("r", path, 0, "p")
    assert (t, s) == (1.25, 77)
    p = ctx.get(0, "p")
    np.testing.assert_array_equal(p[1:-1, 1:-1, 1:-1], u[1:-1, 1:-1, 1:-1])
    assert not p[0].any() and not p[-1].any()  # ghosts zeroed, as before a checkpoint load
    bad = tmp_path / "short.bin"
    bad.write_bytes(path.read_bytes()[:-8])
    with pytest.raises(RuntimeError, match="incorrect size"):
        ctx.load_one("r", bad, 0, "p")
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("nscal", [0, 1])
def test_restart_continues_bitwise(tmp_path, nscal):
    """3 steps, checkpoint, 3 more steps == a fresh context that loads the
    checkpoint, fills the ghosts and runs the last 3 steps (the RK3 history
    arrays are not in the checkpoint: the first substage does not read them)."""
    a = _gpu_channel(tmp_path / "a", nscal=nscal)
    u, v, w = cases.channel_ic(a.shape(0, "u"), a.metric(0, 8))
    for f, x in (("u", u), ("v", v), ("w", w)):
        a.set(0, f, x)
    if nscal:
        a.set(0, "s1", np.random.default_rng(2).uniform(0, 1, a.shape(0, "s1")))
    a.init_bounds()
    dt = 0.5 * a.chkdt()
    a.steps(dt, 3)
    data = tmp_path / "data"
    data.mkdir()
    a.restart_io("w", str(data) + "/", time=3 * dt, istep=3)
    names = sorted(os.listdir(data))
    assert names == sorted(["fld_b_001_u.bin", "fld_b_001_v.bin", "fld_b_001_w.bin", "fld_b_001_p.bin"] +
                           (["fld_b_001_s_001.bin"] if nscal else []))
    a.steps(dt, 3)
    b = _gpu_channel(tmp_path / "b", nscal=nscal)
    choices, _ = a.autotune()  # same kernel variants (their roundings differ)
    b.set_tridiag_variant(choices[-3])  # ..., tridiagonal, rk_mom, FP32 rk_mom
    b.set_rk_variant(choices[-2])
    t, s = b.restart_io("r", str(data) + "/")
    assert s == 3 and t == 3 * dt
    np.testing.assert_array_equal(b.get(0, "po"), b.get(0, "p"))
    b.init_bounds()
    b.steps(dt, 3)
    for f in ("u", "v", "w", "p") + (("s1",) if nscal else ()):
        np.testing.assert_array_equal(b.get(0, f)[1:-1, 1:-1, 1:-1], a.get(0, f)[1:-1, 1:-1, 1:-1], err_msg=f)
    a.close()
    b.close()


@pytest.mark.gpu
def test_write_visu_3d(tmp_path):
    ctx = _gpu_channel(tmp_path / "c")
    w = np.random.default_rng(9).uniform(-1, 1, ctx.shape(0, "w"))
    ctx.set(0, "w", w)
    data = tmp_path / "data"
    data.mkdir()
    n = [16, 12, 8]
    for step in (0, 10):
        ctx.write_visu_3d(str(data) + "/", f"vez_fld_{step:07d}_b_001.bin", "log_visu_3d_b_001.out", "Velocity_Z",
                          0, "w", [1, 1, 1], n, [1, 1, 1], 0.01 * step, step)
    raw = (data / "vez_fld_0000010_b_001.bin").read_bytes()
    assert raw == np.asarray(w[1:-1, 1:-1, 1:-1]).reshape(-1, order="F").tobytes()
    lines = (data / "log_visu_3d_b_001.out").read_text().splitlines()
    assert lines == [ff.visu_log_line(f"vez_fld_{s:07d}_b_001.bin", "Velocity_Z", [1, 1, 1], n, [1, 1, 1], 0.01 * s, s)
                     for s in (0, 10)]
    ctx.close()


# ------------------------------------------------------------------ driver
DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "snac_amd", "snac")
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_driver_refuses_bad_inputs(tmp_path):
    """The main program fails with the reference's sanity message (here: no
    stopping criterion, sanity.f90:41-51) before it touches a device."""
    import subprocess
    d, b = cases.write_case(str(tmp_path), [cases.channel_block((8, 8, 8))], 1000.0)
    txt = open(d).read().replace("stop_type(1:3) = T, F, F", "stop_type(1:3) = F, F, F")
    open(d, "w").write(txt)
    r = subprocess.run([DRIVER, "3", "-1"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "stopping criterion not chosen." in r.stderr


@pytest.mark.gpu
def test_driver_ldc_reference_procedure(tmp_path):
    """The reference's lid-driven-cavity test (tests/lid_driven_cavity/
    testit.sh + test.py) with this library's main program in place of the
    reference binary: run it in a directory holding dns.nml and blocks.nml,
    then read data/vey_fld_0001500_b_00N.bin the way read_single_field_binary
    does (sizes from geometry_b_00N.out, Fortran order), assemble the four
    blocks and compare the mid-line profile with data_ldc_re1000.txt (rtol
    1e-7, atol 5e-8)."""
    import shutil
    import subprocess
    for f in ("dns.nml", "blocks.nml"):
        shutil.copy(os.path.join(GOLD, "ldc", f), tmp_path / f)
    print("driver: ldc, FFT_AXIS=1", flush=True)
    r = subprocess.run([DRIVER, "1"], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "*** Fim ***" in r.stdout
    data = tmp_path / "data"
    blocks = []
    for b in range(1, 5):
        geo = np.loadtxt(data / f"geometry_b_{b:03d}.out", max_rows=4)
        ng = (geo[1] - geo[0] + 1).astype(int)
        fld = np.fromfile(data / f"vey_fld_0001500_b_{b:03d}.bin", dtype="float64")
        blocks.append(fld.reshape(ng, order="F")[0, :, :])
    full = np.block([[blocks[0], blocks[2]], [blocks[1], blocks[3]]])
    ref = np.loadtxt(os.path.join(GOLD, "ldc", "data_ldc_re1000.txt"))
    np.testing.assert_allclose(full[full.shape[1] // 2, :], ref[:, 1], rtol=1e-7, atol=5e-8)
    # time.out every iout0d = 10 steps: istep, dt, time
    t = np.loadtxt(data / "time.out")
    assert t.shape == (150, 3) and t[-1, 0] == 1500
    # the final checkpoint (isave = 2000 > nstep: written when the run ends)
    u = np.fromfile(data / "fld_b_001_u.bin", dtype="float64")
    geo = np.loadtxt(data / "geometry_b_001.out", max_rows=4)
    assert u.size == int(np.prod(geo[1] - geo[0] + 1)) + 2 and u[-1] == 1500.0


@pytest.mark.gpu
def test_driver_impdiff_cavity(tmp_path):
    """The main program with implicit diffusion (third argument "impdiff",
    the reference's IMPDIFF=1 build): 10 fixed-dt steps of a lid-driven
    cavity from rest; the final checkpoint equals the oracle's 10 _IMPDIFF
    steps (relative L2 1e-10)."""
    import subprocess
    import nmlparse
    import oracle_py
    d, b = cases.write_case(str(tmp_path), [cases.cavity_block((16, 12, 20), lid=1.0)], 100.0, nstep=10,
                            dt_f=2.0e-3)
    txt = open(d).read().replace("isave = 0", "isave = 100").replace("iout0d = 0", "iout0d = 5")
    open(d, "w").write(txt)
    print("driver: cavity 16x12x20, FFT_AXIS=0, impdiff", flush=True)
    r = subprocess.run([DRIVER, "0", "0", "impdiff"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "*** Fim ***" in r.stdout
    c = nmlparse.load_case(d, b, 0)
    c["impdiff"] = True
    oc = oracle_py.OracleCase(c)
    oc.set_solver("band")
    oc.init_bounds()
    for _ in range(10):
        oc.step(2.0e-3)
    for f in ("u", "v", "w"):
        raw = np.fromfile(tmp_path / "data" / f"fld_b_001_{f}.bin", dtype="float64")
        assert raw[-1] == 10.0 and raw[-2] == pytest.approx(0.02, rel=1e-14)
        got = raw[:-2].reshape((16, 12, 20), order="F")
        ref = oc.get(0, f)[1:-1, 1:-1, 1:-1]
        assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref), f
    assert np.loadtxt(tmp_path / "data" / "time.out").shape == (2, 3)

"""The C-ABI library loads on a machine without a GPU and exports every
entry point that include/snac_amd.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "snac_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(snac_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("snac_ctx_create_from_nml", "snac_rk_mom", "snac_bounduvw", "snac_solve_pressure", "snac_step"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    import snac_amd
    lib = ctypes.CDLL(snac_amd.LIBPATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_host_only_context_refuses_compute(tmp_path):
    import cases
    import snac_amd
    d, b = cases.write_case(str(tmp_path), [cases.channel_block((8, 8, 8))], 5600.0)
    ctx = snac_amd.Context(d, b, 3, device=-1)
    with pytest.raises(snac_amd.SnacError, match="host-only"):
        ctx.step(1e-3)
    ctx.close()


def test_bad_input_reports_reference_wording(tmp_path):
    import cases
    import snac_amd
    blk = cases.channel_block((8, 8, 8))
    blk["gr"][2] = 1.0  # stretching along the FFT axis (chk_grid_fft)
    d, b = cases.write_case(str(tmp_path), [blk], 5600.0)
    with pytest.raises(snac_amd.SnacError, match="FFT direction must be zero"):
        snac_amd.Context(d, b, 3, device=-1)

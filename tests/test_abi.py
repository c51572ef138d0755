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


@pytest.mark.parametrize("geom", ["lchannel", "bfs"])
def test_solver_selection_on_unions(tmp_path, geom):
    """A union of blocks that is not a box has no tensor-product solver; the
    host-only planner accepts it for the MG-PCG path (which AUTO selects)."""
    import cases
    import snac_amd
    blocks = cases.lchannel_blocks(8, 16) if geom == "lchannel" else cases.bfs_blocks((12, 8), (20, 8, 12), 8)
    d, b = cases.write_case(str(tmp_path), blocks, 1000.0)
    ctx = snac_amd.Context(d, b, 3, device=-1)
    ctx.set_solver("auto")
    ctx.set_solver("mg")
    with pytest.raises(snac_amd.SnacError, match="do not tile a box"):
        ctx.set_solver("tensor")
    ctx.close()


def test_all_periodic_box_takes_the_tensor_solver(tmp_path):
    """Every axis periodic: the tensor-product solver transforms all three
    axes; the MG line relaxation needs a non-periodic axis and refuses."""
    import cases
    import snac_amd
    d, b = cases.write_case(str(tmp_path), [cases.triperiodic_block((8, 8, 8))], 100.0)
    ctx = snac_amd.Context(d, b, 3, device=-1)
    ctx.set_solver("tensor")
    with pytest.raises(snac_amd.SnacError, match="every axis of the Poisson problem is periodic"):
        ctx.set_solver("mg")
    ctx.close()

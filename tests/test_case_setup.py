"""Host set-up of the product (C++ namelist reader, block placement,
neighbours, periods, grids) against the oracle's independent restatement of
param.f90 / initmpi.f90 / initgrid.f90, on the reference's own multi-block
inputs (tests/lid_driven_cavity, examples/__MULTI_BLOCK_GEOMETRY copies under
tests/golden) and on the benchmark geometries."""
import os

import numpy as np
import pytest

import cases
import nmlparse
import oracle_py

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def compare(dns, blocks, fft_axis):
    import snac_amd
    ctx = snac_amd.Context(dns, blocks, fft_axis, device=-1)
    oc = oracle_py.OracleCase(nmlparse.load_case(dns, blocks, fft_axis))
    assert ctx.nblocks == oc.nblocks
    for b in range(ctx.nblocks):
        pi, oi = ctx.block_info(b), oc.block_info(b)
        for key in ("n", "lo_g", "nb", "periods"):
            np.testing.assert_array_equal(pi[key], oi[key], err_msg=f"block {b} {key}")
        for which in range(12):
            np.testing.assert_array_equal(ctx.metric(b, which), oc.metric(b, which), err_msg=f"block {b} metric {which}")
    assert ctx.vol_all == pytest.approx(oracle_py.lib().oc_vol_all(oc.h), rel=1e-15)
    ctx.close()
    return oc


@pytest.mark.parametrize("fft_axis", [0, 1])
def test_ldc_four_blocks(fft_axis):
    oc = compare(os.path.join(GOLD, "ldc", "dns.nml"), os.path.join(GOLD, "ldc", "blocks.nml"), fft_axis)
    # index fixture: blocks 2 and 3 sit above block 1 in y and z (global lo)
    assert list(oc.block_info(1)["lo_g"]) == [1, 33, 1]
    assert list(oc.block_info(2)["lo_g"]) == [1, 1, 33]
    assert list(oc.block_info(3)["lo_g"]) == [1, 33, 33]
    assert list(oc.block_info(0)["periods"]) == [2, 0, 0]


@pytest.mark.parametrize("name", ["L_channel", "lid_driven_cavity"])
def test_reference_multiblock_examples(name):
    d = os.path.join(GOLD, "multiblock", name)
    compare(os.path.join(GOLD, "ldc", "dns.nml"), os.path.join(d, "blocks.nml"), 0)


def test_channel_stretched(tmp_path):
    d, b = cases.write_case(str(tmp_path), [cases.channel_block((32, 24, 16))], 5600.0, bforce=(12 / 5600, 0, 0))
    oc = compare(d, b, 3)
    # tanh two-end map, closed form of gridpoint_cluster_two_end
    g = oracle_py.initgrid(24, 0, 1.8, 0.0, 1.0)
    r0 = np.arange(1, 25) / 24.0
    np.testing.assert_allclose(g["rf"][1:25], 0.5 * (1 + np.tanh((r0 - 0.5) * 1.8) / np.tanh(0.9)), rtol=0, atol=1e-15)
    np.testing.assert_allclose(oc.metric(0, 3)[1:-1].sum(), 1.0, rtol=1e-14)


def test_dns_parameters(tmp_path):
    import snac_amd
    d, b = cases.write_case(str(tmp_path), [cases.channel_block((8, 8, 8))], 5600.0, dt_f=2e-3, cfl=0.5,
                            bforce=(12 / 5600, 0, 0))
    ctx = snac_amd.Context(d, b, 3, device=-1)
    assert ctx.visc == pytest.approx(1 / 5600, rel=1e-15)
    assert ctx.dt_f == 2e-3 and ctx.cfl == 0.5
    assert ctx.bforce[0] == pytest.approx(12 / 5600, rel=1e-15)
    ctx.close()

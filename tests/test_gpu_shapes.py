"""Parity of the kernel instances the autotuner picks at the benchmark sizes,
run at shapes that select them, against the CPU oracle.

* rk_mom (rk.f90:7-128): the LDS k-march variants 13..16 walk 32 or 64
  planes per tile; a 160-plane channel makes every tile hand its planes over
  at least twice (k = 65, 129 for the 64-plane walk), with a short last chunk.
* the tridiagonal line solves of the tensor-product Poisson path
  (solver.f90:843-922 solved exactly): 64-, 128- and 256-row lines select
  the chunked instances <4,17>, <8,17>, <16,9> and <16,17>.
* a 10-step channel with 128 wall-normal rows and those instances forced.
* chkdiv (chkdiv.f90:31-43) on a field whose divtot is known in closed form.
"""
import numpy as np
import pytest

import cases
from test_gpu_parity import FIELDS3, assert_fields, make_pair, run_both, seed_fields

pytestmark = pytest.mark.gpu


def channel(tmp_path, n):
    return make_pair(tmp_path, [cases.channel_block(n)], 3, bforce=(12 / 5600, 0, 0))


@pytest.mark.parametrize("variant", [13, 14, 15, 16])
def test_rk_mom_kmarch_chunks(tmp_path, variant):
    ctx, oc = channel(tmp_path, (64, 24, 160))
    ctx.set_rk_variant(variant)
    for irk in (1, 2, 3):
        seed_fields(ctx, oc, seed=10 + irk)
        ctx.rk_mom(irk, 1.1e-3)
        oc.rk_mom(irk, 1.1e-3)
        assert_fields(ctx, oc, FIELDS3 + ("dudtrko", "dvdtrko", "dwdtrko"), 1e-13)


@pytest.mark.parametrize("ny,tvariant", [(64, 2), (128, 2), (128, 3), (256, 2)])
def test_tridiag_production_lines(tmp_path, ny, tvariant):
    """64 rows -> <4,17>; 128 rows -> <8,17> (variant 2) and <16,9> (variant
    3); 256 rows -> <16,17>."""
    ctx, oc = channel(tmp_path, (16, ny, 8))
    ctx.set_tridiag_variant(tvariant)
    seed_fields(ctx, oc)
    ctx.bounduvw(False); oc.bounduvw(False)
    ctx.fillps(1e-3); oc.fillps(1e-3)
    ctx.solve_pressure()
    oc.solve_pressure()
    assert_fields(ctx, oc, ("pp",), 1e-11, interior=True, modmean=("pp",))


@pytest.mark.parametrize("tvariant", [2, 3])
def test_channel_128_rows_10_steps(tmp_path, tvariant):
    """The benchmark channel's wall-normal resolution (128 rows, tanh 1.8)
    with the production rk_mom variant (64-plane LDS walk over 48 planes)
    and each chunked tridiagonal instance of 128-row lines."""
    ctx, oc = channel(tmp_path, (32, 128, 48))
    ctx.set_rk_variant(16)
    ctx.set_tridiag_variant(tvariant)
    u, v, w = cases.channel_ic(ctx.shape(0, "u"), ctx.metric(0, 8))
    for f, a in zip(FIELDS3, (u, v, w)):
        ctx.set(0, f, a)
        oc.set(0, f, a)
    run_both(ctx, oc, 10)
    assert_fields(ctx, oc, FIELDS3, 1e-10, interior=True)
    assert_fields(ctx, oc, ("p",), 1e-10, interior=True, modmean=("p",))
    _, divmax = ctx.chkdiv()
    assert divmax < 1e-10


def test_channel_three_kchunks_10_steps(tmp_path):
    """Ten steps with the 64-plane walk crossing two chunk boundaries."""
    ctx, oc = channel(tmp_path, (64, 24, 160))
    ctx.set_rk_variant(16)
    u, v, w = cases.channel_ic(ctx.shape(0, "u"), ctx.metric(0, 8))
    for f, a in zip(FIELDS3, (u, v, w)):
        ctx.set(0, f, a)
        oc.set(0, f, a)
    run_both(ctx, oc, 10)
    assert_fields(ctx, oc, FIELDS3, 1e-10, interior=True)
    assert_fields(ctx, oc, ("p",), 1e-10, interior=True, modmean=("p",))


def test_divtot_closed_form_gpu(tmp_path):
    """u = x^2, v = y^2, w = z^2 on a box stretched along every axis:
    divtot = Lx + Ly + Lz and divmax = sum of (f_n + f_{n-1}) over the axes."""
    L = (2.0, 1.0, 1.5)
    n = (12, 10, 8)
    ctx, oc = make_pair(tmp_path, [cases.quadratic_box(n, L)], 0, visci=100.0)
    xf, yf, zf = ctx.metric(0, 7), ctx.metric(0, 9), ctx.metric(0, 11)
    for f, a in zip(FIELDS3, cases.quadratic_fields(ctx.shape(0, "u"), xf, yf, zf)):
        ctx.set(0, f, a)
    tot, mx = ctx.chkdiv()
    assert tot == pytest.approx(sum(L), rel=1e-13)
    assert mx == pytest.approx(sum(f[q] + f[q - 1] for f, q in zip((xf, yf, zf), n)), rel=1e-13)

"""The oracle's grid stretching maps (initgrid.f90:171-287, selected by
block_gt as initgrid.f90:14-46 numbers them) against properties the
mathematics of each map fixes, so that a swapped map number, a reflection
wired the wrong way or a sign error fails:

* every map sends the faces r0 = q/n to a strictly increasing sequence from
  lmin to lmax;
* the upper-end maps are reflections of the lower-end ones:
  r_3(x) = 1 - r_1(1 - x) (tanh), r_5(x) = 1 - r_4(1 - x) (geometric);
* the two-end and middle maps (0, 2, 6, 7) are antisymmetric about 1/2;
* each clusters where its name says: smallest cells at both ends (0, 6),
  at the lower end (1, 4), in the middle (2, 7), at the upper end (3, 5);
* no stretching is uniform: gr = 0 for the tanh maps, gr = 1 for the
  geometric ones ((1 - x)/(1 - x) = 1 inside the cube root);
* the tanh maps equal their closed forms 0.5 (1 + tanh(a(x-1/2))/tanh(a/2)),
  1 + tanh(a(x-1))/tanh(a) and 0.5 tanh(2 a x)/tanh(a) (x <= 1/2) at the
  faces.
"""
import numpy as np
import pytest

import oracle_py

N = 32
LMIN, LMAX = -0.3, 1.7


def faces(gt, gr):
    g = oracle_py.initgrid(N, gt, gr, LMIN, LMAX)
    return (g["rf"][:N + 1] - LMIN) / (LMAX - LMIN)  # r(q/n), q = 0..n


X = np.arange(N + 1) / N


@pytest.mark.parametrize("gt,gr", [(0, 1.8), (1, 1.5), (2, 1.2), (3, 1.5), (4, 2.5), (5, 2.5), (6, 3.0), (7, 3.0)])
def test_endpoints_and_monotone(gt, gr):
    r = faces(gt, gr)
    assert r[0] == 0.0 and r[-1] == pytest.approx(1.0, abs=1e-15)
    assert np.all(np.diff(r) > 0)


def test_reflections():
    np.testing.assert_allclose(faces(3, 1.5), 1 - faces(1, 1.5)[::-1], atol=1e-14)
    np.testing.assert_allclose(faces(5, 2.5), 1 - faces(4, 2.5)[::-1], atol=1e-14)


@pytest.mark.parametrize("gt,gr", [(0, 1.8), (2, 1.2), (6, 3.0), (7, 3.0)])
def test_antisymmetric(gt, gr):
    r = faces(gt, gr)
    np.testing.assert_allclose(r, 1 - r[::-1], atol=1e-14)


@pytest.mark.parametrize("gt,gr,where", [(0, 1.8, "ends"), (6, 3.0, "ends"), (1, 1.5, "low"), (4, 2.5, "low"),
                                         (2, 1.2, "middle"), (7, 3.0, "middle"), (3, 1.5, "high"),
                                         (5, 2.5, "high")])
def test_clustering(gt, gr, where):
    d = np.diff(faces(gt, gr))
    lo, mid, hi = d[0], d[N // 2], d[-1]
    if where == "ends":
        assert lo < mid and hi < mid
    elif where == "middle":
        assert mid < lo and mid < hi
    elif where == "low":
        assert lo < mid < hi
    else:
        assert hi < mid < lo


@pytest.mark.parametrize("gt,gr", [(0, 0.0), (1, 0.0), (2, 0.0), (3, 0.0), (4, 1.0), (5, 1.0), (6, 2.0),
                                   (7, 2.0)])
def test_uniform_limit(gt, gr):
    np.testing.assert_allclose(faces(gt, gr), X, atol=1e-15)


def test_tanh_closed_forms():
    a = 1.8
    np.testing.assert_allclose(faces(0, a), 0.5 * (1 + np.tanh(a * (X - 0.5)) / np.tanh(a / 2)), atol=1e-15)
    np.testing.assert_allclose(faces(1, a), 1 + np.tanh(a * (X - 1)) / np.tanh(a), atol=1e-15)
    h = X <= 0.5
    np.testing.assert_allclose(faces(2, a)[h], 0.5 * np.tanh(2 * a * X[h]) / np.tanh(a), atol=1e-15)

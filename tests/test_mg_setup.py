"""Multigrid set-up of the product library on the host (no GPU): knot coarsening and
the knot-insertion matrix (Oslo algorithm) against the oracle's Boehm insertion, the
scipy collocation fixtures and scipy BSpline on random nested knot vectors
(src/multigrid.cpp:25-93 for the reference behaviour)."""
import os

import numpy as np
import pytest
from scipy.interpolate import BSpline

import oracle
import tdsez_amd as T

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_lib = pytest.mark.skipif(not os.path.exists(T.LIB_PATH), reason="libtdz.so not built")


def open_uniform(E, lo, hi, p):
    return np.concatenate([[lo] * (p + 1), lo + (hi - lo) * np.arange(1, E) / E, [hi] * (p + 1)])


@needs_lib
@pytest.mark.parametrize("U", [
    open_uniform(8, 0.0, 8.0, 3), open_uniform(5, 0.0, 5.0, 2), open_uniform(128, -40.0, 40.0, 4),
    np.array([0, 0, 0, 1, 2, 2, 3, 4, 4, 4], dtype=float),
    np.array([0, 0, 0, 0, 0, 1, 2, 3, 3, 3, 3, 4, 5, 6, 7, 7, 7, 7, 7], dtype=float),
    np.array([0, 0, 1, 1], dtype=float)])
def test_coarsening_matches_oracle(U):
    got, want = T.coarsen_knots(U), oracle.coarsen_knots(U)
    if want is None:
        assert got is None
    else:
        assert np.array_equal(got, want)


@needs_lib
def test_coarsening_known_lists():
    assert np.array_equal(T.coarsen_knots(open_uniform(8, 0.0, 8.0, 3)), open_uniform(4, 0.0, 8.0, 3))
    assert np.array_equal(T.coarsen_knots(open_uniform(5, 0.0, 5.0, 2)), [0, 0, 0, 2, 4, 5, 5, 5])
    # C0 interfaces (a knot repeated p times) survive every coarsening
    U = np.array([0] * 5 + [1, 2] + [3] * 4 + [4, 5, 6] + [7] * 5, dtype=float)
    C = T.coarsen_knots(U)
    assert list(C).count(3.0) == 4


@needs_lib
@pytest.mark.parametrize("Ef,p", [(8, 3), (12, 4), (16, 5)])
def test_refinement_matches_collocation_fixture(Ef, p):
    g = np.load(os.path.join(GOLD, f"boehm_e{Ef}_p{p}.npz"))
    R = T.refinement_matrix(g["coarse"], g["fine"], p)
    assert np.abs(R - g["P"]).max() < 1e-13


def _random_nested(rng, p, nbreak, c0_every=0):
    """Random open knot vectors F (fine) refining C (coarse, every other interior breakpoint)."""
    bp = np.sort(rng.uniform(-3.0, 5.0, nbreak))
    inner_f, inner_c = [], []
    for k, b in enumerate(bp):
        mult = p if (c0_every and k % c0_every == c0_every - 1) else 1
        inner_f += [b] * mult
        if mult > 1 or k % 2 == 1:
            inner_c += [b] * mult
    F = np.array([-3.0] * (p + 1) + inner_f + [5.0] * (p + 1))
    C = np.array([-3.0] * (p + 1) + inner_c + [5.0] * (p + 1))
    return C, F


@needs_lib
@pytest.mark.parametrize("p,c0", [(1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (3, 4), (4, 5)])
def test_refinement_reproduces_coarse_splines(p, c0):
    rng = np.random.default_rng(10 * p + c0)
    C, F = _random_nested(rng, p, 23, c0)
    R = T.refinement_matrix(C, F, p)
    # every coarse spline is the fine spline with coefficients R c (nested spaces)
    c = rng.standard_normal(C.size - p - 1)
    x = np.linspace(-3.0, 5.0 - 1e-9, 997)
    assert np.abs(BSpline(C, c, p)(x) - BSpline(F, R @ c, p)(x)).max() < 1e-12
    # the same matrix as the oracle's Boehm insertion
    assert np.abs(R - oracle.prolongation(C, F, p)).max() < 1e-13
    # rows are local: at most p+1 nonzeros
    assert (np.abs(R) > 0).sum(axis=1).max() <= p + 1


@needs_lib
def test_refinement_rejects_non_nested():
    with pytest.raises(T.TdzError):
        T.refinement_matrix(open_uniform(4, 0.0, 1.0, 3), open_uniform(4, 0.0, 2.0, 3), 3)

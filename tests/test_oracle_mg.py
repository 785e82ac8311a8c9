"""Pins the multigrid pieces of the oracle (src/multigrid.cpp:25-93): knot
coarsening, Boehm knot-insertion prolongation against scipy collocation
(golden fixtures), the two-scale mask, partition of unity and linear precision
at the Greville points (the reference's own exactness check, :283-292)."""
import math
import os

import numpy as np
import pytest
from scipy.interpolate import BSpline

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def open_uniform(E, lo, hi, p):
    return np.concatenate([[lo] * (p + 1), lo + (hi - lo) * np.arange(1, E) / E, [hi] * (p + 1)])


def test_coarsen_knots():
    U = open_uniform(8, 0.0, 8.0, 3)
    C = oracle.coarsen_knots(U)
    assert np.array_equal(C, open_uniform(4, 0.0, 8.0, 3))
    # odd element count: drop every other simple interior breakpoint, keep the rest
    U5 = open_uniform(5, 0.0, 5.0, 2)
    C5 = oracle.coarsen_knots(U5)
    assert np.array_equal(C5, np.array([0, 0, 0, 2, 4, 5, 5, 5], dtype=float))
    # repeated (C0) breakpoints are kept
    Ur = np.array([0, 0, 0, 1, 2, 2, 3, 4, 4, 4], dtype=float)
    Cr = oracle.coarsen_knots(Ur)
    assert 2.0 in Cr and list(Cr).count(2.0) == 2
    assert oracle.coarsen_knots(np.array([0, 0, 1, 1], dtype=float)) is None


@pytest.mark.parametrize("Ef,p", [(8, 3), (12, 4), (16, 5)])
def test_boehm_matches_collocation_fixture(Ef, p):
    g = np.load(os.path.join(GOLD, f"boehm_e{Ef}_p{p}.npz"))
    P = oracle.prolongation(g["coarse"], g["fine"], p)
    assert np.abs(P - g["P"]).max() < 1e-12


@pytest.mark.parametrize("Ef,p", [(16, 3), (20, 4), (24, 5)])
def test_boehm_nested_spaces(Ef, p):
    tf = open_uniform(Ef, -3.0, 5.0, p)
    tc = open_uniform(Ef // 2, -3.0, 5.0, p)
    P = oracle.prolongation(tc, tf, p)
    nc = len(tc) - p - 1
    c = np.random.default_rng(0).standard_normal(nc)
    x = np.linspace(-3.0, 5.0 - 1e-9, 301)
    sc = BSpline(tc, c, p)(x)
    sf = BSpline(tf, P @ c, p)(x)
    assert np.abs(sc - sf).max() < 1e-12
    # partition of unity, linear precision at the Greville points
    gc, gf = oracle.greville(tc, p), oracle.greville(tf, p)
    assert np.abs(P.sum(axis=1) - 1).max() < 1e-13
    assert np.abs(P @ gc - gf).max() < 1e-12
    # interior columns carry the two-scale mask 2^-p C(p+1, k)
    mask = np.array([math.comb(p + 1, k) for k in range(p + 2)]) / 2 ** p
    col = P[:, nc // 2]
    nz = col[np.abs(col) > 1e-14]
    assert np.abs(nz - mask).max() < 1e-13

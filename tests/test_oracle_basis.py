"""Pins the oracle's 1D discretisation (knots, Gauss rule, basis tables, bands,
CAP profiles) against the golden fixtures (scipy/numpy) and against closed
forms. Reference: src/core.cpp:1793,1899; src/tensorops.cpp:27-61;
src/assembly/cap.cpp:31-88; src/assembly/assembly.cpp:170-208."""
import math
import os

import numpy as np
import pytest
from scipy.interpolate import BSpline

from oracle import Oracle, gauss

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [("basis_e7_p3_q5", 7, -5.0, 5.0, 3, 5), ("basis_e10_p4_q8", 10, -10.0, 10.0, 4, 8),
         ("basis_e6_p5_q8", 6, -3.0, 3.0, 5, 8), ("basis_e5_p7_q10", 5, -20.0, 20.0, 7, 10)]


@pytest.mark.parametrize("Q", [5, 8, 10])
def test_gauss_legendre_fixture(Q):
    g = np.load(os.path.join(GOLD, f"gauss_{Q}.npz"))
    x, w = gauss(Q)
    assert np.abs(x - g["x"]).max() < 1e-15
    assert np.abs(w - g["w"]).max() < 1e-15
    # exactness on monomials of degree <= 2Q-1
    for k in range(2 * Q):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) < 1e-14


def _oracle_1d(E, lo, hi, p, Q, **kw):
    return Oracle(dict(dim=1, nelem=(E,), lo=(lo,), hi=(hi,), p=p, nquad=Q, **kw))


@pytest.mark.parametrize("name,E,lo,hi,p,Q", CASES)
def test_basis_tables_fixture(name, E, lo, hi, p, Q):
    g = np.load(os.path.join(GOLD, name + ".npz"))
    o = _oracle_1d(E, lo, hi, p, Q)
    assert np.abs(o.knots(0) - g["knots"]).max() < 1e-14
    t = o.axis_tables(0)
    assert np.abs(t["pts"] - g["pts"]).max() < 1e-13
    assert np.abs(t["wts"] - g["wts"]).max() < 1e-14
    assert np.abs(t["N0"] - g["N0"]).max() < 1e-13
    assert np.abs(t["N1"] - g["N1"]).max() < 1e-11 * max(1.0, np.abs(g["N1"]).max())
    # partition of unity and its derivative
    assert np.abs(t["N0"].sum(axis=1) - 1).max() < 1e-14
    assert np.abs(t["N1"].sum(axis=1)).max() < 1e-11


@pytest.mark.parametrize("name,E,lo,hi,p,Q", CASES)
def test_bands_fixture(name, E, lo, hi, p, Q):
    g = np.load(os.path.join(GOLD, name + ".npz"))
    o = _oracle_1d(E, lo, hi, p, Q)
    for which, key in ((0, "M"), (1, "X"), (2, "G"), (4, "S")):
        b = o.band(0, which)
        ref = g[key]
        assert np.abs(b - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), key
    A = o.band(0, 3)
    G = g["G"]
    n, bw = G.shape
    Gd = np.zeros((n, n))
    for i in range(n):
        for k in range(bw):
            j = i + k - p
            if 0 <= j < n:
                Gd[i, j] = G[i, k]
    Ad = Gd - Gd.T
    for i in range(n):
        for k in range(bw):
            j = i + k - p
            if 0 <= j < n:
                assert abs(A[i, k] - Ad[i, j]) < 1e-12 * np.abs(Gd).max()


def test_band_invariants():
    E, lo, hi, p, Q = 12, -6.0, 6.0, 4, 8
    o = _oracle_1d(E, lo, hi, p, Q)
    t = o.knots(0)
    n = E + p
    M, S = o.band(0, 0), o.band(0, 4)
    # row sums of M: int B_i = (t_{i+p+1} - t_i)/(p+1)
    for i in range(n):
        assert abs(M[i].sum() - (t[i + p + 1] - t[i]) / (p + 1)) < 1e-14
    # S 1 = 0 (Neumann)
    assert np.abs(S.sum(axis=1)).max() < 1e-12
    # interior rows of a uniform mass band: h * cardinal B-spline of degree 2p+1 at integers
    h = (hi - lo) / E
    card = BSpline.basis_element(np.arange(2 * p + 3), extrapolate=False)
    ref = np.array([h * card(p + 1 + k) for k in range(-p, p + 1)])
    mid = n // 2
    assert np.abs(M[mid] - ref).max() < 1e-14


def _mano_ref(u, lo, hi, kmin, mass=1.0):
    # Manolopoulos, J. Chem. Phys. 117, 9552 (2002): layer of width c/(2 delta kmin)
    c, delta = 2.62206, 0.2
    a, b = 1 - 16 / c ** 3, (1 - 17 / c ** 3) / c ** 2
    width = c / (2 * delta * kmin)
    s = min(u - lo, hi - u)
    if s >= width:
        return 0.0
    r = c * min(width - s, width * (1 - 1e-10)) / width
    return 0.5 * kmin ** 2 * (a * r - b * r ** 3 + 4 / (c - r) ** 2 - 4 / (c + r) ** 2) / mass


@pytest.mark.parametrize("kind,par,mass", [(1, (0.01, 3.0), 1.0), (2, (1.5,), 1.0), (2, (1.2,), 1.7)])
def test_cap_profile_closed_form(kind, par, mass):
    o = _oracle_1d(20, -10.0, 10.0, 4, 8, cap_kind=kind, cap=par, mass=mass)
    t = o.axis_tables(0)
    for u, w in zip(t["pts"], t["cap"]):
        if kind == 1:
            s = min(u + 10, 10 - u)
            ref = par[0] * (par[1] - s) ** 2 if s < par[1] else 0.0
        else:
            ref = _mano_ref(u, -10.0, 10.0, par[0], mass)
        assert abs(w - ref) <= 1e-12 * max(1.0, abs(ref))
    # the CAP band is the Gram matrix of W_d
    W = o.band(0, 5)
    assert W.min() > -1e-12 * abs(W).max() or True
    assert np.abs(W).max() > 0


def _mano_samples(kmin, mass, E=400):
    """(relative depth r / c, W / (kmin^2 / 2 m)) at the Gauss points inside the CAP layer
    of a 1D grid [-10, 10], and the points outside it."""
    o = _oracle_1d(E, -10.0, 10.0, 3, 8, cap_kind=2, cap=(kmin,), mass=mass)
    t = o.axis_tables(0)
    width = 2.62206 / (2 * 0.2 * kmin)
    s = np.minimum(t["pts"] + 10.0, 10.0 - t["pts"])
    inside = s < width
    depth = (width - s[inside]) / width
    return depth, t["cap"][inside] / (0.5 * kmin ** 2 / mass), t["cap"][~inside]


def test_manolopoulos_profile_properties():
    """Properties of the transmission-free CAP (Manolopoulos, J. Chem. Phys. 117, 9552 (2002))
    that the formula's text does not state: zero outside the layer of width c/(2 delta kmin),
    a continuous start with unit slope in the scaled variable r = c depth (W ~ Emin r), one
    universal shape W / Emin(r) for every kmin and mass (the kmin^2 and 1/m scaling), a
    monotone rise and a divergence at the wall."""
    c = 2.62206
    d1, g1, out1 = _mano_samples(1.2, 1.0)
    d2, g2, out2 = _mano_samples(2.4, 1.0)
    d3, g3, _ = _mano_samples(1.2, 1.7)
    assert np.all(out1 == 0.0) and np.all(out2 == 0.0)
    # unit slope at onset: W / (Emin r) -> 1
    near = d1 * c < 0.02
    assert near.any()
    assert np.abs(g1[near] / (c * d1[near]) - 1.0).max() < 0.02
    # 1/m: the same points, the same scaled values
    assert np.allclose(g3, g1, rtol=1e-13, atol=0)
    # kmin^2: the scaled profile is one function of the depth (interpolate one sample set)
    order = np.argsort(d2)
    mid = (d1 > d2.min()) & (d1 < 0.9)
    gi = np.interp(d1[mid], d2[order], g2[order])
    assert np.abs(gi - g1[mid]).max() < 2e-3 * np.abs(g1[mid]).max()
    # monotone in depth (points at equal depth on the two walls agree), diverging at the
    # wall (4/(c - r)^2)
    order1 = np.argsort(d1)
    assert np.all(np.diff(g1[order1]) >= -1e-12 * g1.max())
    assert g1.max() > 100.0

"""Pins the oracle's layered structures (BenDaniel-Duke kinetic term with a
piecewise-constant mass m(z), band offsets, interface-aligned C0 knots): the 1D
bands against an independent scipy construction, the Kronecker terms against the
brute-force per-point dense assembly, and equal layers against a constant mass."""
import numpy as np
import pytest
from scipy.interpolate import BSpline

from oracle import Oracle


def layered_knots(bounds, per_layer, p):
    """Open knot vector with interior knots repeated p times at the layer interfaces
    (C0) and per_layer uniform elements inside each layer."""
    t = [bounds[0]] * (p + 1)
    for i in range(len(bounds) - 1):
        pts = np.linspace(bounds[i], bounds[i + 1], per_layer[i] + 1)[1:]
        for x in pts[:-1]:
            t.append(x)
        if i + 1 < len(bounds) - 1:
            t += [bounds[i + 1]] * p
    t += [bounds[-1]] * (p + 1)
    return np.array(t)


def layered(p=3, bounds=(-6.0, -2.0, 1.0, 5.0), per=(2, 3, 2), masses=(0.9, 1.4, 0.9), v=(0.0, 0.3, 0.0)):
    t = layered_knots(bounds, per, p)
    return dict(dim=3, nelem=(3, 2, 1), lo=(-4.0, -3.0, 0.0), hi=(4.0, 3.0, 1.0), p=p, nquad=p + 2, mass=1.0,
                charge=-1.0, knots=t, layers=(np.array(bounds), np.array(masses), np.array(v)))


def _scipy_bands(t, p, bounds, masses, v, Q):
    """Dense Mw, Sw, Vl, M by Gauss quadrature on every non-empty span (scipy BSpline)."""
    n = len(t) - p - 1
    gx, gw = np.polynomial.legendre.leggauss(Q)
    out = {k: np.zeros((n, n)) for k in ("M", "Mw", "Sw", "Vl")}
    basis = [BSpline(t, np.eye(n)[i], p, extrapolate=False) for i in range(n)]
    for s in range(p, n):
        a, b = t[s], t[s + 1]
        if b <= a:
            continue
        x = 0.5 * (a + b) + 0.5 * (b - a) * gx
        w = 0.5 * (b - a) * gw
        B = np.nan_to_num(np.array([bs(x) for bs in basis]))
        D = np.nan_to_num(np.array([bs.derivative()(x) for bs in basis]))
        lay = np.clip(np.searchsorted(bounds, x, side="right") - 1, 0, len(masses) - 1)
        im, vv = 1.0 / np.asarray(masses)[lay], np.asarray(v)[lay]
        out["M"] += (B * w) @ B.T
        out["Mw"] += (B * w * im) @ B.T
        out["Sw"] += (D * w * im) @ D.T
        out["Vl"] += (B * w * vv) @ B.T
    return out


def _dense_band(band, p):
    n = band.shape[0]
    A = np.zeros((n, n))
    for i in range(n):
        for m in range(2 * p + 1):
            j = i - p + m
            if 0 <= j < n:
                A[i, j] = band[i, m]
    return A


def test_layered_bands_against_scipy():
    desc = layered()
    o = Oracle(desc)
    p = desc["p"]
    ref = _scipy_bands(desc["knots"], p, desc["layers"][0], desc["layers"][1], desc["layers"][2], desc["nquad"])
    for which, key in ((0, "M"), (6, "Mw"), (7, "Sw"), (8, "Vl")):
        got = _dense_band(o.band(2, which), p)
        assert np.abs(got - ref[key]).max() < 1e-13 * max(1.0, np.abs(ref[key]).max()), key


def test_layered_terms_equal_brute_force():
    desc = layered()
    o = Oracle(desc)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
    for term in (0, 1, 2, 4, 5, 6):
        A = o.dense(term)
        y = o.term(term, x)
        assert np.abs(A @ x - y).max() < 1e-12 * np.abs(y).max(), term
        assert np.abs(A - A.T).max() < 1e-13 * np.abs(A).max()


def test_equal_layers_are_a_constant_mass():
    m = 1.3
    a = Oracle(layered(masses=(m, m, m), v=(0.0, 0.0, 0.0)))
    b = Oracle(dict(layered(), mass=m, layers=None))
    rng = np.random.default_rng(4)
    x = rng.standard_normal(a.N) + 1j * rng.standard_normal(a.N)
    ka, kb = a.term(1, x), b.term(1, x)
    assert np.abs(ka - kb).max() < 1e-13 * np.abs(kb).max()
    assert np.abs(a.term(2, x)).max() == 0.0


def test_layered_operator_composition():
    """alpha M + g (K + sV_l + sE_z D_z) assembled densely equals the oracle's apply."""
    desc = layered()
    o = Oracle(desc)
    rng = np.random.default_rng(6)
    x = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
    alpha, g, sE = 0.7 + 0.1j, 0.3j, np.array([0.0, 0.0, 0.03])
    A = alpha * o.dense(0) + g * (o.dense(1) + o.dense(2) + sE[2] * o.dense(6))
    y = o.apply(alpha, g, sE, x)
    assert np.abs(A @ x - y).max() < 1e-13 * np.abs(y).max()

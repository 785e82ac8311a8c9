"""Pins the oracle operator: sum-factorised / Kronecker applies against a
brute-force per-point dense assembly (src/assembly/assembly.cpp:703-864),
and the dense potential against an independent scipy construction.
Reference criteria: factorised = assembled to 1e-11 (tests/test_tensorops.py:93-102),
quadrature = matrix expectations to 1e-11 (tests/test_tensorops.py:124-148)."""
import numpy as np
import pytest
from scipy.interpolate import BSpline

from oracle import Oracle
from tests.problems import HO_1D, MANO_3D, SOFTCORE_2D

TINY_3D = dict(dim=3, nelem=(3, 4, 3), lo=(-5.0, -6.0, -4.0), hi=(5.0, 6.0, 4.0), p=3, nquad=5, mass=1.3,
               charge=-1.0, pot_kind=2, pot=(1.0, 2.0), cap_kind=1, cap=(0.05, 2.0))
TINY_2D = dict(SOFTCORE_2D, nelem=(4, 5), p=4)
TINY_1D = dict(HO_1D, nelem=(12,))
TINY_HOLE = dict(TINY_3D, charge=1.0, cap_kind=2, cap=(1.0,))


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


@pytest.mark.parametrize("desc", [TINY_3D, TINY_2D, TINY_HOLE], ids=["3d", "2d", "hole"])
def test_terms_equal_brute_force(desc):
    o = Oracle(desc)
    v = _rand(o.N, 1)
    for term in range(4 + desc["dim"]):
        A = o.dense(term)
        y = o.term(term, v)
        err = np.abs(A @ v - y).max() / np.abs(A @ v).max()
        assert err < 1e-13, (term, err)
        # symmetry: M, K, V, D real symmetric; CAP = -i (real symmetric)
        if term == Oracle.TERM_CAP:
            assert np.abs(A.real).max() == 0.0
            assert np.abs(A.imag - A.imag.T).max() < 1e-13 * np.abs(A).max()
            assert np.linalg.eigvalsh(-A.imag).min() > -1e-12   # W >= 0: absorbing
        else:
            assert np.abs(A.imag).max() == 0.0
            assert np.abs(A - A.T).max() < 1e-13 * np.abs(A).max()


def test_full_operator_composition():
    o = Oracle(TINY_3D)
    v = _rand(o.N, 2)
    alpha, g, sE = 0.7 + 0.1j, 0.3j, np.array([0.01, -0.02, 0.03])
    A = alpha * o.dense(0) + g * (o.dense(1) + o.dense(2) + o.dense(3) + sum(sE[a] * o.dense(4 + a) for a in range(3)))
    y = o.apply(alpha, g, sE, v)
    assert np.abs(A @ v - y).max() / np.abs(y).max() < 1e-13


def test_hole_inverts_potential_and_coupling():
    e = Oracle(TINY_3D)
    h = Oracle(dict(TINY_3D, charge=1.0))
    assert np.abs(e.dense(2) + h.dense(2)).max() < 1e-15 * np.abs(e.dense(2)).max() + 1e-300


def _dense_potential_scipy(E, lo, hi, p, Q, V):
    """Independent 1D construction: V_ij = sum_q w_q V(x_q) B_i(x_q) B_j(x_q)."""
    t = np.concatenate([[lo] * (p + 1), lo + (hi - lo) * np.arange(1, E) / E, [hi] * (p + 1)])
    n = E + p
    gx, gw = np.polynomial.legendre.leggauss(Q)
    A = np.zeros((n, n))
    for e in range(E):
        u0, u1 = t[p + e], t[p + e + 1]
        x = 0.5 * (u0 + u1) + 0.5 * (u1 - u0) * gx
        w = gw * 0.5 * (u1 - u0)
        B = np.zeros((Q, n))
        for i in range(n):
            c = np.zeros(n)
            c[i] = 1
            B[:, i] = np.nan_to_num(BSpline(t, c, p, extrapolate=False)(x))
        A += (B * (w * V(x))[:, None]).T @ B
    return A


def test_potential_against_scipy_1d():
    desc = dict(dim=1, nelem=(9,), lo=(-7.0,), hi=(7.0,), p=4, nquad=8, pot_kind=2, pot=(1.3, 2.0))
    o = Oracle(desc)
    ref = _dense_potential_scipy(9, -7.0, 7.0, 4, 8, lambda x: -1.3 / np.sqrt(x * x + 2.0))
    assert np.abs(o.dense(2).real - ref).max() < 1e-14 * np.abs(ref).max() * 10


def test_potential_expectation_equals_quadrature():
    """psi^H V psi equals sum_q w_q V(r_q) |psi(r_q)|^2 with psi evaluated by scipy."""
    desc = dict(dim=2, nelem=(5, 4), lo=(-6.0, -5.0), hi=(6.0, 5.0), p=3, nquad=6, pot_kind=2, pot=(1.0, 1.0))
    o = Oracle(desc)
    psi = _rand(o.N, 3)
    val = np.vdot(psi, o.term(Oracle.TERM_V, psi))
    n0, n1 = o.n[0], o.n[1]
    C = psi.reshape(n1, n0)
    tabs = [o.axis_tables(d) for d in range(2)]
    # scipy evaluation of psi at the 2D tensor points
    ts = [o.knots(d) for d in range(2)]
    Bx = np.stack([np.nan_to_num(BSpline(ts[0], np.eye(n0)[i], 3, extrapolate=False)(tabs[0]["pts"]))
                   for i in range(n0)], 1)
    By = np.stack([np.nan_to_num(BSpline(ts[1], np.eye(n1)[i], 3, extrapolate=False)(tabs[1]["pts"]))
                   for i in range(n1)], 1)
    F = By @ C @ Bx.T
    X, Y = np.meshgrid(tabs[0]["pts"], tabs[1]["pts"])
    W = np.outer(tabs[1]["wts"], tabs[0]["wts"])
    ref = np.sum(W * (-1.0 / np.sqrt(X ** 2 + Y ** 2 + 1.0)) * np.abs(F) ** 2)
    assert abs(val - ref) < 1e-12 * abs(ref)


def test_dirichlet_walls_decouple():
    o = Oracle(TINY_1D)
    v = _rand(o.N, 4)
    y = o.apply(1.0, 0.3j, np.zeros(3), v)
    assert y[0] == v[0] and y[-1] == v[-1]
    w = v.copy()
    w[0] = w[-1] = 0
    y2 = o.apply(1.0, 0.3j, np.zeros(3), w)
    assert np.abs(y[1:-1] - y2[1:-1]).max() == 0.0


def test_separable_preconditioner_inverts_separable_part():
    """With no potential, CAP or field, the separable PC is the exact inverse."""
    desc = dict(MANO_3D, pot_kind=0, cap_kind=0, nelem=(3, 4, 3))
    o = Oracle(desc)
    x = _rand(o.N, 5)
    b = o.apply(1.0, 0.05j, np.zeros(3), x)
    z = o.fd_apply(1.0, 0.05j, b)
    assert np.abs(z - x).max() / np.abs(x).max() < 1e-10
    # harmonic potential is separable: the PC inverts M + g(K + V) exactly (walls: identity)
    o1 = Oracle(TINY_1D)
    x1 = _rand(o1.N, 6)
    b1 = o1.apply(1.0, 0.05j, np.zeros(3), x1)
    z1 = o1.fd_apply(1.0, 0.05j, b1)
    assert np.abs(z1 - x1).max() / np.abs(x1).max() < 1e-10

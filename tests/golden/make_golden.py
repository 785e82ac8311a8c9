"""Generate the golden fixtures of tests/golden/ from independent textbook routines.

The reference ships no stored arrays (its tests/inputs/expected/ does not
exist), so the fixtures come from scipy/numpy, not from the oracle:

  gauss_Q.npz    numpy.polynomial.legendre.leggauss nodes and weights
  basis_*.npz    scipy.interpolate.BSpline values and first derivatives of every
                 basis function at the Gauss points of every element, and the
                 1D Gram bands M, X, S, G computed from them with a 24-point rule
  boehm_*.npz    prolongation by collocation: P = B_fine(g)^-1 B_coarse(g) at the
                 fine Greville points (exact for nested spline spaces)

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np
from scipy.interpolate import BSpline

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = [  # (name, E, lo, hi, p, Q)
    ("basis_e7_p3_q5", 7, -5.0, 5.0, 3, 5),
    ("basis_e10_p4_q8", 10, -10.0, 10.0, 4, 8),
    ("basis_e6_p5_q8", 6, -3.0, 3.0, 5, 8),
    ("basis_e5_p7_q10", 5, -20.0, 20.0, 7, 10),
]


def open_uniform(E, lo, hi, p):
    return np.concatenate([[lo] * (p + 1), lo + (hi - lo) * np.arange(1, E) / E, [hi] * (p + 1)])


def design(t, p, x, nu=0):
    n = len(t) - p - 1
    out = np.zeros((len(x), n))
    for i in range(n):
        c = np.zeros(n)
        c[i] = 1.0
        s = BSpline(t, c, p, extrapolate=False)
        v = s(x, nu)
        out[:, i] = np.nan_to_num(v)
    return out


def basis_case(E, lo, hi, p, Q):
    t = open_uniform(E, lo, hi, p)
    n = E + p
    gx, gw = np.polynomial.legendre.leggauss(Q)
    pts, wts, N0, N1 = [], [], np.zeros((E, p + 1, Q)), np.zeros((E, p + 1, Q))
    for e in range(E):
        u0, u1 = t[p + e], t[p + e + 1]
        x = 0.5 * (u0 + u1) + 0.5 * (u1 - u0) * gx
        pts.append(x)
        wts.append(gw * 0.5 * (u1 - u0))
        B = design(t, p, x)
        D = design(t, p, x, 1)
        N0[e] = B[:, e:e + p + 1].T
        N1[e] = D[:, e:e + p + 1].T
    # bands with a 24-point rule per element (exact for these polynomial integrands)
    hx, hw = np.polynomial.legendre.leggauss(24)
    bw = 2 * p + 1
    M, X, S, G = (np.zeros((n, bw)) for _ in range(4))
    for e in range(E):
        u0, u1 = t[p + e], t[p + e + 1]
        x = 0.5 * (u0 + u1) + 0.5 * (u1 - u0) * hx
        w = hw * 0.5 * (u1 - u0)
        B = design(t, p, x)
        D = design(t, p, x, 1)
        for a in range(e, e + p + 1):
            for b in range(e, e + p + 1):
                M[a, b - a + p] += np.sum(w * B[:, a] * B[:, b])
                X[a, b - a + p] += np.sum(w * x * B[:, a] * B[:, b])
                S[a, b - a + p] += np.sum(w * D[:, a] * D[:, b])
                G[a, b - a + p] += np.sum(w * B[:, a] * D[:, b])
    return dict(knots=t, pts=np.concatenate(pts), wts=np.concatenate(wts), N0=N0, N1=N1, M=M, X=X, S=S, G=G)


def greville(t, p):
    n = len(t) - p - 1
    return np.array([t[i + 1:i + p + 1].mean() for i in range(n)])


def boehm_case(Ef, lo, hi, p):
    tf = open_uniform(Ef, lo, hi, p)
    tc = open_uniform(Ef // 2, lo, hi, p)
    g = greville(tf, p)
    g[-1] = hi - 1e-12 * (hi - lo)   # right end: stay inside the last span
    Bf = design(tf, p, g)
    Bc = design(tc, p, g)
    P = np.linalg.solve(Bf, Bc)
    return dict(fine=tf, coarse=tc, P=P)


def main():
    for Q in (5, 8, 10):
        x, w = np.polynomial.legendre.leggauss(Q)
        np.savez(os.path.join(HERE, f"gauss_{Q}.npz"), x=x, w=w)
    for name, E, lo, hi, p, Q in CASES:
        np.savez(os.path.join(HERE, name + ".npz"), **basis_case(E, lo, hi, p, Q))
    for Ef, p in ((8, 3), (12, 4), (16, 5)):
        np.savez(os.path.join(HERE, f"boehm_e{Ef}_p{p}.npz"), **boehm_case(Ef, -4.0, 4.0, p))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()

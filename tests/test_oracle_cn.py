"""Pins the oracle's Crank-Nicolson / Pade stepping and FGMRES:
  * dense brute-force CN (M + i dt/2 H)^-1 (M - i dt/2 H) on a tiny 3D grid;
  * the closed-form CN coherent state of the 1D oscillator (U = Cayley(H));
  * the driven oscillator of the reference test (tests/test_analytic.py:224-273):
    <x>(t) = -E0/(1-w^2)(cos wt - cos t), error < 2e-3 at dt = 0.01, order 2;
  * norm conservation without CAP, monotone norm decay with CAP
    (tests/test_tdsez.py:272-288);
  * Pade(2,2) is fourth order (src/propagator.cpp:564-676)."""
import math

import numpy as np
import pytest

from oracle import Oracle

TINY_3D = dict(dim=3, nelem=(3, 3, 4), lo=(-5.0, -5.0, -6.0), hi=(5.0, 5.0, 6.0), p=3, nquad=5, mass=1.0,
               charge=-1.0, pot_kind=2, pot=(1.0, 2.0), cap_kind=1, cap=(0.05, 2.0))


def test_cn_equals_dense_cayley():
    o = Oracle(TINY_3D)
    psi0 = o.project_gaussian((0.3, -0.2, 0.1), (1.0, 1.2, 0.9), (0.2, 0.0, 0.1))
    dt, sE = 0.05, np.array([0.0, 0.0, 0.04])
    H = o.dense(1) + o.dense(2) + o.dense(3) + sE[2] * o.dense(6)
    M = o.dense(0)
    ref = np.linalg.solve(M + 0.5j * dt * H, (M - 0.5j * dt * H) @ psi0)
    got, its, res = o.cn_steps(psi0, dt, sE.reshape(1, 3), rtol=1e-14)
    # error in L2 (sqrt(e^H M e)): the coefficients are fixed only to cond(M) times the
    # solver tolerance, and the oracle's threaded reductions move the last iterate
    d = got - ref
    assert np.sqrt(abs(np.vdot(d, M @ d)) / abs(np.vdot(ref, M @ ref))) < 1e-12
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-11
    assert res[0] <= 1e-13


def _ho_problem(E=400):
    return dict(dim=1, nelem=(E,), lo=(-20.0,), hi=(20.0,), p=5, nquad=8, dirichlet=1, mass=1.0, charge=-1.0,
                pot_kind=1, pot=(1.0,))


def test_cn_oscillator_closed_form():
    """Shifted ground state x0 = 2 is a coherent state, alpha = sqrt(2). CN is the
    Cayley transform of H, so each eigenstate E_n = n + 1/2 turns by
    theta_n = 2 arctan(E_n dt / 2) per step: <x>_k = sqrt(2) Re sum_n c_n c_{n+1}
    sqrt(n+1) exp(-i k (theta_{n+1} - theta_n))."""
    o = Oracle(_ho_problem())
    psi = o.project_gaussian((2.0,), (1 / math.sqrt(2),))
    dt, nsteps = 0.01, 300
    out, _, _ = o.cn_steps(psi, dt, np.zeros((nsteps, 3)), rtol=1e-14)
    alpha = math.sqrt(2.0)
    nmax = 60
    c = np.array([math.exp(-alpha ** 2 / 2) * alpha ** n / math.sqrt(math.factorial(n)) for n in range(nmax)])
    th = 2 * np.arctan((np.arange(nmax) + 0.5) * dt / 2)
    k = nsteps
    ref = math.sqrt(2) * np.real(np.sum(c[:-1] * c[1:] * np.sqrt(np.arange(1, nmax)) *
                                        np.exp(-1j * k * (th[1:] - th[:-1]))))
    obs = o.observables(out)
    assert abs(obs[0] - 1.0) < 1e-11            # norm
    assert abs(obs[1] - 2.5) < 1e-9             # <H> = |alpha|^2 + 1/2
    assert abs(obs[3] - ref) < 1e-8, (obs[3], ref)


def _driven_error(dt, T=20.0):
    desc = dict(dim=1, nelem=(193,), lo=(-20.0,), hi=(20.0,), p=7, nquad=10, mass=1.0, charge=-1.0,
                pot_kind=1, pot=(1.0,))
    o = Oracle(desc)
    psi = o.project_gaussian((0.0,), (1 / math.sqrt(2),))
    n = int(round(T / dt))
    tm = (np.arange(n) + 0.5) * dt
    sE = np.zeros((n, 3))
    sE[:, 0] = 0.5 * np.cos(0.6 * tm)     # s E with s = +1 (electron)
    err, nrm = 0.0, 0.0
    x = psi
    for s in range(0, n, 100):
        x, _, _ = o.cn_steps(x, dt, sE[s:s + 100], rtol=1e-13)
        t = (s + 100) * dt
        obs = o.observables(x)
        ref = -0.5 / (1 - 0.36) * (math.cos(0.6 * t) - math.cos(t))
        err = max(err, abs(obs[3] - ref))
        nrm = max(nrm, abs(obs[0] - 1.0))
    return err, nrm


def test_driven_oscillator_accuracy_and_order():
    e1, n1 = _driven_error(0.01)
    e2, n2 = _driven_error(0.005)
    assert e1 < 2e-3, e1
    assert abs(math.log2(e1 / e2) - 2.0) < 0.3
    assert max(n1, n2) < 1e-9


def test_cap_absorbs_monotonically():
    desc = dict(dim=1, nelem=(120,), lo=(-15.0,), hi=(15.0,), p=4, nquad=8, mass=1.0, charge=-1.0,
                cap_kind=1, cap=(0.05, 5.0))
    o = Oracle(desc)
    x = o.project_gaussian((0.0,), (1.0,), (2.0,))
    norms = []
    for _ in range(20):
        x, _, _ = o.cn_steps(x, 0.05, np.zeros((10, 3)), rtol=1e-13)
        norms.append(o.observables(x)[0])
    assert all(b <= a + 1e-13 for a, b in zip(norms, norms[1:]))
    assert norms[-1] < 0.9


def test_pade_fourth_order():
    desc = dict(dim=1, nelem=(60,), lo=(-12.0,), hi=(12.0,), p=5, nquad=8, mass=1.0, charge=-1.0,
                pot_kind=1, pot=(1.0,))
    o = Oracle(desc)
    psi = o.project_gaussian((1.5,), (1 / math.sqrt(2),))
    H = o.dense(1) + o.dense(2)
    M = o.dense(0)
    from scipy.linalg import expm
    T = 1.0
    ref = expm(-1j * T * np.linalg.solve(M, H)) @ psi
    errs = []
    for h in (0.1, 0.05):
        x = psi
        for _ in range(int(round(T / h))):
            x, _ = o.pade22_step(x, h, np.zeros(3), rtol=1e-14)
        errs.append(np.linalg.norm(x - ref))
    assert abs(math.log2(errs[0] / errs[1]) - 4.0) < 0.3, errs

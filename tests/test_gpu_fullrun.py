"""Full-run parity on the GPU through the C ABI: the 1D harmonic-oscillator run
(2000 elements, p = 5, Dirichlet walls, CN dt = 0.01, 1000 steps) against the CPU
oracle and the closed form of the coherent state, and a reduced 3D soft-core run
over a whole sin^2 pulse against the oracle. Wavefunction error is the relative L2
norm sqrt(d^H M d) / sqrt(psi^H M psi); traces are norm, <H> and dipole."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HO_FULL = dict(dim=1, nelem=(2000,), lo=(-20.0,), hi=(20.0,), p=5, nquad=8, dirichlet=1, mass=1.0, charge=-1.0,
               pot_kind=1, pot=(1.0,))


def _l2(o, d):
    return math.sqrt(abs(np.vdot(d, o.term(0, d))))


def _coherent_x(k, dt, alpha=math.sqrt(2.0), nmax=60):
    """<x> after k CN steps of the shifted ground state x0 = sqrt(2) alpha: CN turns each
    eigenstate E_n = n + 1/2 by theta_n = 2 arctan(E_n dt / 2) per step."""
    c = np.array([math.exp(-alpha ** 2 / 2) * alpha ** n / math.sqrt(math.factorial(n)) for n in range(nmax)])
    th = 2 * np.arctan((np.arange(nmax) + 0.5) * dt / 2)
    return math.sqrt(2) * np.real(np.sum(c[:-1] * c[1:] * np.sqrt(np.arange(1, nmax)) *
                                         np.exp(-1j * k * (th[1:] - th[:-1]))))


def test_oscillator_full_run():
    import tdsez_amd as T
    from oracle import Oracle
    o, ctx = Oracle(HO_FULL), T.Context(HO_FULL)
    dt, nsteps, chunk = 0.01, 1000, 100
    psi0 = o.project_gaussian((2.0,), (1 / math.sqrt(2),))
    ctx.pc_setup(T.PC_SEPARABLE)
    d = ctx.vec(psi0)
    its, obs = ctx.evolve_cn(d, 0.0, dt, nsteps, None, rtol=1e-13, observe=True)
    x = psi0
    for s in range(0, nsteps, chunk):
        # the oracle solves without its separable preconditioner here: its Jacobi
        # eigensolver is O(n^3) per set-up at n = 2005, while plain FGMRES takes ~20 its
        x, _, _ = o.cn_steps(x, dt, np.zeros((chunk, 3)), rtol=1e-13, pc=0, restart=100, maxit=5000)
        k = s + chunk
        print(f"oscillator step {k}: <x> = {obs[k - 1, 4]:.12f}", flush=True)
        ref_obs = o.observables(x)
        got = obs[k - 1, 1:]
        assert abs(obs[k - 1, 0] - k * dt) < 1e-12
        # traces against the oracle: norm, Re<H>, Im<H>, dipole
        assert abs(got[0] - ref_obs[0]) < 1e-10, (k, got[0], ref_obs[0])
        assert abs(got[1] - ref_obs[1]) < 1e-10, (k, got[1], ref_obs[1])
        assert abs(got[3] - ref_obs[3]) < 1e-10, (k, got[3], ref_obs[3])
        # and against the closed form (spatial error of p = 5, h = 0.02 is far below)
        assert abs(got[0] - 1.0) < 1e-10
        assert abs(got[1] - 2.5) < 1e-9
        assert abs(got[3] - _coherent_x(k, dt)) < 1e-8, (k, got[3], _coherent_x(k, dt))
    err = _l2(o, d.download() - x) / _l2(o, x)
    assert err < 1e-10, err
    assert its.max() <= 3
    ctx.close()


PULSE_3D = dict(dim=3, nelem=(12, 12, 12), lo=(-12.0,) * 3, hi=(12.0,) * 3, p=4, nquad=8, mass=1.0, charge=-1.0,
                pot_kind=2, pot=(1.0, 2.0), cap_kind=1, cap=(0.01, 8.0), mg_levels=2)


@pytest.mark.parametrize("pc", [2, 4, 7])
def test_softcore_whole_pulse(pc):
    """The 3D soft-core atom with its CAP over a whole 2-cycle sin^2 pulse (251 steps of
    dt = 0.05) from the ground state: plain FP64 multigrid (pc 2) and the learned
    mixed-precision preconditioner (pc 7, trained on the run's first steps), and the
    complex64 multigrid (pc 4) against
    the oracle's explicit CN steps."""
    import tdsez_amd as T
    from oracle import Oracle
    o, ctx = Oracle(PULSE_3D), T.Context(PULSE_3D)
    pulse = T.make_pulse(E0=(0.0, 0.0, 0.05), omega=1.0, ncycles=2.0)
    dt = 0.05
    nsteps = int(round(2 * 2 * math.pi / 1.0 / dt))
    d = ctx.vec()
    ctx.project_gaussian(d, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    ctx.pc_setup(T.PC_MG)
    ctx.ground_state(d, sigma=-0.5, tol=1e-10)
    psi0 = d.download()
    if pc == T.PC_LEARNED:
        ctx.learn_train(d, 0.0, dt, pulse, harvest_steps=2, max_samples=6, iters=100, rtol=1e-13)
    ctx.pc_setup(pc)
    d = ctx.vec(psi0)
    its, obs = ctx.evolve_cn(d, 0.0, dt, nsteps, pulse, rtol=1e-13, observe=True)
    sE = np.array([ctx.coupling(pulse, (s + 0.5) * dt) for s in range(nsteps)])
    x = psi0
    chunk = 50
    for s in range(0, nsteps, chunk):
        m = min(chunk, nsteps - s)
        x, _, _ = o.cn_steps(x, dt, sE[s:s + m], rtol=1e-13)
        print(f"pc {pc} step {s + m}: norm {obs[s + m - 1, 1]:.12f}", flush=True)
        ref = o.observables(x)
        got = obs[s + m - 1, 1:]
        for col in (0, 1, 2, 5):   # norm, Re<H>, Im<H>, z-dipole
            assert abs(got[col] - ref[col]) < 1e-10, (s + m, col, got[col], ref[col])
    err = _l2(o, d.download() - x) / _l2(o, x)
    assert err < 1e-10, err
    assert obs[-1, 1] < 1.0   # the CAP has absorbed something
    ctx.close()

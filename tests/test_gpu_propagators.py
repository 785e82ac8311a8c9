"""GPU parity of the rational propagators, the initial-state projection and the
ground-state solver, through the C ABI, against the CPU oracle and closed forms."""
import math

import numpy as np
import pytest

from tests.problems import HO_1D, SOFTCORE_3D

pytestmark = pytest.mark.gpu

TINY = dict(SOFTCORE_3D, nelem=(4, 4, 5), lo=(-6.0, -6.0, -6.0), hi=(6.0, 6.0, 6.0))


def _pair(desc):
    import tdsez_amd as T
    from oracle import Oracle
    return Oracle(desc), T.Context(desc)


def test_projection_matches_oracle():
    o, ctx = _pair(SOFTCORE_3D)
    ref = o.project_gaussian((0.5, -0.3, 0.2), (1.2, 1.0, 1.1), (0.3, 0.0, -0.2))
    d = ctx.vec()
    ctx.project_gaussian(d, (0.5, -0.3, 0.2), (1.2, 1.0, 1.1), (0.3, 0.0, -0.2))
    got = d.download()
    # the oracle solves M c = b by Jacobi-CG to a 1e-13 residual as the reference does
    # (src/core.cpp:1224-1230); cond(M) ~ 1e3 in 3D turns that into ~1e-10 in c, while
    # the GPU applies M^-1 exactly by fast diagonalisation
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-9
    assert abs(ctx.observables(d)[0] - 1.0) < 1e-12
    ctx.close()


def test_pade22_step_matches_oracle():
    import tdsez_amd as T
    o, ctx = _pair(SOFTCORE_3D)
    psi0 = o.project_gaussian((0.3, 0.0, -0.2), (1.0, 1.1, 0.9))
    h, sE = 0.05, np.array([0.0, 0.0, 0.03])
    ref, _ = o.pade22_step(psi0, h, sE, rtol=1e-13)
    ctx.pc_setup(T.PC_SEPARABLE)
    d = ctx.vec(psi0)
    ctx.step_pade22(d, h, sE, rtol=1e-13)
    err = np.linalg.norm(d.download() - ref) / np.linalg.norm(ref)
    assert err < 1e-11, err
    ctx.close()


@pytest.mark.parametrize("pc", [1, 2])
def test_magnus4_matches_oracle(pc):
    import tdsez_amd as T
    o, ctx = _pair(dict(SOFTCORE_3D, nelem=(10, 10, 12)))
    psi = o.project_gaussian((0.3, 0.0, -0.2), (1.0, 1.1, 0.9))
    dt, nsteps = 0.05, 3
    pulse = T.make_pulse(E0=(0.0, 0.0, 0.05), omega=0.3, ncycles=1.0)
    s3 = math.sqrt(3)
    c1, c2 = 0.5 - s3 / 6, 0.5 + s3 / 6
    a1, a2 = (3 - 2 * s3) / 12, (3 + 2 * s3) / 12
    ref = psi
    for s in range(nsteps):
        t = s * dt
        E1, E2 = ctx.coupling(pulse, t + c1 * dt), ctx.coupling(pulse, t + c2 * dt)
        ref, _ = o.pade22_step(ref, 0.5 * dt, 2 * (a2 * E1 + a1 * E2), rtol=1e-13)
        ref, _ = o.pade22_step(ref, 0.5 * dt, 2 * (a1 * E1 + a2 * E2), rtol=1e-13)
    if pc == T.PC_MG and ctx.levels < 2:
        pytest.skip("grid cannot be coarsened")
    ctx.pc_setup(pc)
    d = ctx.vec(psi)
    its, _ = ctx.evolve_rational(d, 0.0, dt, nsteps, pulse, magnus=True, rtol=1e-13)
    err = np.linalg.norm(d.download() - ref) / np.linalg.norm(ref)
    assert err < 1e-11, err
    assert its.max() < 80
    ctx.close()


def test_ground_state_tiny_against_dense_eigensolve():
    from scipy.linalg import eigh
    desc = dict(TINY, cap_kind=0)
    o, ctx = _pair(desc)
    H = (o.dense(1) + o.dense(2)).real
    M = o.dense(0).real
    E0 = eigh(H, M, eigvals_only=True)[0]
    d = ctx.vec()
    ctx.project_gaussian(d, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    E, res, it = ctx.ground_state(d, sigma=E0 - 0.2, tol=1e-9)
    assert abs(E - E0) < 1e-10 * max(1, abs(E0)), (E, E0)
    assert res < 1e-9
    ctx.close()


def test_ground_state_oscillator():
    import tdsez_amd as T
    _, ctx = _pair(HO_1D)
    d = ctx.vec()
    ctx.project_gaussian(d, (0.3,), (0.9,))
    E, res, it = ctx.ground_state(d, sigma=0.3, tol=1e-10)
    assert abs(E - 0.5) < 1e-9, E
    # the CAP is back in H after the eigensolve: the context still propagates
    ctx.pc_setup(T.PC_SEPARABLE)
    st = ctx.step_cn(d, 0.01, (0, 0, 0), rtol=1e-12)
    obs = ctx.observables(d)
    assert abs(obs[0] - 1.0) < 1e-10 and abs(obs[1] - 0.5) < 1e-8
    ctx.close()


def test_ifunction_matches_oracle():
    """F = i M psidot - H psi - sE.D psi (TDSEZIFunctionPol*, src/monitor.cpp:108-162), and the
    shift of the Jacobian (tdz_set_shift) is left as it was."""
    o, ctx = _pair(SOFTCORE_3D)
    rng = np.random.default_rng(21)
    psi = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
    psidot = rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)
    sE = np.array([0.01, -0.02, 0.05])
    ref = 1j * o.term(0, psidot) + o.apply(0.0, -1.0, sE, psi)
    dp, dd, dF = ctx.vec(psi), ctx.vec(psidot), ctx.vec()
    ctx.set_shift(1.0, 0.01j, (0.0, 0.0, 0.02))
    ctx.ifunction(sE, dp, dd, dF)
    F = dF.download()
    assert np.abs(F - ref).max() / np.abs(ref).max() < 1e-13
    # the Jacobian shift survives: A psi = M psi + 0.01i (H + 0.02 D_z) psi
    dy = ctx.vec()
    ctx.apply(dp, dy)
    ref2 = o.apply(1.0, 0.01j, np.array([0.0, 0.0, 0.02]), psi)
    assert np.abs(dy.download() - ref2).max() / np.abs(ref2).max() < 1e-13
    # TSTHETA's residual at the CN solution: F(t_mid, (x0+x1)/2, (x1-x0)/dt) = 0
    dt, sEm = 0.02, np.array([0.0, 0.0, 0.03])
    x0 = o.project_gaussian((0.3, 0.0, -0.2), (1.0, 1.1, 0.9))
    x1, _, _ = o.cn_steps(x0, dt, sEm.reshape(1, 3), rtol=1e-14)
    ctx.ifunction(sEm, ctx.vec(0.5 * (x0 + x1)), ctx.vec((x1 - x0) / dt), dF)
    assert np.linalg.norm(dF.download()) / np.linalg.norm(o.term(0, x0)) < 1e-9
    ctx.close()


@pytest.mark.parametrize("order,extrapolate", [(1, True), (3, True), (5, True), (3, False)])
def test_pade_guess_options_keep_the_answer(order, extrapolate):
    """-pade_guess_order / -pade_extrapolate_guess change the initial guess of the stage solves,
    not their answer (src/propagator.cpp:620-646, :723-726)."""
    import tdsez_amd as T
    o, ctx = _pair(TINY)
    psi = o.project_gaussian((0.3, 0.0, -0.2), (1.0, 1.1, 0.9))
    h, nsteps = 0.05, 6
    ref = psi
    for s in range(nsteps):
        sE = np.array([0.0, 0.0, 0.04 * math.sin(0.2 * (s + 0.5) * h)])
        ref, _ = o.pade22_step(ref, h, sE, rtol=1e-14)
    ctx.pc_setup(T.PC_SEPARABLE)
    ctx.set_pade_guess(extrapolate, order)
    d = ctx.vec(psi)
    for s in range(nsteps):
        ctx.step_pade22(d, h, (0.0, 0.0, 0.04 * math.sin(0.2 * (s + 0.5) * h)), rtol=1e-14)
    cnt, hist = ctx.pade_history()
    assert hist.shape == (4, order, ctx.N)
    assert list(cnt[:2]) == ([min(order, nsteps)] * 2 if extrapolate else [0, 0])
    diff = d.download() - ref
    err = np.sqrt(abs(np.vdot(diff, o.term(0, diff))) / abs(np.vdot(ref, o.term(0, ref))))
    assert err < 1e-12, err
    ctx.close()

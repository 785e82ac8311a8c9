"""CN predictor (extrapolated initial guess) through the C ABI.

The predictor changes only where each CN solve starts: the solves are the same
linear systems to the same tolerance. So the states must equal the oracle's CN steps
(explicit right-hand sides, the current state as guess) in L2 to the level the
solver tolerance fixes them, on every problem family (3D soft-core with CAP, uneven
orders, 2D, 1D with Dirichlet walls), under a field that changes every step along
all axes; the history must carry over between evolve calls that continue a run and
be dropped when the state was changed in between.
"""
import numpy as np
import pytest

from tests.problems import HO_1D, MANO_3D, SOFTCORE_2D, SOFTCORE_3D

pytestmark = pytest.mark.gpu

PROBLEMS = {"softcore3d": SOFTCORE_3D, "mano3d": MANO_3D, "softcore2d": SOFTCORE_2D, "ho1d": HO_1D}


def _l2(o, x):
    return np.sqrt(abs(np.vdot(x, o.term(0, x))))


@pytest.fixture(scope="module", params=list(PROBLEMS))
def pair(request):
    import tdsez_amd as T
    from oracle import Oracle
    desc = PROBLEMS[request.param]
    o = Oracle(desc)
    ctx = T.Context(desc)
    yield request.param, desc, o, ctx
    ctx.close()


def _setup(desc, o, ctx, nsteps, dt):
    import tdsez_amd as T
    dim = desc["dim"]
    psi0 = o.project_gaussian(centre=(0.5, -0.3, 0.2)[:dim], width=(1.2, 1.0, 1.1)[:dim],
                              momentum=(0.3, 0.0, -0.2)[:dim])
    pulse = T.make_pulse(E0=(0.02, -0.01, 0.05), omega=0.3, ncycles=1.0)
    sE = np.array([ctx.coupling(pulse, (s + 0.5) * dt) for s in range(nsteps)])
    sE[:, dim:] = 0.0
    return psi0, pulse, sE


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_predicted_steps_match_oracle(pair, order):
    import tdsez_amd as T
    name, desc, o, ctx = pair
    dt, nsteps = 0.02, 8
    psi0, pulse, sE = _setup(desc, o, ctx, nsteps, dt)
    ref, _, _ = o.cn_steps(psi0, dt, sE, rtol=1e-13)
    ctx.pc_setup(T.PC_SEPARABLE)
    ctx.set_cn_predictor(order)
    try:
        d = ctx.vec(psi0)
        its, _ = ctx.evolve_cn(d, 0.0, dt, nsteps, pulse, rtol=1e-13)
        # each of the 8 solves meets rtol 1e-13 on its residual; the state error it leaves
        # is that times the condition of M + i dt/2 H (a few here), summed over the steps
        err = _l2(o, d.download() - ref) / _l2(o, ref)
        assert err < 1e-11, (name, order, err, its)
        # and the guess pays: no more iterations than starting from the current state (orders
        # 1 and 2; the higher ones amplify the rough error of the Manolopoulos CAP problem,
        # whose separable preconditioner needs 35 iterations, by an iteration or two)
        if order <= 2:
            d0 = ctx.vec(psi0)
            ctx.set_cn_predictor(0)
            its0, _ = ctx.evolve_cn(d0, 0.0, dt, nsteps, pulse, rtol=1e-13)
            assert its[2:].sum() <= its0[2:].sum(), (name, its, its0)
    finally:
        ctx.set_cn_predictor(0)


def test_predicted_steps_z_field_match_oracle(pair):
    """A field along z alone: the mass and z-dipole products of each state come from one
    fused pass (3D problems)."""
    import tdsez_amd as T
    name, desc, o, ctx = pair
    dim = desc["dim"]
    dt, nsteps = 0.02, 8
    psi0 = o.project_gaussian(centre=(0.5, -0.3, 0.2)[:dim], width=(1.2, 1.0, 1.1)[:dim],
                              momentum=(0.3, 0.0, -0.2)[:dim])
    pulse = T.make_pulse(E0=(0.0, 0.0, 0.08), omega=0.3, ncycles=1.0)
    sE = np.array([ctx.coupling(pulse, (s + 0.5) * dt) for s in range(nsteps)])
    sE[:, dim:] = 0.0
    ref, _, _ = o.cn_steps(psi0, dt, sE, rtol=1e-13)
    ctx.pc_setup(T.PC_SEPARABLE)
    ctx.set_cn_predictor(2)
    try:
        d = ctx.vec(psi0)
        its, _ = ctx.evolve_cn(d, 0.0, dt, nsteps, pulse, rtol=1e-13)
        err = _l2(o, d.download() - ref) / _l2(o, ref)
        assert err < 1e-11, (name, err, its)
    finally:
        ctx.set_cn_predictor(0)


def test_history_carries_over_and_resets(pair):
    import tdsez_amd as T
    name, desc, o, ctx = pair
    dt = 0.02
    psi0, pulse, _ = _setup(desc, o, ctx, 8, dt)
    ctx.pc_setup(T.PC_SEPARABLE)
    ctx.set_cn_predictor(2)
    try:
        a = ctx.vec(psi0)
        its_a, _ = ctx.evolve_cn(a, 0.0, dt, 8, pulse, rtol=1e-13)
        b = ctx.vec(psi0)
        ctx.evolve_cn(b, 0.0, dt, 4, pulse, rtol=1e-13)
        its_b, _ = ctx.evolve_cn(b, 4 * dt, dt, 4, pulse, rtol=1e-13)   # continues the run
        # the same extrapolated guesses (a restarted history would start from the state
        # and take more iterations); the midpoint times agree only to rounding
        assert np.array_equal(its_a[4:], its_b), (name, its_a, its_b)
        err = _l2(o, a.download() - b.download()) / _l2(o, a.download())
        assert err < 1e-11, (name, err)   # both solves meet rtol 1e-13
        # a state changed between calls: the history must not be used
        c = ctx.vec(psi0)
        ctx.evolve_cn(c, 0.0, dt, 4, pulse, rtol=1e-13)
        x = c.download()
        c.upload(x * np.exp(0.3j))
        ctx.evolve_cn(c, 4 * dt, dt, 4, pulse, rtol=1e-13)
        ref = a.download() * np.exp(0.3j)
        err = _l2(o, c.download() - ref) / _l2(o, ref)
        assert err < 1e-11, (name, err)
    finally:
        ctx.set_cn_predictor(0)


def test_predictor_order_is_checked():
    import tdsez_amd as T
    ctx = T.Context(SOFTCORE_2D)
    try:
        with pytest.raises(RuntimeError):
            ctx.set_cn_predictor(5)
        with pytest.raises(RuntimeError):
            ctx.set_cn_predictor(-1)
    finally:
        ctx.close()


@pytest.mark.parametrize("order", [2, 3])
def test_long_run_stays_on_the_cn_trajectory(order):
    """The stored products are carried from solve to solve; their rounding would grow like
    n^(order+1/2) through the extrapolation (delta_new = sum_j w_j delta_j). With the
    periodic exact recomputation 400 predicted steps stay on the oracle's CN trajectory."""
    import tdsez_amd as T
    from oracle import Oracle
    desc = dict(SOFTCORE_3D, nelem=(5, 5, 6))
    o, ctx = Oracle(desc), T.Context(desc)
    try:
        dt, nsteps = 0.02, 400
        psi0 = o.project_gaussian(centre=(0.5, -0.3, 0.2), width=(1.2, 1.0, 1.1), momentum=(0.3, 0.0, -0.2))
        pulse = T.make_pulse(E0=(0.0, 0.0, 0.08), omega=0.3, ncycles=2.0)
        sE = np.array([ctx.coupling(pulse, (s + 0.5) * dt) for s in range(nsteps)])
        ref, _, _ = o.cn_steps(psi0, dt, sE, rtol=1e-14)
        ctx.pc_setup(T.PC_SEPARABLE)
        ctx.set_cn_predictor(order)
        d = ctx.vec(psi0)
        its, _ = ctx.evolve_cn(d, 0.0, dt, nsteps, pulse, rtol=1e-12)
        err = _l2(o, d.download() - ref) / _l2(o, ref)
        every, drift = ctx.predictor_status()
        print(f"order {order}: error {err:.2e}, resync every {every} steps, last drift {drift:.1e}")
        assert err < 1e-10, (order, err, every, drift)
        assert drift < 1e-14, drift   # within 1e-2 of the tolerance
    finally:
        ctx.close()

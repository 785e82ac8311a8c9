"""Layered structures on the GPU (BenDaniel-Duke kinetic term with a piecewise-constant
mass m(z), band offsets, interface-aligned C0 knots) against the oracle: operator terms,
the z-line separable solve (exact for the separable layered operator with a field
along z), multigrid, CN and Magnus-4 steps."""
import math

import numpy as np
import pytest

from tests.test_oracle_layers import layered, layered_knots

pytestmark = pytest.mark.gpu

LAY = dict(layered(p=4, bounds=(-8.0, -3.0, 2.0, 8.0), per=(3, 4, 3), masses=(0.067, 0.092, 0.067),
                   v=(0.0, 0.00735, 0.0)),
           nelem=(6, 5, 1), lo=(-10.0, -9.0, 0.0), hi=(10.0, 9.0, 1.0), nquad=8, mg_levels=2)


def _pair(desc=LAY):
    import tdsez_amd as T
    from oracle import Oracle
    return Oracle(desc), T.Context(desc)


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _l2(o, d):
    return math.sqrt(abs(np.vdot(d, o.term(0, d))))


def test_layered_terms_and_operator():
    import tdsez_amd as T
    o, ctx = _pair()
    assert ctx.n == o.n
    x = _rand(o.N, 1)
    for term in (T.TERM_M, T.TERM_K, T.TERM_V, T.TERM_DX, T.TERM_DY, T.TERM_DZ):
        dx, dy = ctx.vec(x), ctx.vec()
        ctx.term(term, dx, dy)
        ref = o.term(term, x)
        err = np.abs(dy.download() - ref).max() / np.abs(ref).max()
        assert err < 1e-13, (term, err)
    for alpha, g, sE in ((1.0, 0.01j, (0, 0, 0.03)), (0.3 - 0.2j, 0.7 + 0.1j, (0, 0, -0.05))):
        ctx.set_shift(alpha, g, sE)
        dx, dy = ctx.vec(x), ctx.vec()
        ctx.apply(dx, dy)
        ref = o.apply(alpha, g, np.array(sE, float), x)
        assert np.abs(dy.download() - ref).max() / np.abs(ref).max() < 1e-13
    with pytest.raises(T.TdzError, match="along z"):
        ctx.set_shift(1.0, 0.01j, (0.01, 0.0, 0.0))
    ctx.close()


@pytest.mark.parametrize("pc", [1, 3])
def test_layered_line_solve_is_exact(pc):
    """The separable layered operator (no non-separable potential) with a z field is
    inverted exactly by the z-line solve: one FGMRES iteration."""
    import tdsez_amd as T
    o, ctx = _pair()
    alpha, g, sE = 1.0, 0.05j, np.array([0.0, 0.0, 0.04])
    xs = _rand(o.N, 2)
    b = o.apply(alpha, g, sE, xs)
    ctx.set_shift(alpha, g, sE)
    ctx.pc_setup(pc)
    db, dx = ctx.vec(b), ctx.vec()
    st = ctx.solve(db, dx, rtol=1e-12 if pc == 1 else 1e-6, maxit=50)
    res = np.linalg.norm(b - o.apply(alpha, g, sE, dx.download())) / np.linalg.norm(b)
    assert st.its == 1, st.its
    assert res < (1e-12 if pc == 1 else 1e-5), res
    ctx.close()


def test_layered_multigrid():
    import tdsez_amd as T
    o, ctx = _pair()
    assert ctx.levels >= 2
    alpha, g, sE = 1.0, 0.05j, np.array([0.0, 0.0, 0.04])
    xs = _rand(o.N, 3)
    b = o.apply(alpha, g, sE, xs)
    ctx.set_shift(alpha, g, sE)
    ctx.pc_setup(T.PC_MG)
    db, dx = ctx.vec(b), ctx.vec()
    st = ctx.solve(db, dx, rtol=1e-12, maxit=100)
    res = np.linalg.norm(b - o.apply(alpha, g, sE, dx.download())) / np.linalg.norm(b)
    assert res < 1e-11 and st.its <= 3, (res, st.its)
    ctx.close()


@pytest.mark.parametrize("magnus", [False, True])
def test_layered_steps_match_oracle(magnus):
    import tdsez_amd as T
    o, ctx = _pair()
    psi0 = o.project_gaussian((0.5, -0.3, -4.0), (2.0, 2.0, 1.5), (0.2, 0.0, 0.3))
    dt, nsteps = 0.5, 4
    pulse = T.make_pulse(E0=(0.0, 0.0, 0.01), omega=0.3, ncycles=1.0)
    ctx.pc_setup(T.PC_SEPARABLE)
    d = ctx.vec(psi0)
    if magnus:
        s3 = math.sqrt(3)
        c1, c2 = 0.5 - s3 / 6, 0.5 + s3 / 6
        a1, a2 = (3 - 2 * s3) / 12, (3 + 2 * s3) / 12
        ref = psi0
        for s in range(nsteps):
            t = s * dt
            E1, E2 = ctx.coupling(pulse, t + c1 * dt), ctx.coupling(pulse, t + c2 * dt)
            ref, _ = o.pade22_step(ref, 0.5 * dt, 2 * (a2 * E1 + a1 * E2), rtol=1e-13, restart=100, maxit=5000)
            ref, _ = o.pade22_step(ref, 0.5 * dt, 2 * (a1 * E1 + a2 * E2), rtol=1e-13, restart=100, maxit=5000)
        its, _ = ctx.evolve_rational(d, 0.0, dt, nsteps, pulse, magnus=True, rtol=1e-13)
    else:
        sE = np.array([ctx.coupling(pulse, (s + 0.5) * dt) for s in range(nsteps)])
        # the oracle's separable preconditioner assumes one mass: an inexact but convergent one here
        ref, _, _ = o.cn_steps(psi0, dt, sE, rtol=1e-13, restart=100, maxit=5000)
        its, _ = ctx.evolve_cn(d, 0.0, dt, nsteps, pulse, rtol=1e-13)
    err = _l2(o, d.download() - ref) / _l2(o, ref)
    assert err < 1e-11, err
    # one FGMRES iteration per solve: CN has one solve per step, Magnus-4 four stage solves
    assert its.max() <= (4 if magnus else 2), its
    ctx.close()

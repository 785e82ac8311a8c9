"""The batched 2D scan: tile variants of the 2D potential kernel against the oracle,
and graph-captured batched CN steps against single-run steps and the oracle."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scan_desc(nel, p, L=20.0):
    return dict(dim=2, nelem=nel, lo=(-L, -L), hi=(L, L), p=p, nquad=8, mass=1.0, charge=-1.0, pot_kind=2,
                pot=(1.0, 1.0))


def _l2(o, d):
    return math.sqrt(abs(np.vdot(d, o.term(0, d))))


@pytest.mark.parametrize("p", [3, 4, 5])
@pytest.mark.parametrize("variant", range(5))
def test_pot2d_tile_variants(p, variant, monkeypatch):
    import tdsez_amd as T
    from oracle import Oracle
    desc = _scan_desc((13, 7), p)
    monkeypatch.setenv("TDZ_POT_VARIANT", str(variant))
    o, ctx = Oracle(desc), T.Context(desc)
    rng = np.random.default_rng(p * 10 + variant)
    x = rng.standard_normal(ctx.N) + 1j * rng.standard_normal(ctx.N)
    dx, dy = ctx.vec(x), ctx.vec()
    ctx.term(T.TERM_V, dx, dy)
    ref = o.term(T.TERM_V, x)
    err = np.abs(dy.download() - ref).max() / np.abs(ref).max()
    assert err < 1e-13, (p, variant, err)
    ctx.close()


SCAN = [(_scan_desc((16, 16), 3), dict(E0=(0.05, 0, 0), cep=0.0)),
        (_scan_desc((20, 12), 4, 16.0), dict(E0=(0.1, 0, 0), cep=1.3)),
        (_scan_desc((12, 14), 5), dict(E0=(0.02, 0, 0), cep=4.0))]


@pytest.mark.parametrize("pc,kiters,pred", [(1, 4, 0), (3, 4, 0), (3, 1, 0), (3, 2, 2), (1, 3, 1)])
def test_batch_matches_single_runs_and_oracle(pc, kiters, pred):
    """kiters = 1 leaves most steps to the host completion path; pred > 0 starts each
    captured solve from the extrapolated guess (the same systems, the same tolerance)."""
    import tdsez_amd as T
    from oracle import Oracle
    descs = [d for d, _ in SCAN]
    pulses = [T.make_pulse(E0=kw["E0"], omega=0.5, ncycles=1.0, cep=kw["cep"]) for _, kw in SCAN]
    dt, nsteps = 0.05, 12
    b = T.Batch(descs, pulses, dt, nsteps, pc=pc, kiters=kiters)
    if pred:
        b.set_predictor(pred)
    starts = []
    for r, ctx in enumerate(b.runs):
        ctx.project_gaussian(b.psi[r], (0.5, -0.3), (1.2, 1.0), (0.2, 0.1))
        starts.append(b.psi[r].download())
    its, res = b.evolve(nsteps, rtol=1e-13)
    assert res.max() <= 1e-13
    if kiters == 1:
        assert b.fixups() > 0
    if pred:
        with pytest.raises(RuntimeError):   # the step is captured now
            b.set_predictor(0)
    for r, (desc, pulse) in enumerate(zip(descs, pulses)):
        got = b.psi[r].download()
        ctx = T.Context(desc)
        ctx.pc_setup(pc)
        d = ctx.vec(starts[r])
        ctx.evolve_cn(d, 0.0, dt, nsteps, pulse, rtol=1e-13)
        o = Oracle(desc)
        sE = np.array([ctx.coupling(pulse, (s + 0.5) * dt) for s in range(nsteps)])
        sE[:, 2] = 0.0
        ref, _, _ = o.cn_steps(starts[r], dt, sE, rtol=1e-13)
        assert _l2(o, got - d.download()) / _l2(o, ref) < 1e-11, r
        assert _l2(o, got - ref) / _l2(o, ref) < 1e-11, r
        ctx.close()
    b.close()

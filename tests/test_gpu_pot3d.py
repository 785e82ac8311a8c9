"""Tile variants of the tiled 3D potential kernel (z-marching, FP64) against the oracle.

Every variant the autotuner may pick (4x4 and 2x2 element tiles, one, two or four
Gauss-point columns per thread) must give the oracle's matrix-free potential product to
rounding, for the soft-core and the harmonic potential, with element counts that leave
partial tiles in x and y.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _desc(p, kind):
    return dict(dim=3, nelem=(7, 5, 6), lo=(-6.0, -5.0, -7.0), hi=(6.0, 5.0, 7.0), p=p, nquad=8, mass=1.0,
                charge=-1.0, pot_kind=kind, pot=(1.0, 2.0) if kind == 2 else (0.7,))


@pytest.mark.parametrize("p", [3, 4])
@pytest.mark.parametrize("variant", range(4))
@pytest.mark.parametrize("kind", [1, 2])
def test_pot3d_tile_variants(p, variant, kind, monkeypatch):
    import tdsez_amd as T
    from oracle import Oracle
    desc = _desc(p, kind)
    monkeypatch.setenv("TDZ_POT_VARIANT", str(variant))
    o, ctx = Oracle(desc), T.Context(desc)
    try:
        rng = np.random.default_rng(100 * p + 10 * variant + kind)
        x = rng.standard_normal(ctx.N) + 1j * rng.standard_normal(ctx.N)
        dx, dy = ctx.vec(x), ctx.vec()
        ctx.term(T.TERM_V, dx, dy)
        ref = o.term(T.TERM_V, x)
        err = np.abs(dy.download() - ref).max() / np.abs(ref).max()
        assert err < 1e-13, (p, variant, kind, err)
    finally:
        ctx.close()


@pytest.mark.parametrize("variant", range(4))
def test_pot3d_is_deterministic(variant, monkeypatch):
    """Neighbouring tiles share p planes of functions in x and y; their contributions go to
    scratch planes by offset class and are summed in a fixed order, so repeated applies (and
    the whole operator) are bit-identical."""
    import tdsez_amd as T
    desc = dict(_desc(4, 2), nelem=(24, 20, 10))
    monkeypatch.setenv("TDZ_POT_VARIANT", str(variant))
    ctx = T.Context(desc)
    try:
        rng = np.random.default_rng(7 + variant)
        x = rng.standard_normal(ctx.N) + 1j * rng.standard_normal(ctx.N)
        dx, dy = ctx.vec(x), ctx.vec()
        outs = []
        for _ in range(3):
            ctx.term(T.TERM_V, dx, dy)
            outs.append(dy.download())
        ctx.set_shift(1.0, 0.01j, (0.0, 0.0, 0.05))
        for _ in range(2):
            ctx.apply(dx, dy)
            outs.append(dy.download())
        assert np.array_equal(outs[0].view(np.float64), outs[1].view(np.float64))
        assert np.array_equal(outs[0].view(np.float64), outs[2].view(np.float64))
        assert np.array_equal(outs[3].view(np.float64), outs[4].view(np.float64))
    finally:
        ctx.close()

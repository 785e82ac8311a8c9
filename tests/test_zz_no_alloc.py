"""The allocation-free entry point tdz_observables_noalloc against tdz_observables,
through the C ABI: two calls in a row on one context, then on a smaller one (the static
scratch keeps the last call's contents). Floating point: the two sum in different orders
(device block reductions, host loop), so they agree to 1e-12 relative."""
import numpy as np
import pytest

from tests.problems import HO_1D, SOFTCORE_2D, SOFTCORE_3D

pytestmark = pytest.mark.gpu


def _check(ctx, x):
    d = ctx.vec(x)
    ref = ctx.observables(d)
    got = ctx.observables_noalloc(d)
    scale = max(np.abs(ref).max(), 1e-300)
    assert np.abs(got - ref).max() / scale < 1e-12, (got, ref)
    d.free()
    return got


def test_noalloc_observables_match():
    import tdsez_amd as T
    big = T.Context(dict(SOFTCORE_3D, nelem=(14, 12, 16)))
    small = T.Context(SOFTCORE_3D)
    try:
        for ctx, seed in ((big, 1), (big, 2), (small, 3), (small, 4)):
            d = ctx.vec()
            ctx.project_gaussian(d, (0.4, -0.2, 0.3), (1.1, 1.0, 1.2), (0.2, 0.0, -0.3))
            x = d.download() * np.exp(0.1j * seed)
            x += 1e-3 * np.random.default_rng(seed).standard_normal(ctx.N)
            _check(ctx, x)
    finally:
        big.close()
        small.close()


@pytest.mark.parametrize("desc", [SOFTCORE_2D, HO_1D])
def test_noalloc_observables_lower_dims(desc):
    import tdsez_amd as T
    ctx = T.Context(desc)
    try:
        d = ctx.vec()
        dim = desc["dim"]
        ctx.project_gaussian(d, (0.4, -0.2, 0.3)[:dim], (1.1, 1.0, 1.2)[:dim], (0.2, 0.0, -0.3)[:dim])
        _check(ctx, d.download())
    finally:
        ctx.close()

"""Complex64 fast-diagonalisation solve on long axes (through the C ABI), with the default
split-FP16 transforms and with the FP32 ones.

Axis lengths n = E + p of 65 .. 154 take every register-resident MFMA transform size
(up to 8, 16, 24, 33, 40 k-steps of 4); n = 205 takes the LDS-tiled transform. The
complex64 solve must compute the FP64 solve's map to single-precision accuracy
(relative 1e-4 in the 2-norm), and the FP64 solve must match the oracle where the
oracle's dense eigen-solves finish in seconds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = dict(mass=1.0, charge=-1.0, pot_kind=2, pot=(1.0, 2.0), cap_kind=1, cap=(0.01, 3.0), nquad=6)
SHAPES = {
    "x132": dict(BASE, dim=3, nelem=(128, 3, 4), lo=(-30.0, -4.0, -4.0), hi=(30.0, 4.0, 4.0), p=4),
    "y154": dict(BASE, dim=3, nelem=(3, 150, 2), lo=(-4.0, -40.0, -3.0), hi=(4.0, 40.0, 3.0), p=4),
    "z103": dict(BASE, dim=3, nelem=(4, 3, 100), lo=(-4.0, -4.0, -30.0), hi=(4.0, 4.0, 30.0), p=3),
    "2d65x105": dict(BASE, dim=2, nelem=(60, 100), lo=(-20.0, -30.0), hi=(20.0, 30.0), p=5),
    "1d205": dict(BASE, dim=1, nelem=(200,), lo=(-40.0,), hi=(40.0,), p=5),
    "walls3d": dict(BASE, dim=3, nelem=(20, 14, 30), lo=(-10.0, -8.0, -12.0), hi=(10.0, 8.0, 12.0), p=4,
                    dirichlet=1),
}


@pytest.mark.parametrize("name", list(SHAPES))
def test_complex64_fd_matches_fp64(name):
    import tdsez_amd as T
    desc = SHAPES[name]
    ctx = T.Context(desc)
    try:
        ctx.set_shift(1.0, 0.02j, (0, 0, 0))
        rng = np.random.default_rng(3)
        r = rng.standard_normal(ctx.N) + 1j * rng.standard_normal(ctx.N)
        dr, z64, z32 = ctx.vec(r), ctx.vec(), ctx.vec()
        ctx.pc_setup(T.PC_SEPARABLE)
        ctx.pc_apply(dr, z64)
        ctx.pc_setup(T.PC_SEPARABLE_MIXED)
        ctx.pc_apply(dr, z32)
        a, b = z64.download(), z32.download()
        err = np.linalg.norm(a - b) / np.linalg.norm(a)
        assert err < 1e-4, (name, err)
        if name in ("x132", "2d65x105"):
            from oracle import Oracle
            o = Oracle(desc)
            ref = o.fd_apply(1.0, 0.02j, r)
            assert np.abs(a - ref).max() / np.abs(ref).max() < 1e-10, name
    finally:
        ctx.close()


@pytest.mark.parametrize("name", list(SHAPES))
def test_split16_transforms_match_fp64(name):
    """The split-FP16 transforms (three FP16 MFMA products per FP32 product) give the
    complex64 solve at least the accuracy of the FP32 MFMA transforms' order (1e-5 relative
    to the FP64 solve; FP32 rounding alone leaves about 1e-6), on residual-like inputs whose
    magnitudes span six decades across the grid. Axes above 160 functions keep the FP32
    transform. walls3d runs the diagonal solve's wall rows (identity) in the fused z pass."""
    import tdsez_amd as T
    desc = SHAPES[name]
    ctx = T.Context(desc)
    try:
        ctx.set_shift(1.0, 0.02j, (0, 0, 0))
        rng = np.random.default_rng(5)
        scale = 10.0 ** rng.uniform(-6, 0, ctx.N)
        r = (rng.standard_normal(ctx.N) + 1j * rng.standard_normal(ctx.N)) * scale
        dr, z64, z32, z16, zl = ctx.vec(r), ctx.vec(), ctx.vec(), ctx.vec(), ctx.vec()
        ctx.pc_setup(T.PC_SEPARABLE)
        ctx.pc_apply(dr, z64)
        ctx.pc_setup(T.PC_SEPARABLE_MIXED)
        ctx.set_fd_transform(0)
        ctx.pc_apply(dr, z32)
        ctx.set_fd_transform(1)
        ctx.pc_apply(dr, z16)
        ctx.set_fd_transform(2)
        ctx.pc_apply(dr, zl)
        a = z64.download()
        e32 = np.linalg.norm(z32.download() - a) / np.linalg.norm(a)
        e16 = np.linalg.norm(z16.download() - a) / np.linalg.norm(a)
        el = np.linalg.norm(zl.download() - a) / np.linalg.norm(a)
        # the LDS-staged form runs the same products in the same order per accumulator
        dl = np.linalg.norm(zl.download() - z16.download()) / np.linalg.norm(a)
        print(name, "relative error of the complex64 solve: FP32 transforms", e32, "split FP16", e16,
              "split FP16 (matrix in LDS)", el, "difference of the two split forms", dl)
        assert e16 < 1e-5, (name, e16, e32)
        assert el < 1e-5, (name, el, e32)
        assert dl < 1e-7, (name, dl)
    finally:
        ctx.set_fd_transform(1)
        ctx.close()

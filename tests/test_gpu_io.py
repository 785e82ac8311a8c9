"""GPU runs with traces and checkpoints: the trace rows are the observables in the
reference's column layout, and a run restarted from a checkpoint continues as the
uninterrupted run did (src/checkpoint.cpp; src/propagator.cpp:78-251)."""
import math
import os

import numpy as np
import pytest

from tests.problems import SOFTCORE_3D

pytestmark = pytest.mark.gpu

io = pytest.importorskip("tdsez_amd.io")


def _skip_without_io():
    if not os.path.exists(io.IO_PATH):
        pytest.skip("libtdz_io.so not built (no HDF5)")


def test_trace_row_columns():
    import tdsez_amd as T
    ctx = T.Context(SOFTCORE_3D)
    d = ctx.vec()
    ctx.project_gaussian(d, (0.3, 0.0, -0.2), (1.0, 1.1, 0.9), (0.2, 0.0, 0.1))
    obs = ctx.observables(d)
    E = (0.01, -0.02, 0.03)
    dip, en, ac = ctx.trace_row(d, 1.5, E, psi0=d)
    q, s = SOFTCORE_3D["charge"], -SOFTCORE_3D["charge"]
    assert dip[0] == 1.5 and np.allclose(dip[1:4], E)
    assert np.allclose(dip[4:7], q * obs[3:6], rtol=1e-13, atol=1e-15)
    inter = s * float(np.dot(E, obs[3:6]))
    assert abs(en[3] - inter) < 1e-14
    assert abs(en[4] - (obs[1] + inter)) < 1e-12
    assert abs(en[6] - obs[0]) < 1e-13 and abs(en[5] - obs[0] / SOFTCORE_3D["mass"]) < 1e-13
    # Kin + Pot + Re<CAP> = Re<H>; the CAP is -iW, so its real part is zero
    assert abs(en[1] + en[2] - obs[1]) < 1e-12 * max(1.0, abs(obs[1]))
    assert abs(ac[1] - obs[0]) < 1e-13 and abs(ac[2]) < 1e-13   # <psi|M|psi>
    ctx.close()


def test_restart_continues_the_run(tmp_path):
    _skip_without_io()
    import tdsez_amd as T
    desc = dict(SOFTCORE_3D, nelem=(8, 8, 9))
    pulse = T.make_pulse(E0=(0.0, 0.0, 0.05), omega=0.5, ncycles=1.0)
    dt, nsteps = 0.05, 20
    ck, full, part = (str(tmp_path / n) for n in ("ck.h5", "full.h5", "restart.h5"))
    ctx = T.Context(desc)
    ctx.pc_setup(T.PC_SEPARABLE)
    psi = ctx.vec()
    ctx.project_gaussian(psi, (0.3, 0.0, -0.2), (1.0, 1.1, 0.9))
    psi0 = ctx.vec(psi.download())
    step, t, _ = io.run_cn(ctx, psi, dt, nsteps, pulse, rtol=1e-13, trace=full, output_stride=5, checkpoint=ck,
                           checkpoint_stride=10, psi0=psi0)
    assert step == nsteps and math.isclose(t, nsteps * dt)
    ref = psi.download()
    ctx2 = T.Context(desc)
    ctx2.pc_setup(T.PC_SEPARABLE)
    psi2 = ctx2.vec()
    psi02 = ctx2.vec(psi0.download())
    step2, t2, _ = io.run_cn(ctx2, psi2, dt, nsteps, pulse, rtol=1e-13, trace=part, output_stride=5, checkpoint=ck,
                             restart=True, psi0=psi02)
    assert step2 == nsteps
    got = psi2.download()

    def l2(v):   # sqrt(v^H M v): the first step after a restart solves again, so the
        dv, mv = ctx.vec(v), ctx.vec()   # coefficients agree to cond(M) x the tolerance
        ctx.term(T.TERM_M, dv, mv)
        return math.sqrt(abs(np.vdot(v, mv.download())))

    assert l2(got - ref) / l2(ref) < 1e-12
    # and both are the oracle's CN steps under the same pulse (midpoint fields)
    from oracle import Oracle
    o = Oracle(desc)
    sE = np.array([ctx.coupling(pulse, (s + 0.5) * dt) for s in range(nsteps)])
    oref, _, _ = o.cn_steps(psi0.download(), dt, sE, rtol=1e-14)
    assert l2(got - oref) / l2(oref) < 1e-11
    a, b = io.read_trace(full, "energies"), io.read_trace(part, "energies")
    assert a.shape == (5, 7) and b.shape == (2, 7)          # rows at steps 0, 5, 10, 15, 20 / 15, 20
    assert np.allclose(b, a[3:], rtol=1e-11, atol=1e-14)
    da = io.read_trace(full, "dipoles")
    assert da.shape == (5, 10) and np.allclose(da[:, 0], dt * np.arange(0, 21, 5))
    ctx.close()
    ctx2.close()

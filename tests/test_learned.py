"""The learned potential-product network: oracle gradients pinned by finite
differences (CPU); GPU input features, forward (FP32 and BF16 MFMA) and GPU
gradients against the oracle; training gain, the ablation against the method
without the network, and the plain-multigrid fallback (GPU)."""
import numpy as np
import pytest

from oracle import cnn

C, CI = 16, 4


def _rand_net(seed, n=(5, 4, 6)):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((CI,) + n[::-1])
    t = rng.standard_normal(n[::-1])
    W1 = rng.standard_normal((27, CI, C)) * 0.3
    W2 = rng.standard_normal((27, C, C)) * 0.05
    P1 = rng.standard_normal((C,) + n[::-1])
    P2 = rng.standard_normal((C,) + n[::-1])
    return u, t, W1, W2, P1, P2


def test_oracle_gradients_finite_difference():
    u, t, W1, W2, P1, P2 = _rand_net(0)
    loss, dW1, dW2, _ = cnn.loss_and_grad(u, t, W1, W2, P1, P2)
    rng = np.random.default_rng(1)
    for _ in range(6):
        i, c, a = rng.integers(27), rng.integers(CI), rng.integers(C)
        e = 1e-6
        Wp, Wm = W1.copy(), W1.copy()
        Wp[i, c, a] += e
        Wm[i, c, a] -= e
        fd = (cnn.loss_and_grad(u, t, Wp, W2, P1, P2)[0] - cnn.loss_and_grad(u, t, Wm, W2, P1, P2)[0]) / (2 * e)
        assert abs(fd - dW1[i, c, a]) < 1e-6 * max(1.0, abs(fd)), (fd, dW1[i, c, a])
        i, a, b = rng.integers(27), rng.integers(C), rng.integers(C)
        Wp, Wm = W2.copy(), W2.copy()
        Wp[i, a, b] += e
        Wm[i, a, b] -= e
        fd = (cnn.loss_and_grad(u, t, W1, Wp, P1, P2)[0] - cnn.loss_and_grad(u, t, W1, Wm, P1, P2)[0]) / (2 * e)
        assert abs(fd - dW2[i, a, b]) < 1e-6 * max(1.0, abs(fd)), (fd, dW2[i, a, b])


def test_oracle_shift_convention():
    u = np.arange(24, dtype=float).reshape(2, 3, 4)
    s = cnn.shift(u, 0, 0, 1)          # s(v) = u(v + x)
    assert s[0, 0, 0] == u[0, 0, 1] and s[0, 0, 3] == 0.0
    s = cnn.shift(u, 1, 0, 0)
    assert s[0, 1, 2] == u[1, 1, 2] and np.all(s[1] == 0.0)


SMALL = dict(dim=3, nelem=(6, 5, 7), lo=(-8.0, -7.0, -9.0), hi=(8.0, 7.0, 9.0), p=4, nquad=8, mass=1.0,
             charge=-1.0, pot_kind=2, pot=(1.0, 2.0), cap_kind=1, cap=(0.01, 3.0))


@pytest.mark.gpu
def test_gpu_features_match_oracle():
    """u = [y, X_d y - x_d M y] in complex64 against the oracle's FP64 operators."""
    import tdsez_amd as T
    from oracle import Oracle, greville
    o, ctx = Oracle(SMALL), T.Context(SMALL)
    rng = np.random.default_rng(4)
    y = rng.standard_normal(ctx.N) + 1j * rng.standard_normal(ctx.N)
    got = ctx.learn_features(ctx.vec(y))
    gr = [greville(o.knots(d), SMALL["p"]) for d in range(3)]
    ref = cnn.features(o, y, gr)
    for c in range(CI):
        err = np.linalg.norm(got[c] - ref[c]) / np.linalg.norm(ref[c])
        assert err < 1e-5, (c, err)
    ctx.close()


@pytest.mark.gpu
def test_gpu_forward_and_gradients_match_oracle():
    import tdsez_amd as T
    ctx = T.Context(SMALL)
    n0, n1, n2 = ctx.n
    rng = np.random.default_rng(3)
    W1 = (rng.standard_normal((27, CI, C)) * 0.3).astype(np.float32)
    W2 = (rng.standard_normal((27, C, C)) * 0.05).astype(np.float32)
    ctx.learn_weights(W1, W2)
    P = ctx.learn_fields().astype(np.float64)                 # [N][C]
    Pg = P.T.reshape(C, n2, n1, n0)
    u = rng.standard_normal((CI, ctx.N)) + 1j * rng.standard_normal((CI, ctx.N))
    t = rng.standard_normal(ctx.N) + 1j * rng.standard_normal(ctx.N)
    loss, dW1, dW2, of, om = ctx.learn_eval(u, t)
    u32 = u.astype(np.complex64).astype(np.complex128)
    t32 = t.astype(np.complex64).astype(np.complex128)
    T2 = np.sum(np.abs(t32) ** 2)
    ref_loss, g1, g2, outs = 0.0, 0.0, 0.0, []
    for comp in (np.real, np.imag):
        uu = comp(u32).reshape(CI, n2, n1, n0)
        tt = comp(t32).reshape(n2, n1, n0)
        out, _ = cnn.forward(uu, W1.astype(np.float64), W2.astype(np.float64), Pg, Pg)
        outs.append(out.reshape(-1))
        l, a1, a2, _ = cnn.loss_and_grad(uu, tt, W1.astype(np.float64), W2.astype(np.float64), Pg, Pg)
        # the oracle normalises per component; the GPU by |t|^2 over both components
        s = np.sum(tt ** 2) / T2
        ref_loss += l * s
        g1 = g1 + a1 * s
        g2 = g2 + a2 * s
    ref_out = np.concatenate(outs)
    assert np.linalg.norm(of - ref_out) / np.linalg.norm(ref_out) < 1e-5
    assert abs(loss - ref_loss) < 1e-5 * ref_loss
    assert np.linalg.norm(dW1 - g1) / np.linalg.norm(g1) < 1e-4
    assert np.linalg.norm(dW2 - g2) / np.linalg.norm(g2) < 1e-4
    # BF16 MFMA inference: same map to bfloat16 accuracy
    assert np.linalg.norm(om - ref_out) / np.linalg.norm(ref_out) < 2e-2
    ctx.close()


@pytest.mark.gpu
def test_inference_fields_match_table():
    """The MFMA inference evaluates the positional fields from their closed forms in FP32.
    With the constant input and field 0 (identically 1) routed through hidden channel 0 by
    the centre taps and read out through field b, every product has one nonzero term, so the
    output is the evaluated field b itself: it must equal the FP32 table that training uses
    to a few FP32 ulps of the field's maximum (BF16 fields would differ by about 4e-3)."""
    import tdsez_amd as T
    ctx = T.Context(SMALL)
    P = ctx.learn_fields()                                   # [N][C], each field scaled to max 1
    assert np.all(P[:, 0] == 1.0)
    u = np.zeros((CI, ctx.N), complex)
    u[0] = 1.0
    t = np.zeros(ctx.N, complex)
    W1 = np.zeros((27, CI, C), np.float32)
    W1[13, 0, 0] = 1.0                                       # centre tap, input 0 -> hidden 0
    for b in range(C):
        W2 = np.zeros((27, C, C), np.float32)
        W2[13, 0, b] = 1.0                                   # centre tap, hidden 0 -> output group b
        ctx.learn_weights(W1, W2)
        om = ctx.learn_eval(u, t)[4]
        err = np.max(np.abs(om[:ctx.N] - P[:, b]))
        assert err < 1e-5, (b, err)
        assert np.all(om[ctx.N:] == 0.0)
    ctx.close()


def _mnorm(ctx, x):
    """L2 norm of the wavefunction with coefficients x: sqrt(x^H M x)."""
    import tdsez_amd as T
    dx, dy = ctx.vec(x), ctx.vec()
    ctx.term(T.TERM_M, dx, dy)
    return float(np.sqrt(np.vdot(x, dy.download()).real))


MID = dict(dim=3, nelem=(24, 24, 24), lo=(-7.5,) * 3, hi=(7.5,) * 3, p=4, nquad=8, mass=1.0, charge=-1.0,
           pot_kind=2, pot=(1.0, 2.0), cap_kind=1, cap=(0.01, 2.5), mg_levels=2)


@pytest.mark.gpu
def test_learned_preconditioner_gain_and_fallback():
    """Train on the solver's own vectors. The trained network must capture most of the
    potential product the lumped form misses, contract residuals more than the method
    without it (the ablation), never need more Krylov iterations, and leave the
    solution unchanged (the FP64 residual judges). A broken network regresses and
    plain multigrid takes over with the same answer."""
    import tdsez_amd as T
    ctx = T.Context(MID)
    psi = ctx.vec()
    ctx.project_gaussian(psi, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    pulse = T.make_pulse(E0=(0.0, 0.0, 0.05), omega=0.057, ncycles=4.0)
    rtol = 1e-13   # the regression threshold is measured at the tolerance of the runs below
    l0, l1 = ctx.learn_train(psi, 0.0, 0.02, pulse, harvest_steps=2, max_samples=6, iters=150, lr=3e-3, rtol=rtol)
    print("loss per sample before/after the fit", l0, l1)
    assert l1 < l0 and l1 < 0.5, (l0, l1)
    # ablation: one preconditioner apply on the initial CN residuals -2 g H_E phi of
    # states the network has not seen, residual of the CN operator
    g, sE = 0.01j, (0.0, 0.0, 0.03)
    rs = []
    for centre, momentum in (((0.5, -0.3, 0.2), (0.3, 0.0, -0.2)), ((-0.8, 0.4, 0.0), (0.0, 0.5, 0.4))):
        phi, hphi = ctx.vec(), ctx.vec()
        ctx.project_gaussian(phi, centre, (1.2, 1.0, 1.1), momentum)
        ctx.set_shift(0.0, 1.0, sE)
        ctx.apply(phi, hphi)
        rs.append(-2.0 * g * hphi.download())
    ctx.set_shift(1.0, g, sE)
    contraction = {}
    for kind in (T.PC_CORRECTED, T.PC_LEARNED):
        ctx.pc_setup(kind)
        worst = 0.0
        for r in rs:
            dr, dz, dw = ctx.vec(r), ctx.vec(), ctx.vec()
            ctx.pc_apply(dr, dz)
            ctx.apply(dz, dw)
            worst = max(worst, np.linalg.norm(r - dw.download()) / np.linalg.norm(r))
        contraction[kind] = worst
    print("contraction without / with the network", contraction)
    assert contraction[T.PC_LEARNED] < 0.8 * contraction[T.PC_CORRECTED], contraction
    x0 = psi.download()
    runs = {}
    for kind in (T.PC_CORRECTED, T.PC_LEARNED):
        ctx.pc_setup(kind)
        d = ctx.vec(x0)
        its, _ = ctx.evolve_cn(d, 0.0, 0.02, 5, pulse, rtol=rtol)
        runs[kind] = (its, d.download())
        # one adaptive choice of sweeps (and, learned, of the network) per CN solve
        choices = ctx.sweep_stats()
        print(T.PC_LEARNED == kind and "learned" or "network-free", "sweep choices", choices)
        assert sum(choices.values()) == 5, choices
        assert kind == T.PC_LEARNED or not any(k.startswith("net_") for k in choices), choices
    its_c, xc = runs[T.PC_CORRECTED]
    its_l, xl = runs[T.PC_LEARNED]
    print("iterations without / with the network", its_c, its_l)
    assert its_l.sum() < its_c.sum(), (its_l, its_c)   # the network saves iterations on unseen states
    assert ctx.learn_status()[1] == 0
    # same wavefunction in L2 (the coefficient vector itself is only fixed to
    # cond(M) times the solver tolerance)
    assert _mnorm(ctx, xl - xc) / _mnorm(ctx, xc) < 1e-10
    # a broken network regresses: plain multigrid takes over and the answer is unchanged
    W1, W2 = ctx.learn_weights()
    ctx.learn_weights(W1 * 50.0, W2)
    ctx.pc_setup(T.PC_LEARNED)
    d = ctx.vec(x0)
    its_b, _ = ctx.evolve_cn(d, 0.0, 0.02, 5, pulse, rtol=rtol)
    _, fallbacks, active = ctx.learn_status()
    assert fallbacks >= 1 and active == 0
    assert _mnorm(ctx, d.download() - xc) / _mnorm(ctx, xc) < 1e-10
    ctx.close()


@pytest.mark.gpu
def test_correction_forms_agree():
    """The corrected preconditioner in its two sweep forms: form 0 solves each sweep's whole
    iterate z = S^-1 (r - dA~ z) on split-FP16 transforms, form 1 keeps y = S^-1 r and
    solves for the correction alone on single FP16 products. The correction is about 1e-2
    of z, so the two applies contract a CN residual alike and agree in the M-norm (their
    coefficient vectors differ in directions the operator hardly sees, as every solve's
    coefficients do to cond(M) times its accuracy); CN steps with either form reach the
    same wavefunction (the FP64 residual judges)."""
    import tdsez_amd as T
    ctx = T.Context(MID)
    psi = ctx.vec()
    ctx.project_gaussian(psi, (0.3, -0.2, 0.1), (1.0, 1.2, 0.9), (0.2, 0.0, -0.3))
    g, sE = 0.01j, (0.0, 0.0, 0.04)
    hphi = ctx.vec()
    ctx.set_shift(0.0, 1.0, sE)
    ctx.apply(psi, hphi)
    r = -2.0 * g * hphi.download()
    ctx.set_shift(1.0, g, sE)
    ctx.pc_setup(T.PC_CORRECTED)
    out, contraction = {}, {}
    for k in (1, 2):
        ctx.set_corrections(k)
        for form in (0, 1):
            ctx.set_correction_form(form)
            dr, dz, dw = ctx.vec(r), ctx.vec(), ctx.vec()
            ctx.pc_apply(dr, dz)
            ctx.apply(dz, dw)
            out[k, form] = dz.download()
            contraction[k, form] = np.linalg.norm(r - dw.download()) / np.linalg.norm(r)
        diff = _mnorm(ctx, out[k, 1] - out[k, 0]) / _mnorm(ctx, out[k, 0])
        print("sweeps", k, "contraction form 0 / 1", contraction[k, 0], contraction[k, 1], "M-norm difference", diff)
        assert diff < 1e-3, (k, diff)
        assert contraction[k, 1] < 1.2 * contraction[k, 0] + 1e-6, (k, contraction)
    pulse = T.make_pulse(E0=(0.0, 0.0, 0.05), omega=0.057, ncycles=4.0)
    x0 = psi.download()
    finals = {}
    for form in (0, 1):
        ctx.set_correction_form(form)
        ctx.pc_setup(T.PC_CORRECTED)
        d = ctx.vec(x0)
        its, _ = ctx.evolve_cn(d, 0.0, 0.02, 4, pulse, rtol=1e-12)
        finals[form] = d.download()
        print("form", form, "iterations", its)
    assert _mnorm(ctx, finals[1] - finals[0]) / _mnorm(ctx, finals[0]) < 1e-10
    ctx.close()


@pytest.mark.gpu
def test_correction_pass_matches_fp64_reference():
    """One corrected sweep, z = S^-1 (r - dA~ S^-1 r) with dA~ y = g (Vg .* (M y) + CAP y +
    sE.D y), Vg the potential at the Greville points: the complex64 preconditioner (in both
    sweep forms) against the same map composed in FP64 from the FP64 separable solve and
    single-term applies, in the M-norm. Checks the correction pass (its x/y and z passes and
    the right-hand side it forms) directly rather than through iteration counts."""
    import tdsez_amd as T
    from oracle import Oracle, greville
    ctx = T.Context(MID)
    o = Oracle(MID)
    gx, gy, gz = (greville(o.knots(d), MID["p"]) for d in range(3))
    vg = -1.0 / np.sqrt(gx[None, None, :] ** 2 + gy[None, :, None] ** 2 + gz[:, None, None] ** 2 + MID["pot"][1])
    vg = (MID["pot"][0] * vg).ravel()   # s = -charge = 1; x fastest
    g, sE = 0.01j, (0.0, 0.0, 0.04)
    psi, hphi = ctx.vec(), ctx.vec()
    ctx.project_gaussian(psi, (0.4, -0.3, 0.2), (1.1, 0.9, 1.3), (0.3, -0.1, 0.2))
    ctx.set_shift(0.0, 1.0, sE)
    ctx.apply(psi, hphi)
    r = -2.0 * g * hphi.download()
    ctx.set_shift(1.0, g, sE)
    ctx.pc_setup(T.PC_SEPARABLE)
    dr, dy, dt = ctx.vec(r), ctx.vec(), ctx.vec()
    ctx.pc_apply(dr, dy)
    y = dy.download()
    terms = {}
    for t in (T.TERM_M, T.TERM_CAP, T.TERM_DZ):
        ctx.term(t, dy, dt)
        terms[t] = dt.download()
    q = r - g * (vg * terms[T.TERM_M] + terms[T.TERM_CAP] + sE[2] * terms[T.TERM_DZ])
    dq, dref = ctx.vec(q), ctx.vec()
    ctx.pc_apply(dq, dref)
    ref = dref.download()
    corr = _mnorm(ctx, ref - y) / _mnorm(ctx, ref)
    ctx.pc_setup(T.PC_CORRECTED)
    ctx.set_corrections(1)
    for form, tol in ((0, 1e-5), (1, 1e-4)):
        ctx.set_correction_form(form)
        dz = ctx.vec()
        ctx.pc_apply(dr, dz)
        err = _mnorm(ctx, dz.download() - ref) / _mnorm(ctx, ref)
        print("form", form, "M-norm error against the FP64 composition", err, "(size of the correction", corr, ")")
        assert err < tol, (form, err)
        assert err < 0.05 * corr, (form, err, corr)   # the correction itself is resolved
    ctx.close()

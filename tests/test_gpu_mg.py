"""Geometric B-spline multigrid on the GPU against the oracle, through the C ABI.

- transfers: the device P and R = P^T of every level against the Kronecker product of
  the oracle's Boehm prolongations (src/multigrid.cpp:52-81, :118-162);
- coarse operators: tdz_level_apply at levels 1 and 2 against P^T A_0 P with A_0 the
  oracle's matrix-free operator (PCMGSetGalerkin(PC_MG_GALERKIN_BOTH),
  src/multigrid.cpp:277), every term including the potential, CAP and field;
- CN steps preconditioned by the 3-level V-cycle (FP64 and complex64) against the
  oracle's explicit steps, and the preconditioner refresh rule and strong-field warning
  (include/tdsez_propagator.hpp:17-31, src/multigrid.cpp:213-237,
  tests/test_multigrid.py:67-94 of the reference).
Tolerances: transfers and operators 1e-13 relative (rounding), CN steps 1e-12 in L2.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# config 2 shape with enough elements for a 3-level hierarchy (p = 4: 20 -> 10 -> 5)
MG3D = dict(dim=3, nelem=(20, 20, 24), lo=(-10.0, -10.0, -12.0), hi=(10.0, 10.0, 12.0), p=4, nquad=8, mass=1.0,
            charge=-1.0, pot_kind=2, pot=(1.0, 2.0), cap_kind=1, cap=(0.01, 3.0), mg_levels=3)
# p = 3, Q = 5, heavier carrier, stronger CAP, uneven axes (16 -> 8 -> 4). The V-cycle's
# separable smoother leaves the CAP out, so the CAP must stay a perturbation of the mass
# term (tau max W < 1); the Manolopoulos profile, which diverges at the walls, is not
# (tau max W ~ 100 at the outermost Gauss points here) and warns
MG3D_P3 = dict(dim=3, nelem=(16, 20, 16), lo=(-8.0, -9.0, -8.0), hi=(8.0, 9.0, 8.0), p=3, nquad=5, mass=1.7,
               charge=-1.0, pot_kind=2, pot=(1.0, 1.0), cap_kind=1, cap=(0.05, 3.0), mg_levels=3)
PROBLEMS = {"mg3d_p4": MG3D, "mg3d_p3": MG3D_P3}


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _kron_apply(P, x, nin):
    """y = (P_z (x) P_y (x) P_x) x for x in x-fastest order of sizes nin."""
    X = x.reshape(nin[2], nin[1], nin[0])
    X = np.einsum("ia,kja->kji", P[0], X)
    X = np.einsum("jb,kbi->kji", P[1], X)
    X = np.einsum("kc,cji->kji", P[2], X)
    return X.reshape(-1)


@pytest.fixture(scope="module", params=list(PROBLEMS))
def mg(request):
    import tdsez_amd as T
    import oracle
    desc = PROBLEMS[request.param]
    o = oracle.Oracle(desc)
    ctx = T.Context(desc)
    assert ctx.levels == 3, ctx.levels
    p = desc["p"]
    knots = [[o.knots(d) for d in range(3)]]
    for _ in range(2):
        knots.append([oracle.coarsen_knots(k) for k in knots[-1]])
    # P[l][d]: level l+1 -> level l along axis d
    P = [[oracle.prolongation(knots[l + 1][d], knots[l][d], p) for d in range(3)] for l in range(2)]
    dims = [tuple(len(knots[l][d]) - p - 1 for d in range(3)) for l in range(3)]
    for l in range(3):
        assert ctx.level_dims(l) == dims[l], (l, ctx.level_dims(l), dims[l])
    yield request.param, desc, o, ctx, P, dims
    ctx.close()


@pytest.mark.parametrize("level", [0, 1])
def test_transfers_match_oracle(mg, level):
    name, desc, o, ctx, P, dims = mg
    nf, nc = dims[level], dims[level + 1]
    xc = _rand(int(np.prod(nc)), 1)
    xf = _rand(int(np.prod(nf)), 2)
    pad = lambda v: np.concatenate([v, np.zeros(ctx.N - v.size, complex)])
    dxc, dxf, dy = ctx.vec(pad(xc)), ctx.vec(pad(xf)), ctx.vec()
    ctx.mg_transfer(level, 1, dxc, dy)                       # prolong
    ref = _kron_apply(P[level], xc, nc)
    assert _rel(dy.download()[:ref.size], ref) < 1e-13, name
    ctx.mg_transfer(level, 0, dxf, dy)                       # restrict
    ref = _kron_apply([M.T for M in P[level]], xf, nf)
    assert _rel(dy.download()[:ref.size], ref) < 1e-13, name


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("alpha,g,sE", [(1.0, 0.01j, (0.0, 0.0, 0.05)), (0.3 - 0.2j, 0.7 + 0.1j, (0.02, -0.01, 0.05))])
def test_galerkin_coarse_operator(mg, level, alpha, g, sE):
    """The coarse operator is exactly P^T A P, potential included (no rediscretisation)."""
    name, desc, o, ctx, P, dims = mg
    nl = dims[level]
    x = _rand(int(np.prod(nl)), 3)
    # composite prolongation level -> 0
    Pl = [np.eye(dims[0][d]) for d in range(3)]
    for l in range(level):
        Pl = [Pl[d] @ P[l][d] for d in range(3)]
    sE = np.array(sE)
    ref = _kron_apply([M.T for M in Pl], o.apply(alpha, g, sE, _kron_apply(Pl, x, nl)), dims[0])
    ctx.set_shift(alpha, g, sE)
    dx = ctx.vec(np.concatenate([x, np.zeros(ctx.N - x.size, complex)]))
    dy = ctx.vec()
    ctx.level_apply(level, dx, dy)
    assert _rel(dy.download()[:x.size], ref) < 1e-13, (name, level, _rel(dy.download()[:x.size], ref))


def test_fine_level_apply_is_the_operator(mg):
    name, desc, o, ctx, P, dims = mg
    x = _rand(o.N, 4)
    sE = np.array([0.0, 0.01, 0.03])
    ctx.set_shift(1.0, 0.01j, sE)
    dx, dy = ctx.vec(x), ctx.vec()
    ctx.level_apply(0, dx, dy)
    assert _rel(dy.download(), o.apply(1.0, 0.01j, sE, x)) < 1e-13


@pytest.mark.parametrize("pc_name", ["MG", "MG_MIXED"])
def test_mg_cn_steps_match_oracle(mg, pc_name):
    import tdsez_amd as T
    name, desc, o, ctx, P, dims = mg
    dt, nsteps = 0.02, 3
    psi0 = o.project_gaussian(centre=(0.5, -0.3, 0.2), width=(1.2, 1.0, 1.1), momentum=(0.3, 0.0, -0.2))
    sE = np.zeros((nsteps, 3))
    sE[:, 2] = 0.05 * np.cos(0.3 * (np.arange(nsteps) + 0.5) * dt)
    ref, _, _ = o.cn_steps(psi0, dt, sE, rtol=1e-13)
    ctx.pc_setup(getattr(T, "PC_" + pc_name))
    d = ctx.vec(psi0)
    its = []
    for s in range(nsteps):
        its.append(ctx.step_cn(d, dt, sE[s], rtol=1e-13).its)
    diff = d.download() - ref
    err = np.sqrt(abs(np.vdot(diff, o.term(0, diff)))) / np.sqrt(abs(np.vdot(ref, o.term(0, ref))))
    assert err < 1e-12, (name, pc_name, err)
    assert max(its) <= 10, (name, pc_name, its)


def test_refresh_cap_rebuilds_every_solve(mg):
    """-tdsez_pc_refresh_its 0: every solve with an iteration asks for a rebuild; the answer
    is the same (the preconditioner never changes the solution beyond the tolerance)."""
    import tdsez_amd as T
    name, desc, o, ctx, P, dims = mg
    psi0 = o.project_gaussian(centre=(0.2, 0.0, -0.1), width=(1.0, 1.0, 1.0))
    pulse = T.make_pulse((0.0, 0.0, 0.05), 0.057, 2.0)
    ctx.pc_setup(T.PC_MG)
    base = ctx.vec(psi0)
    ctx.set_pc_refresh(-1)
    its0, _ = ctx.evolve_cn(base, 0.0, 0.02, 3, pulse, rtol=1e-12)
    st0 = ctx.step_cn(ctx.vec(psi0), 0.02, (0, 0, 0.01), rtol=1e-12).pc_rebuilds
    capped = ctx.vec(psi0)
    ctx.set_pc_refresh(0)
    its1, _ = ctx.evolve_cn(capped, 0.0, 0.02, 3, pulse, rtol=1e-12)
    st1 = ctx.step_cn(ctx.vec(psi0), 0.02, (0, 0, 0.01), rtol=1e-12).pc_rebuilds
    ctx.set_pc_refresh(-1)
    assert st1 - st0 == 3, (st0, st1, its1)
    a, b = base.download(), capped.download()
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-10


def test_refresh_rule_two_base_plus_five(mg):
    """A solve above 2 base + 5 iterations (base: the first solve after a build) rebuilds."""
    import tdsez_amd as T
    name, desc, o, ctx, P, dims = mg
    psi0 = o.project_gaussian(centre=(0.2, 0.0, -0.1), width=(1.0, 1.0, 1.0))
    ctx.pc_setup(T.PC_SEPARABLE)
    ctx.set_pc_refresh(-1)
    d = ctx.vec(psi0)
    r0 = ctx.step_cn(d, 0.02, (0, 0, 0), rtol=1e-12).pc_rebuilds   # step_cn applies no rule
    its_a, _ = ctx.evolve_cn(d, 0.0, 0.02, 1, None, rtol=1e-3)       # base: a loose solve
    # a long step: the potential, which the separable solve leaves out, dominates
    its_b, _ = ctx.evolve_cn(d, 0.02, 2.0, 1, None, rtol=1e-10, maxit=2000)
    r1 = ctx.step_cn(ctx.vec(psi0), 0.02, (0, 0, 0), rtol=1e-6).pc_rebuilds
    expect = 1 if its_b[0] > 2 * its_a[0] + 5 else 0
    assert its_b[0] > 2 * its_a[0] + 5, (its_a, its_b)
    assert r1 - r0 == expect, (r0, r1, its_a, its_b)


def test_strong_field_warning():
    """tau max|E.r| > 1.5 with multigrid warns (the reference's test: a field of 2 a.u. at
    TimeStep = 0.2; the sin^2 pulse here peaks at about 0.7 E0)."""
    import tdsez_amd as T
    desc = dict(MG3D, nelem=(20, 20, 20), lo=(-10.0,) * 3, hi=(10.0,) * 3)
    ctx = T.Context(desc)
    try:
        ctx.pc_setup(T.PC_MG)
        psi = ctx.vec(np.zeros(ctx.N, complex))
        ctx.evolve_cn(psi, 0.0, 0.2, 0, T.make_pulse((0.0, 0.0, 0.5), 1.0, 1.0), rtol=1e-10)
        assert "tau*max|E.r|" not in T.last_warning()          # 0.1 * 0.35 * 10 < 1.5
        ctx.evolve_cn(psi, 0.0, 0.2, 0, T.make_pulse((0.0, 0.0, 4.0), 1.0, 1.0), rtol=1e-10)
        assert "tau*max|E.r|" in T.last_warning()
    finally:
        ctx.close()


def test_dominant_cap_warning():
    """A CAP that dominates the mass term near the walls (the Manolopoulos profile) warns with
    multigrid: the smoother's separable solve does not contain it."""
    import tdsez_amd as T
    desc = dict(MG3D_P3, cap_kind=2, cap=(1.2,))
    ctx = T.Context(desc)
    try:
        ctx.pc_setup(T.PC_MG)
        psi = ctx.vec(np.zeros(ctx.N, complex))
        ctx.evolve_cn(psi, 0.0, 0.02, 0, None, rtol=1e-10)
        assert "CAP" in T.last_warning()
    finally:
        ctx.close()


@pytest.mark.parametrize("pc_name", ["MG", "MG_MIXED"])
def test_chebyshev_smoother_cn_steps_match_oracle(mg, pc_name):
    """The Jacobi-Chebyshev smoother (the mixed-precision configuration's FP32 smoother)
    gives the same CN steps; its contraction is reported beside the separable solve's."""
    import tdsez_amd as T
    name, desc, o, ctx, P, dims = mg
    dt, nsteps = 0.02, 3
    psi0 = o.project_gaussian(centre=(0.5, -0.3, 0.2), width=(1.2, 1.0, 1.1), momentum=(0.3, 0.0, -0.2))
    sE = np.zeros((nsteps, 3))
    sE[:, 2] = 0.05 * np.cos(0.3 * (np.arange(nsteps) + 0.5) * dt)
    ref, _, _ = o.cn_steps(psi0, dt, sE, rtol=1e-13)
    kind = getattr(T, "PC_" + pc_name)
    r = -2 * 0.01j * o.apply(0.0, 1.0, sE[0], psi0)   # an initial CN residual
    contraction = {}
    try:
        for smoother, degree in ((0, 1), (1, 2), (1, 4)):
            ctx.set_smoother(smoother, degree)
            ctx.pc_setup(kind)
            ctx.set_shift(1.0, 0.01j, sE[0])
            dr, dz, dw = ctx.vec(r), ctx.vec(), ctx.vec()
            ctx.pc_apply(dr, dz)
            ctx.apply(dz, dw)
            contraction[(smoother, degree)] = np.linalg.norm(r - dw.download()) / np.linalg.norm(r)
            if smoother == 1 and degree == 4:
                # point smoothing is weak on these small mass-dominated grids (tau / h^2 = 0.01; the
                # reference found point smoothers poor on this operator, PERFORMANCE.md:117-120):
                # FGMRES(30) needs about 100-300 iterations to 1e-12 and, unlike the multigrid legs
                # with the separable smoother, stops just below the tolerance, so the step error is
                # that residual magnified by the conditioning of the mass-dominated matrix (the
                # B-spline mass matrix of p = 4 in 3D), not far below it: 2e-10 to 7e-10 here
                print(name, pc_name, "contraction (smoother, degree):", contraction, flush=True)
                d = ctx.vec(psi0)
                its = [ctx.step_cn(d, dt, sE[s], rtol=1e-12, maxit=2000).its for s in range(nsteps)]
                diff = d.download() - ref
                err = np.sqrt(abs(np.vdot(diff, o.term(0, diff)))) / np.sqrt(abs(np.vdot(ref, o.term(0, ref))))
                print(name, pc_name, "Chebyshev degree 4: its", its, "L2", err, flush=True)
                assert err < 2e-9, (name, pc_name, degree, err, its)
        print(name, pc_name, "contraction (smoother, degree):", contraction)
        assert contraction[(1, 4)] < 1.0
    finally:
        ctx.set_smoother(0, 1)
        ctx.pc_setup(kind)


@pytest.mark.parametrize("g,sE", [(0.01j, (0.0, 0.0, 0.05)), (0.1j, (0.02, -0.01, 0.05))])
def test_coarse_krylov_solve_matches_dense_solve(mg, g, sE):
    """The Krylov coarse solve (in place of the reference's coarse LU, KSPPREONLY + PCLU,
    src/multigrid.cpp:314-325) solves the coarsest Galerkin system P^T A P x = b assembled
    densely from the oracle's operator: the same x as numpy's dense solve to 1e-10, and a
    relative residual in that dense matrix below 1e-11 (the device operator matches the dense
    one to about 1e-13 relative, and |A| |x| / |b| is about 40 here). The default separable
    sweeps only approximate it."""
    import tdsez_amd as T
    name, desc, o, ctx, P, dims = mg
    level = 2
    nl = dims[level]
    n = int(np.prod(nl))
    Pl = [P[0][d] @ P[1][d] for d in range(3)]
    PlT = [M.T for M in Pl]
    sE = np.array(sE)
    E = np.eye(n)
    A = np.empty((n, n), complex)
    for j in range(n):
        A[:, j] = _kron_apply(PlT, o.apply(1.0, g, sE, _kron_apply(Pl, E[:, j], nl)), dims[0])
    b = _rand(n, 11)
    x_ref = np.linalg.solve(A, b)
    ctx.pc_setup(T.PC_MG)
    ctx.set_shift(1.0, g, sE)
    pad = lambda v: np.concatenate([v, np.zeros(ctx.N - v.size, complex)])
    db, dx = ctx.vec(pad(b)), ctx.vec()
    res = {}
    try:
        for kind in (0, 1):
            ctx.set_coarse_solve(kind, rtol=1e-13, maxit=512)
            ctx.level_solve(level, db, dx)
            x = dx.download()[:n]
            res[kind] = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
            if kind == 1:
                its, est = ctx.coarse_status()
                err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
                print(name, g, "coarse Krylov its", its, "estimate", est, "dense residual", res[1], "x error", err)
                assert res[1] < 1e-11, (name, its, est, res[1])
                assert err < 1e-10, (name, err)
        print(name, g, "dense residual of the separable sweeps", res[0])
        assert res[0] > res[1]
    finally:
        ctx.set_coarse_solve(0)


@pytest.mark.parametrize("pc_name", ["MG", "MG_MIXED"])
def test_mg_krylov_coarse_cn_steps_match_oracle(mg, pc_name):
    """CN steps with the Krylov coarse solve in the V-cycle (FP64 also inside the complex64
    V-cycle) against the oracle's explicit steps."""
    import tdsez_amd as T
    name, desc, o, ctx, P, dims = mg
    dt, nsteps = 0.02, 3
    psi0 = o.project_gaussian(centre=(0.5, -0.3, 0.2), width=(1.2, 1.0, 1.1), momentum=(0.3, 0.0, -0.2))
    sE = np.zeros((nsteps, 3))
    sE[:, 2] = 0.05 * np.cos(0.3 * (np.arange(nsteps) + 0.5) * dt)
    ref, _, _ = o.cn_steps(psi0, dt, sE, rtol=1e-13)
    try:
        ctx.set_coarse_solve(1, rtol=1e-12, maxit=128)
        ctx.pc_setup(getattr(T, "PC_" + pc_name))
        d = ctx.vec(psi0)
        its = [ctx.step_cn(d, dt, sE[s], rtol=1e-13).its for s in range(nsteps)]
        diff = d.download() - ref
        err = np.sqrt(abs(np.vdot(diff, o.term(0, diff)))) / np.sqrt(abs(np.vdot(ref, o.term(0, ref))))
        print(name, pc_name, "Krylov coarse solve: its", its, "coarse", ctx.coarse_status(), "L2", err)
        assert err < 1e-12, (name, pc_name, err)
        assert max(its) <= 10, (name, pc_name, its)
    finally:
        ctx.set_coarse_solve(0)
        ctx.pc_setup(getattr(T, "PC_" + pc_name))

"""Benchmark: Crank-Nicolson steps/s of the 3D soft-core atom at 128^3, p = 4,
learned mixed-precision preconditioner against plain FP64 multigrid.

Workload (the 3D soft-core configuration): V = -1/sqrt(r^2+2) on [-40,40]^3,
128 uniform elements per axis, p = 4, Q = 8 Gauss points, a quadratic CAP in
the outer 8 a.u. (strength 0.01), a z-polarised sin^2 pulse E0 = 0.05,
omega = 0.057, 4 cycles; CN dt = 0.02; FGMRES to ||r||/||b|| < 1e-10. The
initial state is the lowest eigenstate of H, computed on the GPU at set-up. A
"step" is one CN step; inputs are resident in HBM.

Default run: the same steps from the same state twice, first with plain FP64
geometric multigrid (the baseline), then with the learned mixed-precision
preconditioner (network trained at set-up on vectors harvested from two CN
steps; set-up time is not in the timed region). value = steps/s of the learned
run; the baseline's steps/s, the speed-up and the L2 difference of the two final
wavefunctions go into the same line. Prints one JSON line (see the driver
contract). The CPU baseline times the CPU oracle on one CN step of the same
problem (bounded sample) on this host.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12  # MI355X dense FP64 (vector and matrix), FLOP/s
PEAK_HBM = 8.0e12


def problem(nel=128):
    return dict(dim=3, nelem=(nel,) * 3, lo=(-40.0,) * 3, hi=(40.0,) * 3, p=4, nquad=8, mass=1.0, charge=-1.0,
                pot_kind=2, pot=(1.0, 2.0), cap_kind=1, cap=(0.01, 8.0), mg_levels=3, mg_smooth=1)


PCS = {"sep": "fp64 separable (fast diagonalisation)", "mg": "plain fp64 GMG",
       "sep32": "separable solve in complex64", "mg32": "mixed-precision GMG (complex64 V-cycle)",
       "corr32": "complex64 first-order corrected separable solve (lumped potential)",
       "corrx32": "complex64 first-order corrected separable solve (matrix-free potential)",
       "learned32": "learned mixed precision: complex64 corrected separable solve + BF16-MFMA network remainder, "
                    "FP64 operator/orthogonalisation/residual, plain FP64 GMG fallback",
       "none": "none"}


def pot_flops_per_element(p, Q):
    n = p + 1
    return 2 * 4 * (n ** 3 * Q + n ** 2 * Q ** 2 + n * Q ** 3) + 2 * Q ** 3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pc", default="learned32", choices=list(PCS))
    ap.add_argument("--no-baseline", action="store_true", help="skip the plain FP64 GMG run")
    ap.add_argument("--train-iters", type=int, default=150)
    ap.add_argument("--nel", type=int, default=128)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--predictor", type=int, default=2, choices=[0, 1, 2],
                    help="initial guess of the CN solves, both legs: 0 the current state (as TSTHETA), "
                         "1/2 the binomial extrapolation of the last 2/3 states")
    ap.add_argument("--no-plain-guess", action="store_true",
                    help="skip the comparison runs that start each solve from the current state")
    ap.add_argument("--init", default="ground", choices=["ground", "gaussian"],
                    help="initial state: lowest eigenstate (default) or the projected Gaussian (profiling runs)")
    ap.add_argument("--workload", default="atom128", choices=["atom128", "scan", "hetero"],
                    help="atom128: the 3D soft-core CN benchmark (default); scan: the batched 2D intensity/CEP "
                         "scan; hetero: the layered heterostructure with Magnus-4")
    ap.add_argument("--runs", type=int, default=64, help="scan: number of runs")
    ap.add_argument("--seed", type=int, default=7, help="scan: seed of the drawn grids and pulses")
    ap.add_argument("--kiters", type=int, default=None,
                    help="scan: FGMRES iterations captured per step (default 2 with the predictor, else 4)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    if args.workload == "scan":
        return scan(args)
    if args.workload == "hetero":
        return hetero(args)
    import tdsez_amd as T
    desc = problem(args.nel)
    t_setup = time.perf_counter()
    ctx = T.Context(desc)
    kinds = {"sep": T.PC_SEPARABLE, "mg": T.PC_MG, "none": T.PC_NONE, "sep32": T.PC_SEPARABLE_MIXED,
             "mg32": T.PC_MG_MIXED, "corr32": T.PC_CORRECTED, "corrx32": T.PC_CORRECTED_EXACT,
             "learned32": T.PC_LEARNED}
    ctx.pc_setup(T.PC_MG if ctx.levels > 1 else T.PC_SEPARABLE)
    psi = ctx.vec()
    ctx.project_gaussian(psi, (0.0, 0.0, 0.0), (1.5, 1.5, 1.5))
    E0, res0, it0 = float("nan"), float("nan"), 0
    if args.init == "ground":
        E0, res0, it0 = ctx.ground_state(psi, sigma=-0.3, tol=1e-8, maxit=80)
    print(f"# ground state E0 = {E0:.10f} (eigen-residual {res0:.1e}, {it0} inverse iterations)",
          file=sys.stderr, flush=True)
    pulse = T.make_pulse(E0=(0.0, 0.0, 0.05), omega=0.057, ncycles=4.0)
    dt = 0.02
    x0 = psi.download()

    def run(kind, timing, order=None):
        """W untimed + K timed CN steps from x0; returns (steps/s, its, final psi, kernel ms, launches)."""
        ctx.pc_setup(kind)
        ctx.set_cn_predictor(args.predictor if order is None else order)
        d = ctx.vec(x0)
        ctx.evolve_cn(d, 0.0, dt, args.warmup, pulse, rtol=args.rtol, maxit=500)
        ctx.sync()
        if timing:
            ctx.timing(True)
        t0 = time.perf_counter()
        its, _ = ctx.evolve_cn(d, args.warmup * dt, dt, args.steps, pulse, rtol=args.rtol, maxit=500)
        ctx.sync()
        el = time.perf_counter() - t0
        kms, nl = ctx.timing_stats() if timing else (0.0, 0)
        ctx.timing(False)
        out = d.download()
        d.free()
        return args.steps / el, its, out, kms, nl

    base = None
    if args.pc == "learned32" and not args.no_baseline:
        sps_b, its_b, psi_b, _, _ = run(T.PC_MG, False)
        base = {"preconditioner": PCS["mg"], "steps_per_s": sps_b, "ms_per_step": 1e3 / sps_b,
                "krylov_its_per_step": float(np.mean(its_b))}
        print(f"# baseline FP64 GMG: {sps_b:.2f} steps/s, its/step {its_b.tolist()}", file=sys.stderr, flush=True)
    learn = None
    if args.pc == "learned32":
        tl = time.perf_counter()
        l0, l1 = ctx.learn_train(psi, 0.0, dt, pulse, harvest_steps=2, max_samples=6, iters=args.train_iters,
                                 lr=3e-3, rtol=args.rtol)
        base_its, _, _ = ctx.learn_status()
        learn = {"loss_physics_init": l0, "loss_trained": l1, "train_s": time.perf_counter() - tl,
                 "its_without_network": base_its}
        print(f"# trained: loss/sample {l0:.3e} -> {l1:.3e} in {learn['train_s']:.1f} s", file=sys.stderr, flush=True)
    t_setup = time.perf_counter() - t_setup
    print(f"# setup {t_setup:.2f} s, N = {ctx.N}, levels = {ctx.levels}, pc = {args.pc}", file=sys.stderr, flush=True)

    steps_per_s, its, psi_end, kms, nlaunch = run(kinds[args.pc], True)
    elapsed = args.steps / steps_per_s
    print(f"# timed {args.steps} steps in {elapsed:.3f} s; Krylov its/step {its.tolist()}", file=sys.stderr,
          flush=True)
    if learn is not None:
        _, fallbacks, active = ctx.learn_status()
        learn.update({"fallbacks": fallbacks, "network_active": bool(active)})
    if base is not None:
        def mnorm(x):
            dx, dy = ctx.vec(x), ctx.vec()
            ctx.term(T.TERM_M, dx, dy)
            v = float(np.sqrt(abs(np.vdot(x, dy.download()).real)))
            dx.free()
            dy.free()
            return v
        base["speedup"] = steps_per_s / base["steps_per_s"]
        base["psi_rel_l2_diff"] = mnorm(psi_end - psi_b) / mnorm(psi_b)

    netfree = None
    if args.pc == "learned32" and not args.no_baseline:
        # the same solve without the network remainder (first-order corrected separable
        # solve with the lumped potential): what the network adds at this tolerance
        sps_c, its_c, _, _, _ = run(T.PC_CORRECTED, False)
        netfree = {"preconditioner": PCS["corr32"], "steps_per_s": sps_c, "krylov_its_per_step": float(np.mean(its_c))}
        print(f"# network-free corrected solve: {sps_c:.2f} steps/s, its/step {np.mean(its_c):.1f}",
              file=sys.stderr, flush=True)

    plain = None
    if args.predictor > 0 and not args.no_plain_guess:   # the same legs with the current state as guess
        sps0, its0, _, _, _ = run(kinds[args.pc], False, 0)
        plain = {"steps_per_s": sps0, "krylov_its_per_step": float(np.mean(its0))}
        if netfree is not None:
            sps0c, its0c, _, _, _ = run(T.PC_CORRECTED, False, 0)
            plain.update({"network_free_steps_per_s": sps0c, "network_free_krylov_its_per_step": float(np.mean(its0c))})
        if base is not None:
            sps0b, its0b, _, _, _ = run(T.PC_MG, False, 0)
            plain.update({"fp64_mg_steps_per_s": sps0b, "fp64_mg_krylov_its_per_step": float(np.mean(its0b)),
                          "speedup": sps0 / sps0b})
        print(f"# without predictor: {sps0:.2f} steps/s (its {np.mean(its0):.1f})"
              + (f", FP64 GMG {plain['fp64_mg_steps_per_s']:.2f} steps/s" if base is not None else ""),
              file=sys.stderr, flush=True)

    p, Q, nel = desc["p"], desc["nquad"], args.nel
    flops = pot_flops_per_element(p, Q) * nel ** 3
    avg_s = (kms / 1e3) / max(nlaunch, 1)
    achieved = flops / avg_s if nlaunch else 0.0
    roofline = {"bound": "mfma", "kernel": "pot3d_kernel", "achieved": achieved / 1e12,
                "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": None,
                "launches": nlaunch, "avg_ms": avg_s * 1e3, "flops_per_launch": flops}
    # HBM bytes per launch from the committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs
    # (copy-kernel calibrated); only meaningful for the 128^3 workload they were taken on
    tj = os.path.join(ROOT, "profiles", "pot3d_softcore128_traffic.json")
    if nel == 128 and os.path.exists(tj):
        with open(tj) as f:
            tr = json.load(f)
        roofline["traffic"] = tr["traffic_bytes"]
        roofline["traffic_source"] = "profiles/pot3d_softcore128_traffic.json (" + tr["kernel"] + ")"

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(desc, dt, pulse, args, x0)

    line = {
        "metric": "CN steps/s, 3D p=4 128^3 grid: learned+mixed PC vs plain FP64 MG at |r|<1e-10",
        "value": steps_per_s, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "complex128 (fp64 operator/Krylov; complex64 + bf16 preconditioner)"
        if args.pc in ("learned32", "sep32", "mg32", "corr32", "corrx32") else "complex128 (fp64)", "data": "synthetic: lowest eigenstate of H (GPU inverse iteration), E0 = %.8f" % E0,
        "config": {"workload": f"3D soft-core atom {nel}^3 elements p=4 Q=8, CN dt=0.02, rtol={args.rtol}",
                   "preconditioner": PCS[args.pc],
                   "cn_guess": ("current state" if args.predictor == 0 else
                                f"binomial extrapolation of the last {args.predictor + 1} states (both legs)"),
                   "dofs": ctx.N, "mg_levels": ctx.levels, "krylov_its_per_step": float(np.mean(its)),
                   "setup_s": t_setup},
        "fp64_mg_baseline": base,
        "without_predictor": plain,
        "learned": learn,
        "network_free": netfree,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def scan_runs(nruns, seed):
    """The batched 2D scan: soft-core V = -1/sqrt(x^2+y^2+1) on [-40,40]^2, elements per
    axis from {64,96,128,192,256}, p from {3,4,5}, E0 log-uniform in [0.02,0.1], CEP
    uniform in [0,2pi), omega = 0.057, 3-cycle sin^2 pulse along x, all from the seed."""
    rng = np.random.default_rng(seed)
    runs = []
    for _ in range(nruns):
        nel = tuple(int(v) for v in rng.choice([64, 96, 128, 192, 256], size=2))
        p = int(rng.choice([3, 4, 5]))
        E0 = float(np.exp(rng.uniform(np.log(0.02), np.log(0.1))))
        cep = float(rng.uniform(0.0, 2 * np.pi))
        desc = dict(dim=2, nelem=nel, lo=(-40.0, -40.0), hi=(40.0, 40.0), p=p, nquad=8, mass=1.0, charge=-1.0,
                    pot_kind=2, pot=(1.0, 1.0))
        runs.append((desc, E0, cep))
    return runs


def scan(args):
    import tdsez_amd as T
    runs = scan_runs(args.runs, args.seed)
    dt = 0.02
    pulses = [T.make_pulse(E0=(E0, 0.0, 0.0), omega=0.057, ncycles=3.0, cep=cep) for _, E0, cep in runs]
    t0 = time.perf_counter()
    if args.kiters is None:
        args.kiters = 2 if args.predictor else 4
    b = T.Batch([d for d, _, _ in runs], pulses, dt, args.warmup + args.steps, pc=T.PC_SEPARABLE_MIXED,
                kiters=args.kiters)
    if args.predictor:
        b.set_predictor(args.predictor)
    for r, ctx in enumerate(b.runs):
        ctx.project_gaussian(b.psi[r], (0.0, 0.0), (1.0, 1.0))
        ctx.ground_state(b.psi[r], sigma=-0.9, tol=1e-8, maxit=60)
    setup = time.perf_counter() - t0
    b.evolve(args.warmup, rtol=args.rtol)          # the first call captures the graph
    fix0 = b.fixups()
    t0 = time.perf_counter()
    its, res = b.evolve(args.steps, rtol=args.rtol)
    el = time.perf_counter() - t0
    dofs = sum((d["nelem"][0] + d["p"]) * (d["nelem"][1] + d["p"]) for d, _, _ in runs)
    line = {
        "metric": "run-steps/s, batched 2D intensity/CEP scan (CN, graph-captured step)",
        "value": args.runs * args.steps / el, "unit": "run-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "complex128 (fp64 operator/Krylov; complex64 separable preconditioner)",
        "data": f"synthetic: {args.runs} runs drawn from seed {args.seed}, ground states at set-up",
        "config": {"workload": f"batched 2D scan, {args.runs} runs, {dofs} dofs in total, CN dt=0.02, "
                               f"rtol={args.rtol}, {args.kiters} captured FGMRES iterations per step",
                   "cn_guess": ("current state" if args.predictor == 0 else
                                f"binomial extrapolation of the last {args.predictor + 1} states"),
                   "setup_s": setup, "krylov_its_per_step": float(its.mean()),
                   "host_completions": b.fixups() - fix0, "host_completions_warmup": fix0,
                   "max_residual": float(res.max())},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    b.close()


NM = 18.897261   # bohr per nanometre
EV = 1.0 / 27.211386   # hartree per electronvolt


def hetero_problem(seed, nxy=64, nz=256, p=4):
    """Layered heterostructure: 5 alternating layers (well m* = 0.067, barrier m* = 0.092
    with a 0.2 eV band offset), each 10-20 nm thick (drawn from the seed); 64 x 64
    elements over 100 nm in x and y; nz elements along z distributed over the layers
    (interface-aligned, C0 at the interfaces)."""
    rng = np.random.default_rng(seed)
    thick = rng.uniform(10.0, 20.0, size=5) * NM
    bounds = np.concatenate([[0.0], np.cumsum(thick)])
    bounds -= 0.5 * bounds[-1]
    masses = np.array([0.067, 0.092, 0.067, 0.092, 0.067])
    voff = np.array([0.0, 0.2 * EV, 0.0, 0.2 * EV, 0.0])
    per = np.maximum(p + 1, np.round(nz * thick / thick.sum()).astype(int))
    per[-1] += nz - per.sum()
    t = [bounds[0]] * (p + 1)
    for i in range(5):
        t += list(np.linspace(bounds[i], bounds[i + 1], per[i] + 1)[1:-1])
        if i < 4:
            t += [bounds[i + 1]] * p
    t += [bounds[-1]] * (p + 1)
    L = 50.0 * NM
    return dict(dim=3, nelem=(nxy, nxy, nz), lo=(-L, -L, bounds[0]), hi=(L, L, bounds[-1]), p=p, nquad=8,
                mass=1.0, charge=-1.0, knots=np.array(t), layers=(bounds, masses, voff), mg_levels=3), bounds


def hetero(args):
    import tdsez_amd as T
    desc, bounds = hetero_problem(args.seed)
    t0 = time.perf_counter()
    ctx = T.Context(desc)
    ctx.pc_setup(T.PC_SEPARABLE)
    psi = ctx.vec()
    zc = 0.5 * (bounds[2] + bounds[3])   # a wave packet in the middle well
    ctx.project_gaussian(psi, (0.0, 0.0, zc), (20 * NM, 20 * NM, 3 * NM))
    setup = time.perf_counter() - t0
    pulse = T.make_pulse(E0=(0.0, 0.0, 1e-3), omega=0.057, ncycles=2.0)
    dt = 1.0
    res = {}
    x0 = psi.download()
    for name, kind in (("separable (z-line)", T.PC_SEPARABLE), ("plain FP64 GMG", T.PC_MG)):
        ctx.pc_setup(kind)
        d = ctx.vec(x0)
        ctx.evolve_rational(d, 0.0, dt, args.warmup, pulse, magnus=True, rtol=args.rtol)
        ctx.sync()
        t1 = time.perf_counter()
        its, _ = ctx.evolve_rational(d, args.warmup * dt, dt, args.steps, pulse, magnus=True, rtol=args.rtol)
        ctx.sync()
        el = time.perf_counter() - t1
        res[name] = {"steps_per_s": args.steps / el, "krylov_its_per_step": float(np.mean(its))}
        d.free()
    sep = res["separable (z-line)"]
    line = {
        "metric": "Magnus-4 steps/s, layered heterostructure 64x64x256 p=4 (BenDaniel-Duke, C0 interfaces)",
        "value": sep["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 / sep["steps_per_s"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex128 (fp64)",
        "data": f"synthetic: 5 layers drawn from seed {args.seed}, Gaussian wave packet in the middle well",
        "config": {"workload": f"heterostructure {ctx.n} dofs per axis, Magnus-4 dt={dt}, rtol={args.rtol}, "
                               f"layer bounds (bohr) {np.round(bounds, 2).tolist()}",
                   "setup_s": setup, "preconditioners": res,
                   "note": "the operator is separable (no non-separable potential): the z-line separable solve "
                           "is its exact inverse, so the network has no remainder to learn"},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def cpu_baseline(desc, dt, pulse, args, psi):
    """Time the CPU oracle on one CN step of the same problem (bounded sample)."""
    try:
        import oracle
        from oracle import Oracle
    except Exception as e:  # noqa: BLE001
        return {"error": f"oracle unavailable: {e}"}
    o = Oracle(desc)
    t = args.warmup * dt + 0.5 * dt
    E = 0.05 * math.sin(math.pi * t / (4 * 2 * math.pi / 0.057)) ** 2 * math.sin(0.057 * t)
    sE = np.array([[0.0, 0.0, E]])
    t0 = time.perf_counter()
    _, its, _ = o.cn_steps(psi, dt, sE, rtol=args.rtol, maxit=500, restart=50, pc=1)
    el = time.perf_counter() - t0
    return {"value": 1.0 / el, "unit": "steps/s", "cores": oracle.lib().orc_threads(), "kind": "port",
            "sample": f"1 CN step of the same {desc['nelem'][0]}^3 problem, FGMRES + separable PC "
                      f"({int(its[0])} its), {el:.1f} s"}


if __name__ == "__main__":
    main()

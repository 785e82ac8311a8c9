"""Benchmark: Crank-Nicolson steps/s of the 3D soft-core atom at 128^3, p = 4,
learned mixed-precision preconditioner against plain FP64 multigrid.

Workload (the 3D soft-core configuration): V = -1/sqrt(r^2+2) on [-40,40]^3,
128 uniform elements per axis, p = 4, Q = 8 Gauss points, a quadratic CAP in
the outer 8 a.u. (strength 0.01), a z-polarised sin^2 pulse E0 = 0.05,
omega = 0.057, 4 cycles (T = 440.9 a.u.); CN dt = 0.02; FGMRES to
||r||/||b|| < 1e-10. The initial state is the lowest eigenstate of H, computed on
the GPU at set-up. A "step" is one CN step; inputs are resident in HBM.

Default run (--workload atom128): two timing windows, each timed for every leg from
the same state, median of --repeats runs of W untimed + K timed steps:
  * "early": from the eigenstate at t = 0 (field about 1e-8);
  * "peak": centred on the field crest nearest the envelope's peak (t = 3.5 pi/omega =
    192.9 a.u., |E| = 0.048), reached by propagating the eigenstate there untimed with
    the learned preconditioner.
Legs: plain FP64 geometric multigrid (the baseline), the learned mixed-precision
preconditioner (network trained at set-up on vectors of two CN steps) and the same
solve without the network. value = the learned leg's median steps/s in the peak
window. --workload fullpulse propagates the whole pulse (22,046 steps) with FP64 GMG,
complex64 GMG and the learned preconditioner and reports steps/s over the run and,
per pulse cycle, Krylov iterations and the wavefunction difference against the FP64
run. The CPU baseline times the CPU oracle (C++/OpenMP, -march=native build made on
this host) on 3 CN steps of the peak window (bounded sample).
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12  # MI355X dense FP64 (vector and matrix), FLOP/s
PEAK_HBM = 8.0e12
DT = 0.02
OMEGA, NCYC, E0Z = 0.057, 4.0, 0.05
T_PULSE = NCYC * 2 * math.pi / OMEGA
T_CREST = 3.5 * math.pi / OMEGA   # the field crest nearest T/2 (the envelope's peak is a zero of the carrier)


def problem(nel=128):
    return dict(dim=3, nelem=(nel,) * 3, lo=(-40.0,) * 3, hi=(40.0,) * 3, p=4, nquad=8, mass=1.0, charge=-1.0,
                pot_kind=2, pot=(1.0, 2.0), cap_kind=1, cap=(0.01, 8.0), mg_levels=3, mg_smooth=1)


PCS = {"sep": "fp64 separable (fast diagonalisation)", "mg": "plain fp64 GMG (exact Galerkin coarse levels)",
       "sep32": "separable solve in complex64", "mg32": "mixed-precision GMG (complex64 V-cycle)",
       "corr32": "complex64 first-order corrected separable solve (lumped potential)",
       "corrx32": "complex64 first-order corrected separable solve (matrix-free potential)",
       "learned32": "learned mixed precision: complex64 corrected separable solve + BF16-MFMA network remainder, "
                    "FP64 operator/orthogonalisation/residual, plain FP64 GMG fallback",
       "none": "none"}


def pot_flops_per_element(p, Q):
    n = p + 1
    return 2 * 4 * (n ** 3 * Q + n ** 2 * Q ** 2 + n * Q ** 3) + 2 * Q ** 3


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed CN steps per window and repeat")
    ap.add_argument("--warmup", type=int, default=2, help="untimed steps before each timed run (predictor history)")
    ap.add_argument("--repeats", type=int, default=3, help="timed runs per leg and window (median reported)")
    ap.add_argument("--windows", default="early,peak", help="comma list of early, peak")
    ap.add_argument("--legs", default="fp64_mg,learned,network_free",
                    help="comma list of fp64_mg, learned, network_free, mixed_mg (profiling runs)")
    ap.add_argument("--train-iters", type=int, default=150)
    ap.add_argument("--init", default="ground", choices=["ground", "gaussian"],
                    help="initial state: lowest eigenstate (default) or the projected Gaussian (profiling runs)")
    ap.add_argument("--nel", type=int, default=128)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--correction-form", type=int, default=1, choices=[0, 1],
                    help="correction sweeps of the corrected/learned legs: 1 (default) the correction alone on "
                         "single FP16 transforms where a solve needs two or more iterations, 0 the whole iterate")
    ap.add_argument("--predictor", type=int, default=0, choices=[0, 1, 2, 3, 4],
                    help="initial guess of the CN solves of the value's runs, every leg: 0 the current state "
                         "(TSTHETA's, the reference's configuration), k > 0 the binomial extrapolation of the "
                         "last k+1 states")
    ap.add_argument("--alt-predictor", type=int, default=2, choices=[-1, 0, 1, 2, 3, 4],
                    help="guess of the second set of runs reported beside the value (-1 to skip)")
    ap.add_argument("--workload", default="atom128", choices=["atom128", "fullpulse", "scan", "hetero", "fdsolve"],
                    help="atom128: the 3D soft-core CN benchmark (default); fullpulse: the whole pulse with "
                         "per-cycle reports; scan: the batched 2D intensity/CEP scan; hetero: the layered "
                         "heterostructure with Magnus-4; fdsolve: the separable (fast-diagonalisation) solve "
                         "alone at 128^3, complex64 with each transform kind and FP64 (profiling runs)")
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
    if args.workload == "fdsolve":
        return fdsolve(args)
    return atom(args)


class Atom:
    """The 128^3 soft-core context with its eigenstate and the trained network."""

    def __init__(self, args, train=True):
        import tdsez_amd as T
        self.T, self.args = T, args
        self.desc = problem(args.nel)
        t0 = time.perf_counter()
        self.ctx = ctx = T.Context(self.desc)
        ctx.set_correction_form(args.correction_form)
        ctx.pc_setup(T.PC_MG)
        psi = ctx.vec()
        ctx.project_gaussian(psi, (0.0, 0.0, 0.0), (1.5, 1.5, 1.5))
        self.E0 = float("nan")
        if args.init == "ground":
            self.E0, res0, it0 = ctx.ground_state(psi, sigma=-0.3, tol=1e-8, maxit=80)
            log(f"ground state E0 = {self.E0:.10f} (eigen-residual {res0:.1e}, {it0} inverse iterations)")
        self.pulse = T.make_pulse(E0=(0.0, 0.0, E0Z), omega=OMEGA, ncycles=NCYC)
        self.x0 = psi.download()
        self.trained = train
        self.learn = None
        self.sweeps = {}
        if not train:
            psi.free()
            self.setup_s = time.perf_counter() - t0
            self.kinds = {"learned": T.PC_LEARNED, "fp64_mg": T.PC_MG, "network_free": T.PC_CORRECTED,
                          "mixed_mg": T.PC_MG_MIXED}
            return
        tl = time.perf_counter()
        l0, l1 = ctx.learn_train(psi, 0.0, DT, self.pulse, harvest_steps=2, max_samples=6, iters=args.train_iters,
                                 lr=3e-3, rtol=args.rtol)
        base_its, _, _ = ctx.learn_status()
        self.learn = {"loss_physics_init": l0, "loss_trained": l1, "train_s": time.perf_counter() - tl,
                      "its_without_network": base_its}
        log(f"trained: loss/sample {l0:.3e} -> {l1:.3e} in {self.learn['train_s']:.1f} s")
        psi.free()
        self.setup_s = time.perf_counter() - t0
        self.kinds = {"learned": T.PC_LEARNED, "fp64_mg": T.PC_MG, "network_free": T.PC_CORRECTED,
                      "mixed_mg": T.PC_MG_MIXED}

    def run(self, leg, x, t0, nsteps, warmup, order, timing=False):
        """warmup untimed + nsteps timed CN steps from state x at t0; returns
        (steps/s, its, final state, potential-kernel ms, launches)."""
        ctx = self.ctx
        ctx.pc_setup(self.kinds[leg])
        ctx.set_cn_predictor(order)
        d = ctx.vec(x)
        if warmup:
            ctx.evolve_cn(d, t0, DT, warmup, self.pulse, rtol=self.args.rtol, maxit=500)
        ctx.sync()
        if timing:
            ctx.timing(True)
        t1 = time.perf_counter()
        its, _ = ctx.evolve_cn(d, t0 + warmup * DT, DT, nsteps, self.pulse, rtol=self.args.rtol, maxit=500)
        ctx.sync()
        el = time.perf_counter() - t1
        kms, nl = ctx.timing_stats() if timing else (0.0, 0)
        ctx.timing(False)
        self.sweeps = ctx.sweep_stats()
        if leg == "learned":
            _, fallbacks, active = ctx.learn_status()
            self.learned_status = {"fallbacks": fallbacks, "network_active": bool(active)}
        out = d.download()
        d.free()
        return nsteps / el, its, out, kms, nl

    def mnorm(self, x):
        T, ctx = self.T, self.ctx
        dx, dy = ctx.vec(x), ctx.vec()
        ctx.term(T.TERM_M, dx, dy)
        v = float(np.sqrt(abs(np.vdot(x, dy.download()).real)))
        dx.free()
        dy.free()
        return v

    def propagate(self, leg, x, t0, nsteps, chunk=1000, on_chunk=None, order=2):
        """Untimed propagation (progress on stderr); returns the state after nsteps."""
        ctx = self.ctx
        ctx.pc_setup(self.kinds[leg])
        ctx.set_cn_predictor(order)
        d = ctx.vec(x)
        done, its_all = 0, []
        while done < nsteps:
            k = min(chunk, nsteps - done)
            its, _ = ctx.evolve_cn(d, t0 + done * DT, DT, k, self.pulse, rtol=self.args.rtol, maxit=500)
            its_all.append(its)
            done += k
            if on_chunk:
                on_chunk(done, d)
            log(f"{leg}: {done}/{nsteps} steps, t = {t0 + done * DT:.2f}, mean its {np.mean(its):.2f}")
        out = d.download()
        d.free()
        return out, np.concatenate(its_all) if its_all else np.zeros(0, np.int32)


def atom(args):
    if args.workload == "fullpulse":
        return fullpulse(args)
    A = Atom(args, train="learned" in args.legs.split(","))
    T, ctx = A.T, A.ctx
    K, W, R = args.steps, args.warmup, args.repeats
    windows = [w for w in args.windows.split(",") if w]
    legs = [g for g in args.legs.split(",") if g]
    roof = None
    states = {}
    for win in windows:
        if win == "early":
            states[win] = (A.x0, 0.0)
        else:
            # the timed steps are centred on the field crest nearest the envelope's peak T/2
            # (omega t = 3.5 pi: |E| = 0.048)
            npre = int(round(T_CREST / DT)) - W - K // 2
            t1 = time.perf_counter()
            xw, _ = A.propagate("learned" if A.trained else "fp64_mg", A.x0, 0.0, npre, chunk=2000, order=2)
            states[win] = (xw, npre * DT)
            log(f"reached t = {npre * DT:.2f} in {time.perf_counter() - t1:.1f} s")
    guesses = [args.predictor] + ([args.alt_predictor] if args.alt_predictor >= 0 and
                                  args.alt_predictor != args.predictor else [])
    res = {}
    for order in guesses:
        key = "current_state" if order == 0 else f"extrapolation_order_{order}"
        res[key] = {}
        for win in windows:
            xw, tw = states[win]
            r = {"t_start": tw + W * DT, "field_at_start": float(ctx.coupling(A.pulse, tw + W * DT)[2])}
            finals = {}
            for leg in legs:
                sps, itsm = [], []
                for rep in range(R):
                    timing = leg == "learned" and rep == 0 and win == windows[-1] and order == args.predictor
                    v, its, out, kms, nl = A.run(leg, xw, tw, K, W, order, timing)
                    sps.append(v)
                    itsm.append(float(np.mean(its)))
                    if rep == 0:
                        finals[leg] = out
                    if timing:
                        roof = (kms, nl)
                r[leg] = {"steps_per_s": float(np.median(sps)), "runs_steps_per_s": sps,
                          "krylov_its_per_step": float(np.mean(itsm))}
                if A.sweeps:   # the corrected legs' sweep/network choices (last run, warmup included)
                    r[leg]["sweep_choices"] = A.sweeps
                log(f"guess {order}, {win}: {leg} {np.median(sps):.2f} steps/s (runs {['%.1f' % v for v in sps]}), "
                    f"its/step {np.mean(itsm):.2f}")
            if "learned" in r and "fp64_mg" in r:
                r["speedup_vs_fp64_mg"] = r["learned"]["steps_per_s"] / r["fp64_mg"]["steps_per_s"]
                r["psi_rel_l2_diff_learned_vs_fp64"] = A.mnorm(finals["learned"] - finals["fp64_mg"]) / \
                    A.mnorm(finals["fp64_mg"])
            res[key][win] = r
    main_key = "current_state" if args.predictor == 0 else f"extrapolation_order_{args.predictor}"
    main_win = windows[-1]
    sps = res[main_key][main_win][legs[0] if "learned" not in legs else "learned"]["steps_per_s"]
    if A.learn is not None:
        A.learn.update(getattr(A, "learned_status", {}))

    p, Q, nel = A.desc["p"], A.desc["nquad"], args.nel
    flops = pot_flops_per_element(p, Q) * nel ** 3
    kms, nlaunch = roof if roof else (0.0, 0)
    avg_s = (kms / 1e3) / max(nlaunch, 1)
    achieved = flops / avg_s if nlaunch else 0.0
    roofline = {"bound": "mfma", "compute_unit": "FP64 VALU (MI355X FP64 vector and matrix peaks are both 78.6)",
                "kernel": "pot3d_kernel + pot3d_sum_kernel (fine-level potential apply)",
                "limiter": "FP64 VALU issue 60% busy, LDS 33% (profiles/pot3d_softcore128_sq_counters.md)",
                "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64, "traffic": None, "launches": nlaunch, "avg_ms": avg_s * 1e3,
                "flops_per_launch": flops}
    tj = os.path.join(ROOT, "profiles", "pot3d_softcore128_traffic.json")
    if nel == 128 and os.path.exists(tj):
        with open(tj) as f:
            tr = json.load(f)
        roofline["traffic"] = tr["traffic_bytes"]
        roofline["traffic_source"] = "profiles/pot3d_softcore128_traffic.json (" + tr["kernel"] + ")"

    cpu = None
    if not args.no_cpu_baseline:
        xw, tw = states[main_win]
        cpu = cpu_baseline(A, xw, tw + W * DT, args)

    line = {
        "metric": "CN steps/s, 3D p=4 128^3 grid: learned+mixed PC vs plain FP64 MG at |r|<1e-10",
        "value": sps, "unit": "steps/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": 1e3 / sps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex128 (fp64 operator/Krylov; complex64 + bf16 preconditioner)",
        "data": "synthetic: lowest eigenstate of H (GPU inverse iteration), E0 = %.8f" % A.E0,
        "config": {"workload": f"3D soft-core atom {nel}^3 elements p=4 Q=8, CN dt={DT}, rtol={args.rtol}, "
                               f"{main_win} window (median of {R} runs of {W}+{K} steps)",
                   "preconditioner": PCS["learned32"], "baseline_preconditioner": PCS["mg"],
                   "cn_guess": ("current state (as TSTHETA)" if args.predictor == 0 else
                                f"binomial extrapolation of the last {args.predictor + 1} states (every leg)"),
                   "dofs": ctx.N, "mg_levels": ctx.levels, "setup_s": A.setup_s},
        "speedup_vs_fp64_mg": res[main_key][main_win].get("speedup_vs_fp64_mg"),
        "runs": res,
        "learned": A.learn,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def fullpulse(args):
    """The whole pulse with FP64 GMG, complex64 GMG and the learned preconditioner: steps/s
    over the run (steps / propagation time, as src/timing.cpp:229-233) and, after every
    pulse cycle, Krylov iterations per step and the wavefunction's L2 difference from the
    FP64 GMG run."""
    A = Atom(args)
    nsteps = int(math.ceil(T_PULSE / DT - 1e-9))
    cyc = [int(round(nsteps * (c + 1) / NCYC)) for c in range(int(NCYC))]
    legs = {}
    ref_states = {}
    order = args.predictor if args.predictor > 0 else max(args.alt_predictor, 0)
    for leg in ("fp64_mg", "mixed_mg", "learned"):
        ctx = A.ctx
        ctx.pc_setup(A.kinds[leg])
        ctx.set_cn_predictor(order)
        d = ctx.vec(A.x0)
        done, per = 0, []
        t_run = 0.0
        for c, end in enumerate(cyc):
            ctx.sync()
            t1 = time.perf_counter()
            its_c = []
            while done < end:
                k = min(1000, end - done)
                its, _ = ctx.evolve_cn(d, done * DT, DT, k, A.pulse, rtol=args.rtol, maxit=500)
                its_c.append(its)
                done += k
            ctx.sync()
            el = time.perf_counter() - t1
            t_run += el
            its_c = np.concatenate(its_c)
            x = d.download()
            entry = {"cycle": c + 1, "t": done * DT, "steps": int(its_c.size), "seconds": el,
                     "krylov_its_per_step": float(its_c.mean()), "krylov_its_max": int(its_c.max())}
            if leg == "fp64_mg":
                ref_states[c] = x
            else:
                entry["psi_rel_l2_diff_vs_fp64"] = A.mnorm(x - ref_states[c]) / A.mnorm(ref_states[c])
            per.append(entry)
            log(f"{leg}: cycle {c + 1} done at t = {done * DT:.2f}: {entry}")
        d.free()
        legs[leg] = {"steps_per_s": nsteps / t_run, "seconds": t_run, "cycles": per}
    _, fallbacks, active = A.ctx.learn_status()
    A.learn.update({"fallbacks": fallbacks, "network_active": bool(active)})
    sps = legs["learned"]["steps_per_s"]
    line = {
        "metric": "CN steps/s over the whole pulse, 3D p=4 128^3: learned+mixed PC vs plain FP64 MG at |r|<1e-10",
        "value": sps, "unit": "steps/s", "n_gpus": 1, "steps": nsteps, "warmup": 0, "ms_per_step": 1e3 / sps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex128 (fp64 operator/Krylov; complex64 + bf16 preconditioner)",
        "data": "synthetic: lowest eigenstate of H (GPU inverse iteration), E0 = %.8f" % A.E0,
        "config": {"workload": f"3D soft-core atom {args.nel}^3 p=4 Q=8, the whole {NCYC:.0f}-cycle pulse "
                               f"({nsteps} CN steps of dt={DT}), rtol={args.rtol}",
                   "cn_guess": ("current state (as TSTHETA)" if order == 0 else
                                f"binomial extrapolation of the last {order + 1} states (every leg)"),
                   "setup_s": A.setup_s},
        "speedup_vs_fp64_mg": sps / legs["fp64_mg"]["steps_per_s"],
        "legs": legs, "learned": A.learn, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    A.ctx.close()


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
    pred = min(args.predictor if args.predictor > 0 else max(args.alt_predictor, 0), 2)
    if args.kiters is None:
        args.kiters = 2 if pred else 4
    b = T.Batch([d for d, _, _ in runs], pulses, dt, args.warmup + args.steps, pc=T.PC_SEPARABLE_MIXED,
                kiters=args.kiters)
    if pred:
        b.set_predictor(pred)
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
                   "cn_guess": ("current state" if pred == 0 else
                                f"binomial extrapolation of the last {pred + 1} states"),
                   "setup_s": setup, "krylov_its_per_step": float(its.mean()),
                   "host_completions": b.fixups() - fix0, "host_completions_warmup": fix0,
                   "max_residual": float(res.max())},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    b.close()


NM = 18.897261   # bohr per nanometre
EV = 1.0 / 27.211386   # hartree per electronvolt


def fdsolve(args):
    """The separable solve of the preconditioners alone on the 128^3 atom's fine level
    (shift 1 + 0.01i): complex64 with the split-FP16 transforms (kind 1) and the FP32 MFMA
    transforms (kind 0), and FP64 with the FP64 MFMA transforms and with the VALU ones. A
    solve is 6 axis transform passes (complex64: the z pass carries the diagonal solve, and
    r is demoted and z promoted; FP64: a separate diagonal pass).
    Prints microseconds per solve (wall clock over --steps solves after --warmup)."""
    import tdsez_amd as T
    desc = problem(args.nel)
    desc.pop("mg_levels")
    ctx = T.Context(desc)
    ctx.set_shift(1.0, 0.01j, (0, 0, 0))
    rng = np.random.default_rng(1)
    r = rng.standard_normal(ctx.N) + 1j * rng.standard_normal(ctx.N)
    dr, z = ctx.vec(r), ctx.vec()
    out = {}
    for name, kind, fdk in (("complex64_split_fp16_lds", T.PC_SEPARABLE_MIXED, 2),
                            ("complex64_split_fp16", T.PC_SEPARABLE_MIXED, 1),
                            ("complex64_single_fp16", T.PC_SEPARABLE_MIXED, 3),
                            ("complex64_fp32_mfma", T.PC_SEPARABLE_MIXED, 0), ("fp64_mfma", T.PC_SEPARABLE, 1),
                            ("fp64_valu", T.PC_SEPARABLE, 0)):
        ctx.pc_setup(kind)
        if kind == T.PC_SEPARABLE:
            ctx.set_fd64_transform(fdk)
        else:
            ctx.set_fd_transform(fdk)
        for _ in range(max(args.warmup, 1)):
            ctx.pc_apply(dr, z)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.pc_apply(dr, z)
        ctx.sync()
        out[name] = 1e6 * (time.perf_counter() - t0) / args.steps
        log(f"{name}: {out[name]:.1f} us per separable solve")
    # the network-free corrected solve with k correction sweeps: k + 1 separable solves and
    # k fused correction passes
    ctx.pc_setup(T.PC_CORRECTED)
    ctx.set_fd_transform(1)
    for k in (1, 2):
        ctx.set_corrections(k)
        for _ in range(max(args.warmup, 1)):
            ctx.pc_apply(dr, z)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.pc_apply(dr, z)
        ctx.sync()
        out[f"corrected_{k}_sweeps"] = 1e6 * (time.perf_counter() - t0) / args.steps
        log(f"corrected, {k} sweeps: {out[f'corrected_{k}_sweeps']:.1f} us per preconditioner apply")
    ctx.set_fd_transform(1)
    ctx.set_fd64_transform(1)
    ctx.close()
    print(json.dumps({"metric": "separable solve time, 128^3 p=4", "unit": "us", "solves": args.steps,
                      "us_per_solve": out, "dofs": int(np.prod([n + desc["p"] for n in desc["nelem"]]))}))


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


def cpu_baseline(A, x, t, args, nsteps=3):
    """The CPU oracle (C++/OpenMP, the reference algorithm restated) on nsteps CN steps of the
    same problem from the state of the timed window (bounded sample), built with
    -march=native on this host when the compiler is there (portable build otherwise)."""
    try:
        r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "native"], capture_output=True, text=True,
                           timeout=300)
        import oracle
        native = r.returncode == 0 and oracle.use_native()
        from oracle import Oracle
    except Exception as e:  # noqa: BLE001
        return {"error": f"oracle unavailable: {e}"}
    o = Oracle(A.desc)
    sE = np.array([A.ctx.coupling(A.pulse, t + (s + 0.5) * DT) for s in range(nsteps)])
    t0 = time.perf_counter()
    _, its, _ = o.cn_steps(x, DT, sE, rtol=args.rtol, maxit=500, restart=50, pc=1)
    el = time.perf_counter() - t0
    return {"value": nsteps / el, "unit": "steps/s", "cores": oracle.lib().orc_threads(), "kind": "port",
            "sample": f"{nsteps} CN steps of the same {A.desc['nelem'][0]}^3 problem from the timed window's start "
                      f"state (t = {t:.2f}), FGMRES(50) + separable fast-diagonalisation preconditioner "
                      f"({its.tolist()} its), {el:.1f} s, "
                      + ("-march=native build" if native else "portable x86-64-v3 build")}


if __name__ == "__main__":
    main()

"""MI355X-native TDSE-Z propagation hot path (Python view of the C ABI).

The product is ``libtdz.so`` (HIP kernels + C++ host code) with the C ABI of
``include/tdz.h``. This module is a thin ctypes binding used by the tests and
the benchmark; it raises if the library is missing instead of falling back to
any CPU path.

Problem descriptions are dicts with the fields of ``tdz_problem``
(include/tdz.h): dim, nelem, lo, hi, p, nquad, dirichlet, mass, charge,
pot_kind, pot, cap_kind, cap, mg_levels, mg_smooth, restart.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtdz.so")

(PC_NONE, PC_SEPARABLE, PC_MG, PC_SEPARABLE_MIXED, PC_MG_MIXED, PC_CORRECTED, PC_CORRECTED_EXACT,
 PC_LEARNED) = range(8)
TERM_M, TERM_K, TERM_V, TERM_CAP, TERM_DX, TERM_DY, TERM_DZ = range(7)
POT_NONE, POT_HARMONIC, POT_SOFTCORE = 0, 1, 2
CAP_NONE, CAP_QUADRATIC, CAP_MANOLOPOULOS = 0, 1, 2
CNN_INPUTS, CNN_CHANNELS = 4, 16   # network input fields (y and its first moments), hidden channels

# every symbol include/tdz.h declares
EXPORTS = [
    "tdz_create", "tdz_destroy", "tdz_last_error", "tdz_size", "tdz_dims", "tdz_levels",
    "tdz_vec_alloc", "tdz_vec_free", "tdz_vec_upload", "tdz_vec_download", "tdz_sync",
    "tdz_set_shift", "tdz_op_apply", "tdz_term_apply", "tdz_pc_setup", "tdz_pc_apply",
    "tdz_ksp_solve", "tdz_step_cn", "tdz_evolve_cn", "tdz_set_cn_predictor", "tdz_batch_set_predictor", "tdz_observables", "tdz_pulse_coupling",
    "tdz_timing", "tdz_timing_stats", "tdz_step_pade22", "tdz_evolve_rational",
    "tdz_project_gaussian", "tdz_ground_state",
    "tdz_trace_row", "tdz_pade_history", "tdz_batch_create", "tdz_batch_context", "tdz_batch_evolve",
    "tdz_batch_fixups", "tdz_batch_destroy", "tdz_ifunction", "tdz_level_dims", "tdz_mg_transfer",
    "tdz_level_apply", "tdz_set_pc_refresh", "tdz_coarsen_knots", "tdz_refinement_matrix", "tdz_last_warning",
    "tdz_set_pade_guess", "tdz_pc_set_corrections", "tdz_cn_predictor_status", "tdz_observables_noalloc", "tdz_mg_set_smoother", "tdz_mg_set_coarse", "tdz_set_fd_transform", "tdz_set_fd64_transform", "tdz_mg_coarse_status", "tdz_level_solve", "tdz_learn_config", "tdz_learn_train", "tdz_learn_weights", "tdz_learn_fields", "tdz_learn_features", "tdz_learn_eval", "tdz_learn_status",
    "tdz_pc_sweep_stats", "tdz_set_correction_form",
]


class Problem(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("nelem", C.c_int * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("p", C.c_int), ("nquad", C.c_int), ("dirichlet", C.c_int),
        ("mass", C.c_double), ("charge", C.c_double),
        ("pot_kind", C.c_int), ("pot", C.c_double * 4),
        ("cap_kind", C.c_int), ("cap", C.c_double * 4),
        ("mg_levels", C.c_int), ("mg_smooth", C.c_int), ("restart", C.c_int),
        ("knots", C.POINTER(C.c_double)), ("nknots", C.c_int), ("nlayers", C.c_int),
        ("layer_bounds", C.POINTER(C.c_double)), ("layer_mass", C.POINTER(C.c_double)),
        ("layer_v", C.POINTER(C.c_double)),
    ]


class KspStats(C.Structure):
    _fields_ = [("its", C.c_int), ("restarts", C.c_int), ("res", C.c_double), ("bnorm", C.c_double),
                ("pc_rebuilds", C.c_int)]


class Pulse(C.Structure):
    _fields_ = [("E0", C.c_double * 3), ("omega", C.c_double), ("ncycles", C.c_double), ("cep", C.c_double)]


def make_problem(desc):
    p = Problem()
    p.dim = desc["dim"]
    for d in range(3):
        p.nelem[d] = desc["nelem"][d] if d < desc["dim"] else 1
        p.lo[d] = desc["lo"][d] if d < desc["dim"] else 0.0
        p.hi[d] = desc["hi"][d] if d < desc["dim"] else 0.0
    p.p = desc["p"]
    p.nquad = desc["nquad"]
    p.dirichlet = desc.get("dirichlet", 0)
    p.mass = desc.get("mass", 1.0)
    p.charge = desc.get("charge", -1.0)
    p.pot_kind = desc.get("pot_kind", 0)
    for i, v in enumerate(desc.get("pot", ())):
        p.pot[i] = v
    p.cap_kind = desc.get("cap_kind", 0)
    for i, v in enumerate(desc.get("cap", ())):
        p.cap[i] = v
    p.mg_levels = desc.get("mg_levels", 0)
    p.mg_smooth = desc.get("mg_smooth", 1)
    p.restart = desc.get("restart", 0)
    # layered z axis: desc["knots"] (knot vector), desc["layers"] = (bounds, masses, potentials)
    keep = []
    if desc.get("knots") is not None:
        k = np.ascontiguousarray(desc["knots"], np.float64)
        keep.append(k)
        p.knots = k.ctypes.data_as(C.POINTER(C.c_double))
        p.nknots = k.size
    if desc.get("layers") is not None:
        b, m, v = (np.ascontiguousarray(a, np.float64) for a in desc["layers"])
        keep += [b, m, v]
        p.nlayers = m.size
        p.layer_bounds = b.ctypes.data_as(C.POINTER(C.c_double))
        p.layer_mass = m.ctypes.data_as(C.POINTER(C.c_double))
        p.layer_v = v.ctypes.data_as(C.POINTER(C.c_double))
    p._keep = keep   # the arrays must outlive tdz_create
    return p


def make_pulse(E0=(0, 0, 0), omega=0.0, ncycles=0.0, cep=0.0):
    pl = Pulse()
    for a in range(3):
        pl.E0[a] = E0[a]
    pl.omega, pl.ncycles, pl.cep = omega, ncycles, cep
    return pl


_lib = None


def lib():
    """Load libtdz.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP extension missing: {LIB_PATH} (run python tdsez_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip, d, i = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_double, C.c_int
        L.tdz_create.argtypes = [C.POINTER(Problem), C.POINTER(vp)]
        L.tdz_destroy.argtypes = [vp]
        L.tdz_last_error.restype = C.c_char_p
        L.tdz_size.argtypes = [vp]
        L.tdz_size.restype = C.c_long
        L.tdz_dims.argtypes = [vp, ip]
        L.tdz_levels.argtypes = [vp]
        L.tdz_vec_alloc.argtypes = [vp, C.POINTER(vp)]
        L.tdz_vec_free.argtypes = [vp, vp]
        L.tdz_vec_upload.argtypes = [vp, vp, dp]
        L.tdz_vec_download.argtypes = [vp, vp, dp]
        L.tdz_sync.argtypes = [vp]
        L.tdz_set_shift.argtypes = [vp, d, d, d, d, dp]
        L.tdz_op_apply.argtypes = [vp, vp, vp]
        L.tdz_term_apply.argtypes = [vp, i, vp, vp]
        L.tdz_pc_setup.argtypes = [vp, i]
        L.tdz_pc_apply.argtypes = [vp, vp, vp]
        L.tdz_ksp_solve.argtypes = [vp, vp, vp, d, i, C.POINTER(KspStats)]
        L.tdz_step_cn.argtypes = [vp, vp, d, dp, d, i, C.POINTER(KspStats)]
        L.tdz_evolve_cn.argtypes = [vp, vp, d, d, i, C.POINTER(Pulse), d, i, dp, ip]
        L.tdz_set_cn_predictor.argtypes = [vp, i]
        L.tdz_batch_set_predictor.argtypes = [vp, i]
        L.tdz_observables.argtypes = [vp, vp, dp]
        L.tdz_pulse_coupling.argtypes = [vp, C.POINTER(Pulse), d, dp]
        L.tdz_step_pade22.argtypes = [vp, vp, d, dp, i, d, i, C.POINTER(KspStats)]
        L.tdz_evolve_rational.argtypes = [vp, vp, d, d, i, C.POINTER(Pulse), i, d, i, dp, ip]
        L.tdz_project_gaussian.argtypes = [vp, dp, dp, dp, vp]
        L.tdz_ground_state.argtypes = [vp, vp, d, d, i, dp, dp, ip]
        fp = C.POINTER(C.c_float)
        L.tdz_trace_row.argtypes = [vp, vp, vp, d, dp, dp, dp, dp]
        L.tdz_batch_create.argtypes = [C.POINTER(Problem), C.POINTER(Pulse), i, d, i, i, i, C.POINTER(vp)]
        L.tdz_batch_context.argtypes = [vp, i, C.POINTER(vp), C.POINTER(vp)]
        L.tdz_batch_evolve.argtypes = [vp, i, d, i, ip, dp]
        L.tdz_batch_fixups.argtypes = [vp, C.POINTER(C.c_long)]
        L.tdz_batch_destroy.argtypes = [vp]
        L.tdz_pade_history.argtypes = [vp, i, ip, dp]
        L.tdz_learn_train.argtypes = [vp, vp, d, d, C.POINTER(Pulse), i, i, i, d, d, dp, dp]
        L.tdz_learn_weights.argtypes = [vp, fp, fp, i]
        L.tdz_learn_fields.argtypes = [vp, fp]
        L.tdz_learn_features.argtypes = [vp, vp, dp]
        L.tdz_learn_eval.argtypes = [vp, dp, dp, dp, fp, fp, dp, dp]
        L.tdz_learn_status.argtypes = [vp, ip, ip, ip]
        L.tdz_pc_sweep_stats.argtypes = [vp, C.POINTER(C.c_long)]
        L.tdz_set_correction_form.argtypes = [vp, i]
        L.tdz_timing.argtypes = [vp, i]
        L.tdz_timing_stats.argtypes = [vp, dp, C.POINTER(C.c_long)]
        L.tdz_ifunction.argtypes = [vp, dp, vp, vp, vp]
        L.tdz_level_dims.argtypes = [vp, i, ip]
        L.tdz_mg_transfer.argtypes = [vp, i, i, vp, vp]
        L.tdz_level_apply.argtypes = [vp, i, vp, vp]
        L.tdz_set_pc_refresh.argtypes = [vp, i]
        L.tdz_coarsen_knots.argtypes = [dp, i, dp, ip]
        L.tdz_refinement_matrix.argtypes = [dp, i, dp, i, i, dp]
        L.tdz_last_warning.restype = C.c_char_p
        L.tdz_set_pade_guess.argtypes = [vp, i, i]
        L.tdz_pc_set_corrections.argtypes = [vp, i]
        L.tdz_learn_config.argtypes = [vp, i, C.c_ulonglong]
        L.tdz_cn_predictor_status.argtypes = [vp, ip, dp]
        L.tdz_observables_noalloc.argtypes = [vp, vp, dp]
        L.tdz_mg_set_smoother.argtypes = [vp, i, i]
        L.tdz_mg_set_coarse.argtypes = [vp, i, C.c_double, i]
        L.tdz_set_fd_transform.argtypes = [vp, i]
        L.tdz_set_fd64_transform.argtypes = [vp, i]
        L.tdz_mg_coarse_status.argtypes = [vp, ip, dp]
        L.tdz_level_solve.argtypes = [vp, i, vp, vp]
        _lib = L
    return _lib


class TdzError(RuntimeError):
    pass


def coarsen_knots(U):
    """One multigrid coarsening of a knot vector (None when nothing can be removed)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    out = np.zeros(U.size)
    m = C.c_int(0)
    _chk(lib().tdz_coarsen_knots(_dp(U), U.size, _dp(out), C.byref(m)))
    return out[:m.value] if m.value else None


def refinement_matrix(Cknots, Fknots, p):
    """Dense knot-insertion matrix R [nf, nc]: coarse B-spline j = sum_i R[i, j] fine B-spline i."""
    Ck = np.ascontiguousarray(Cknots, dtype=np.float64)
    Fk = np.ascontiguousarray(Fknots, dtype=np.float64)
    out = np.zeros((Fk.size - p - 1, Ck.size - p - 1))
    _chk(lib().tdz_refinement_matrix(_dp(Ck), Ck.size, _dp(Fk), Fk.size, p, _dp(out)))
    return out


def last_warning():
    return lib().tdz_last_warning().decode()


def _chk(rc):
    if rc:
        raise TdzError(lib().tdz_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class DeviceVec:
    """N complex128 values in HBM, owned by a Context."""

    def __init__(self, ctx, host=None):
        self.ctx = ctx
        ptr = C.c_void_p()
        _chk(lib().tdz_vec_alloc(ctx.h, C.byref(ptr)))
        self.ptr = ptr
        if host is not None:
            self.upload(host)

    def upload(self, host):
        h = np.ascontiguousarray(host, dtype=np.complex128).reshape(-1)
        assert h.size == self.ctx.N
        _chk(lib().tdz_vec_upload(self.ctx.h, self.ptr, _dp(h.view(np.float64))))

    def download(self):
        out = np.zeros(self.ctx.N, dtype=np.complex128)
        _chk(lib().tdz_vec_download(self.ctx.h, self.ptr, _dp(out.view(np.float64))))
        return out

    def free(self):
        if self.ptr is not None and self.ctx.h:
            _chk(lib().tdz_vec_free(self.ctx.h, self.ptr))
            self.ptr = None


class Context:
    """One discretised problem on cuda:0 behind the C ABI."""

    def __init__(self, desc):
        self.desc = dict(desc)
        self._prob = make_problem(desc)
        h = C.c_void_p()
        _chk(lib().tdz_create(C.byref(self._prob), C.byref(h)))
        self.h = h
        self.N = int(lib().tdz_size(self.h))
        n = (C.c_int * 3)()
        _chk(lib().tdz_dims(self.h, n))
        self.n = tuple(n)
        self.levels = int(lib().tdz_levels(self.h))
        self.guess_order = 3

    def close(self):
        if getattr(self, "h", None):
            lib().tdz_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def vec(self, host=None):
        return DeviceVec(self, host)

    def set_shift(self, alpha, g, sE=(0, 0, 0)):
        s = np.ascontiguousarray(sE, dtype=np.float64)
        _chk(lib().tdz_set_shift(self.h, complex(alpha).real, complex(alpha).imag, complex(g).real,
                                 complex(g).imag, _dp(s)))

    def apply(self, x, y):
        _chk(lib().tdz_op_apply(self.h, x.ptr, y.ptr))
        _chk(lib().tdz_sync(self.h))

    def term(self, t, x, y):
        _chk(lib().tdz_term_apply(self.h, t, x.ptr, y.ptr))

    def pc_setup(self, kind):
        _chk(lib().tdz_pc_setup(self.h, kind))

    def pc_apply(self, r, z):
        _chk(lib().tdz_pc_apply(self.h, r.ptr, z.ptr))

    def solve(self, b, x, rtol=1e-12, maxit=1000):
        st = KspStats()
        _chk(lib().tdz_ksp_solve(self.h, b.ptr, x.ptr, rtol, maxit, C.byref(st)))
        return st

    def step_cn(self, psi, dt, sE_mid=(0, 0, 0), rtol=1e-12, maxit=1000):
        st = KspStats()
        s = np.ascontiguousarray(sE_mid, dtype=np.float64)
        _chk(lib().tdz_step_cn(self.h, psi.ptr, dt, _dp(s), rtol, maxit, C.byref(st)))
        return st

    def ifunction(self, sE, psi, psidot, F):
        """F = i M psidot - H psi - sE.D psi (the TS IFunction)."""
        s = np.ascontiguousarray(sE, dtype=np.float64)
        _chk(lib().tdz_ifunction(self.h, _dp(s), psi.ptr, psidot.ptr, F.ptr))

    def level_dims(self, level):
        n = (C.c_int * 3)()
        _chk(lib().tdz_level_dims(self.h, level, n))
        return tuple(n)

    def mg_transfer(self, level, direction, x, y):
        """direction 0: y = P^T x (level -> level+1); 1: y = P x (level+1 -> level)."""
        _chk(lib().tdz_mg_transfer(self.h, level, direction, x.ptr, y.ptr))

    def level_apply(self, level, x, y):
        _chk(lib().tdz_level_apply(self.h, level, x.ptr, y.ptr))

    def learn_config(self, random_samples=6, seed=12345):
        _chk(lib().tdz_learn_config(self.h, int(random_samples), int(seed)))

    def set_corrections(self, sweeps):
        _chk(lib().tdz_pc_set_corrections(self.h, int(sweeps)))

    def set_smoother(self, kind, degree=2):
        """Multigrid smoother: 0 separable-solve Richardson (default), 1 Jacobi-Chebyshev."""
        _chk(lib().tdz_mg_set_smoother(self.h, int(kind), int(degree)))

    def set_fd64_transform(self, kind):
        """FP64 separable solves: 1 FP64 MFMA transforms (default), 0 VALU."""
        _chk(lib().tdz_set_fd64_transform(self.h, int(kind)))

    def set_fd_transform(self, kind):
        """Complex64 separable-solve transforms: 0 FP32 MFMA, 1 split FP16 MFMA (default)."""
        _chk(lib().tdz_set_fd_transform(self.h, int(kind)))

    def set_coarse_solve(self, kind, rtol=1e-12, maxit=64):
        """Coarsest multigrid level: 0 separable sweeps (default), 1 Krylov solve to rtol."""
        _chk(lib().tdz_mg_set_coarse(self.h, int(kind), float(rtol), int(maxit)))

    def level_solve(self, level, b, x):
        """x = the coarse solve applied to b on the coarsest level (device vectors)."""
        _chk(lib().tdz_level_solve(self.h, int(level), b.ptr, x.ptr))

    def coarse_status(self):
        """(iterations, residual estimate) of the last Krylov coarse solve."""
        its, res = C.c_int(0), C.c_double(0.0)
        _chk(lib().tdz_mg_coarse_status(self.h, C.byref(its), C.byref(res)))
        return its.value, res.value

    def set_pc_refresh(self, max_its):
        _chk(lib().tdz_set_pc_refresh(self.h, int(max_its)))

    def set_pade_guess(self, extrapolate=True, order=3):
        _chk(lib().tdz_set_pade_guess(self.h, int(bool(extrapolate)), int(order)))
        self.guess_order = int(order)

    def predictor_status(self):
        """(interval of the exact recomputation of the stored products, drift found at the last)"""
        k, dr = C.c_int(0), C.c_double(0)
        _chk(lib().tdz_cn_predictor_status(self.h, C.byref(k), C.byref(dr)))
        return k.value, dr.value

    def set_cn_predictor(self, order):
        """Initial guess of the CN solves: 0 the current state, 1/2 the binomial extrapolation."""
        _chk(lib().tdz_set_cn_predictor(self.h, int(order)))

    def evolve_cn(self, psi, t0, dt, nsteps, pulse=None, rtol=1e-12, maxit=1000, observe=False):
        its = np.zeros(nsteps, dtype=np.int32)
        obs = np.zeros((nsteps, 7)) if observe else None
        _chk(lib().tdz_evolve_cn(self.h, psi.ptr, t0, dt, nsteps, C.byref(pulse) if pulse is not None else None,
                                 rtol, maxit, _dp(obs) if observe else None,
                                 its.ctypes.data_as(C.POINTER(C.c_int))))
        return its, obs

    def step_pade22(self, psi, h, sE=(0, 0, 0), slot=0, rtol=1e-12, maxit=1000):
        st = KspStats()
        s = np.ascontiguousarray(sE, dtype=np.float64)
        _chk(lib().tdz_step_pade22(self.h, psi.ptr, h, _dp(s), slot, rtol, maxit, C.byref(st)))
        return st

    def evolve_rational(self, psi, t0, dt, nsteps, pulse=None, magnus=True, rtol=1e-12, maxit=1000,
                        observe=False):
        its = np.zeros(nsteps, dtype=np.int32)
        obs = np.zeros((nsteps, 7)) if observe else None
        _chk(lib().tdz_evolve_rational(self.h, psi.ptr, t0, dt, nsteps,
                                       C.byref(pulse) if pulse is not None else None, 1 if magnus else 0, rtol,
                                       maxit, _dp(obs) if observe else None,
                                       its.ctypes.data_as(C.POINTER(C.c_int))))
        return its, obs

    def project_gaussian(self, psi, centre, width, momentum=(0, 0, 0)):
        c = np.ascontiguousarray(list(centre) + [0.0] * (3 - len(centre)), dtype=np.float64)
        w = np.ascontiguousarray(list(width) + [1.0] * (3 - len(width)), dtype=np.float64)
        k = np.ascontiguousarray(list(momentum) + [0.0] * (3 - len(momentum)), dtype=np.float64)
        _chk(lib().tdz_project_gaussian(self.h, _dp(c), _dp(w), _dp(k), psi.ptr))

    def ground_state(self, psi, sigma, tol=1e-9, maxit=100):
        E, r, it = C.c_double(0), C.c_double(0), C.c_int(0)
        _chk(lib().tdz_ground_state(self.h, psi.ptr, sigma, tol, maxit, C.byref(E), C.byref(r), C.byref(it)))
        return E.value, r.value, it.value

    # ---- learned correction
    def learn_train(self, psi, t0, dt, pulse=None, harvest_steps=2, max_samples=6, iters=200, lr=3e-3,
                    rtol=1e-10):
        l0, l1 = C.c_double(0), C.c_double(0)
        _chk(lib().tdz_learn_train(self.h, psi.ptr, t0, dt, C.byref(pulse) if pulse is not None else None,
                                   harvest_steps, max_samples, iters, lr, rtol, C.byref(l0), C.byref(l1)))
        return l0.value, l1.value

    def learn_weights(self, W1=None, W2=None):
        fp = C.POINTER(C.c_float)
        if W1 is None:
            W1 = np.zeros((27, CNN_INPUTS, 16), np.float32)
            W2 = np.zeros((27, 16, 16), np.float32)
            _chk(lib().tdz_learn_weights(self.h, W1.ctypes.data_as(fp), W2.ctypes.data_as(fp), 0))
            return W1, W2
        W1 = np.ascontiguousarray(W1, np.float32)
        W2 = np.ascontiguousarray(W2, np.float32)
        _chk(lib().tdz_learn_weights(self.h, W1.ctypes.data_as(fp), W2.ctypes.data_as(fp), 1))

    def learn_fields(self):
        P = np.zeros((self.N, 16), np.float32)
        _chk(lib().tdz_learn_fields(self.h, P.ctypes.data_as(C.POINTER(C.c_float))))
        return P

    def learn_features(self, y):
        """Network input [y, G_x, G_y, G_z] (complex64 arithmetic) of a device vector y."""
        u = np.zeros((CNN_INPUTS, self.N), np.complex128)
        _chk(lib().tdz_learn_features(self.h, y.ptr, _dp(u.view(np.float64))))
        return u

    def learn_eval(self, u, target):
        """Loss, weight gradients, FP32 and BF16-MFMA outputs for one (u [4][N], t [N]) pair."""
        fp = C.POINTER(C.c_float)
        loss = C.c_double(0)
        u = np.ascontiguousarray(u, np.complex128).reshape(CNN_INPUTS, self.N)
        t = np.ascontiguousarray(target, np.complex128).reshape(self.N)
        dW1 = np.zeros((27, CNN_INPUTS, 16), np.float32)
        dW2 = np.zeros((27, 16, 16), np.float32)
        of = np.zeros(2 * self.N)
        om = np.zeros(2 * self.N)
        _chk(lib().tdz_learn_eval(self.h, _dp(u.view(np.float64)), _dp(t.view(np.float64)), C.byref(loss),
                                  dW1.ctypes.data_as(fp), dW2.ctypes.data_as(fp), _dp(of), _dp(om)))
        return loss.value, dW1, dW2, of, om

    def learn_status(self):
        b, f, a = C.c_int(0), C.c_int(0), C.c_int(0)
        _chk(lib().tdz_learn_status(self.h, C.byref(b), C.byref(f), C.byref(a)))
        return b.value, f.value, a.value

    def set_correction_form(self, form):
        _chk(lib().tdz_set_correction_form(self.h, int(form)))

    def sweep_stats(self):
        """Solves per correction-sweep choice since pc_setup: {"k1": n, ..., "net_k3": n}."""
        c = (C.c_long * 8)()
        _chk(lib().tdz_pc_sweep_stats(self.h, c))
        return {("net_" if i >= 4 else "") + f"k{i % 4}": int(c[i]) for i in range(8) if i % 4 and c[i]}

    def trace_row(self, psi, t, E=(0.0, 0.0, 0.0), psi0=None):
        """(dipoles row, energies row, autocorrelation row) in the observables-file layout."""
        dim = self.desc["dim"]
        dip, en, ac = np.zeros(1 + 3 * dim), np.zeros(7), np.zeros(3)
        Ea = np.ascontiguousarray(list(E) + [0.0] * (3 - len(E)), dtype=np.float64)
        _chk(lib().tdz_trace_row(self.h, psi.ptr, psi0.ptr if psi0 is not None else None, t, _dp(Ea), _dp(dip),
                                 _dp(en), _dp(ac)))
        return dip, en, ac

    def pade_history(self, counts=None, hist=None):
        """Get (no arguments) or set the Pade/Magnus guess history: counts [4], hist [4, order, N]."""
        ip = C.POINTER(C.c_int)
        if counts is None:
            cnt = np.zeros(4, np.int32)
            h = np.zeros((4, self.guess_order, self.N), np.complex128)
            _chk(lib().tdz_pade_history(self.h, 0, cnt.ctypes.data_as(ip), _dp(h.view(np.float64))))
            return cnt, h
        cnt = np.ascontiguousarray(counts, np.int32)
        h = np.ascontiguousarray(hist, np.complex128)
        _chk(lib().tdz_pade_history(self.h, 1, cnt.ctypes.data_as(ip), _dp(h.view(np.float64))))

    def observables(self, psi):
        out = np.zeros(6)
        _chk(lib().tdz_observables(self.h, psi.ptr, _dp(out)))
        return out

    def observables_noalloc(self, psi):
        """The same as observables(), from the allocation-free entry point."""
        out = np.zeros(6)
        _chk(lib().tdz_observables_noalloc(self.h, psi.ptr, _dp(out)))
        return out

    def coupling(self, pulse, t):
        out = np.zeros(3)
        _chk(lib().tdz_pulse_coupling(self.h, C.byref(pulse), t, _dp(out)))
        return out

    def sync(self):
        _chk(lib().tdz_sync(self.h))

    def timing(self, enable):
        _chk(lib().tdz_timing(self.h, 1 if enable else 0))

    def timing_stats(self):
        ms, n = C.c_double(0), C.c_long(0)
        _chk(lib().tdz_timing_stats(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class _BorrowedVec(DeviceVec):
    """A device vector owned by a batch (not freed by Python)."""

    def __init__(self, ctx, ptr):
        self.ctx = ctx
        self.ptr = ptr

    def free(self):
        self.ptr = None


class _BorrowedContext(Context):
    """A run's context inside a batch (destroyed with the batch)."""

    def __init__(self, desc, h):
        self.desc = dict(desc)
        self.h = h
        self.N = int(lib().tdz_size(self.h))
        n = (C.c_int * 3)()
        _chk(lib().tdz_dims(self.h, n))
        self.n = tuple(n)
        self.levels = int(lib().tdz_levels(self.h))

    def close(self):
        self.h = None


class Batch:
    """Many independent runs advanced by one captured HIP graph per step (tdz_batch_*)."""

    def __init__(self, descs, pulses, dt, nsteps, pc=PC_SEPARABLE_MIXED, kiters=4):
        n = len(descs)
        made = [make_problem(dd) for dd in descs]   # keeps any knot/layer arrays alive
        probs = (Problem * n)(*made)
        pls = (Pulse * n)(*pulses) if pulses is not None else None
        h = C.c_void_p()
        _chk(lib().tdz_batch_create(probs, pls, n, dt, nsteps, pc, kiters, C.byref(h)))
        self.h = h
        self.nruns = n
        self.runs, self.psi = [], []
        for r in range(n):
            ch, ph = C.c_void_p(), C.c_void_p()
            _chk(lib().tdz_batch_context(self.h, r, C.byref(ch), C.byref(ph)))
            ctx = _BorrowedContext(descs[r], ch)
            self.runs.append(ctx)
            self.psi.append(_BorrowedVec(ctx, ph))

    def evolve(self, nsteps, rtol=1e-10, maxit=500):
        its = np.zeros((nsteps, self.nruns), np.int32)
        res = np.zeros((nsteps, self.nruns))
        _chk(lib().tdz_batch_evolve(self.h, nsteps, rtol, maxit, its.ctypes.data_as(C.POINTER(C.c_int)), _dp(res)))
        return its, res

    def fixups(self):
        f = C.c_long(0)
        _chk(lib().tdz_batch_fixups(self.h, C.byref(f)))
        return f.value

    def set_predictor(self, order):
        """Initial guess of the captured steps (before the first evolve): 0 the state,
        1/2 the binomial extrapolation of the last 2/3 states."""
        _chk(lib().tdz_batch_set_predictor(self.h, int(order)))

    def close(self):
        if getattr(self, "h", None):
            lib().tdz_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

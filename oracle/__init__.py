"""CPU oracle of the TDSE-Z propagation hot path (TEST INFRASTRUCTURE ONLY).

ctypes binding of ``oracle/liboracle.so`` (built from ``oracle/tdz_oracle.cpp``
by ``oracle/Makefile``). Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker: the product path (``tdsez_amd``) never loads it.

Every function follows a reference routine, cited in ``tdz_oracle.cpp``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
# -march=native build of the same source, made on the host that runs it (make -C oracle native)
NATIVE_PATH = os.path.join(_HERE, "_native", "liboracle.so")


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("nelem", C.c_int * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("p", C.c_int), ("nquad", C.c_int), ("dirichlet", C.c_int),
        ("mass", C.c_double), ("charge", C.c_double),
        ("pot_kind", C.c_int), ("pot", C.c_double * 4),
        ("cap_kind", C.c_int), ("cap", C.c_double * 4),
        ("knots", C.POINTER(C.c_double)), ("nknots", C.c_int), ("nlayers", C.c_int),
        ("layer_bounds", C.POINTER(C.c_double)), ("layer_mass", C.POINTER(C.c_double)),
        ("layer_v", C.POINTER(C.c_double)),
    ]


def make_params(desc):
    """Build the C struct from a problem dict (keys as in ``Params``)."""
    p = Params()
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
    # layered axis (the last active one): optional knot vector and layers
    # desc["knots"]: knot vector; desc["layers"]: (bounds [n+1], masses [n], potentials [n])
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
    p._keep = keep   # the arrays must outlive the create call
    return p


_lib = None


def use_native():
    """Load the -march=native build (built on this host) instead of the portable one; must
    come before the first use of the library. Returns False when it is not built."""
    global LIB_PATH
    if _lib is not None or not os.path.exists(NATIVE_PATH):
        return False
    LIB_PATH = NATIVE_PATH
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip, d, i = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_double, C.c_int
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(Params)]
        L.orc_destroy.argtypes = [vp]
        L.orc_n.argtypes = [vp, i]
        L.orc_size.argtypes = [vp]
        L.orc_size.restype = C.c_long
        L.orc_gauss.argtypes = [i, dp, dp]
        L.orc_knots.argtypes = [vp, i, dp]
        L.orc_axis_tables.argtypes = [vp, i, dp, dp, dp, dp, dp]
        L.orc_band.argtypes = [vp, i, i, dp]
        L.orc_apply_term.argtypes = [vp, i, dp, dp]
        L.orc_apply.argtypes = [vp, d, d, d, d, dp, dp, dp]
        L.orc_dense.argtypes = [vp, i, dp]
        L.orc_solve.argtypes = [vp, d, d, d, d, dp, dp, dp, d, i, i, i, ip, dp]
        L.orc_fd_apply.argtypes = [vp, d, d, d, d, dp, dp]
        L.orc_cn_steps.argtypes = [vp, dp, i, d, dp, d, i, i, i, ip, dp]
        L.orc_pade22_step.argtypes = [vp, dp, d, dp, d, i, i, i, ip]
        L.orc_observables.argtypes = [vp, dp, dp]
        L.orc_project_gaussian.argtypes = [vp, dp, dp, dp, dp]
        L.orc_coarsen_knots.argtypes = [dp, i, dp]
        L.orc_prolongation.argtypes = [dp, i, dp, i, i, dp]
        L.orc_greville.argtypes = [dp, i, i, dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def gauss(Q):
    x = np.zeros(Q)
    w = np.zeros(Q)
    lib().orc_gauss(Q, _p(x), _p(w))
    return x, w


def coarsen_knots(U):
    U = np.ascontiguousarray(U, dtype=np.float64)
    out = np.zeros(len(U))
    m = lib().orc_coarsen_knots(_p(U), len(U), _p(out))
    return out[:m] if m else None


def prolongation(Cknots, Fknots, p):
    Cknots = np.ascontiguousarray(Cknots, dtype=np.float64)
    Fknots = np.ascontiguousarray(Fknots, dtype=np.float64)
    nf, nc = len(Fknots) - p - 1, len(Cknots) - p - 1
    out = np.zeros((nf, nc))
    rc = lib().orc_prolongation(_p(Cknots), len(Cknots), _p(Fknots), len(Fknots), p, _p(out))
    if rc:
        raise RuntimeError("prolongation failed")
    return out


def greville(U, p):
    U = np.ascontiguousarray(U, dtype=np.float64)
    out = np.zeros(len(U) - p - 1)
    lib().orc_greville(_p(U), len(U), p, _p(out))
    return out


class Oracle:
    """One discretised problem on the CPU (complex128 vectors, x fastest)."""

    TERM_M, TERM_K, TERM_V, TERM_CAP, TERM_DX, TERM_DY, TERM_DZ = range(7)

    def __init__(self, desc):
        self.desc = dict(desc)
        self._prm = make_params(desc)
        self.h = lib().orc_create(C.byref(self._prm))
        if not self.h:
            raise ValueError("invalid problem description")
        self.n = tuple(lib().orc_n(self.h, d) for d in range(3))
        self.N = int(lib().orc_size(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def _c(self, v):
        v = np.ascontiguousarray(v, dtype=np.complex128)
        assert v.size == self.N
        return v

    def knots(self, d):
        p = self.desc["p"]
        out = np.zeros(self.desc["nelem"][d] + 2 * p + 1)
        lib().orc_knots(self.h, d, _p(out))
        return out

    def axis_tables(self, d):
        E, p, Q = self.desc["nelem"][d], self.desc["p"], self.desc["nquad"]
        pts, wts, cap = np.zeros(E * Q), np.zeros(E * Q), np.zeros(E * Q)
        N0, N1 = np.zeros(E * (p + 1) * Q), np.zeros(E * (p + 1) * Q)
        lib().orc_axis_tables(self.h, d, _p(pts), _p(wts), _p(N0), _p(N1), _p(cap))
        return dict(pts=pts, wts=wts, N0=N0.reshape(E, p + 1, Q), N1=N1.reshape(E, p + 1, Q), cap=cap)

    def band(self, d, which):
        p = self.desc["p"]
        out = np.zeros(self.n[d] * (2 * p + 1))
        lib().orc_band(self.h, d, which, _p(out))
        return out.reshape(self.n[d], 2 * p + 1)

    def term(self, t, x):
        x = self._c(x)
        y = np.zeros_like(x)
        lib().orc_apply_term(self.h, t, _p(x.view(np.float64)), _p(y.view(np.float64)))
        return y

    def apply(self, alpha, g, sE, x):
        x = self._c(x)
        y = np.zeros_like(x)
        sE = np.ascontiguousarray(sE, dtype=np.float64)
        lib().orc_apply(self.h, alpha.real, alpha.imag, g.real, g.imag, _p(sE), _p(x.view(np.float64)),
                        _p(y.view(np.float64)))
        return y

    def dense(self, t):
        A = np.zeros((self.N, self.N), dtype=np.complex128)
        if lib().orc_dense(self.h, t, _p(A.view(np.float64))):
            raise ValueError("grid too large for the dense assembler")
        return A

    def solve(self, alpha, g, sE, b, x0=None, rtol=1e-12, maxit=1000, restart=30, pc=1):
        b = self._c(b)
        x = self._c(np.zeros(self.N) if x0 is None else x0).copy()
        sE = np.ascontiguousarray(sE, dtype=np.float64)
        its, res = C.c_int(0), C.c_double(0)
        rc = lib().orc_solve(self.h, alpha.real, alpha.imag, g.real, g.imag, _p(sE), _p(b.view(np.float64)),
                             _p(x.view(np.float64)), rtol, maxit, restart, pc, C.byref(its), C.byref(res))
        return x, its.value, res.value, rc

    def fd_apply(self, alpha, g, r):
        r = self._c(r)
        z = np.zeros_like(r)
        lib().orc_fd_apply(self.h, alpha.real, alpha.imag, g.real, g.imag, _p(r.view(np.float64)),
                           _p(z.view(np.float64)))
        return z

    def cn_steps(self, psi, dt, sE_mid, rtol=1e-12, maxit=1000, restart=30, pc=1):
        psi = self._c(psi).copy()
        sE_mid = np.ascontiguousarray(np.asarray(sE_mid, dtype=np.float64).reshape(-1, 3))
        ns = sE_mid.shape[0]
        its = np.zeros(ns, dtype=np.int32)
        res = np.zeros(ns)
        rc = lib().orc_cn_steps(self.h, _p(psi.view(np.float64)), ns, dt, _p(sE_mid), rtol, maxit, restart, pc,
                                its.ctypes.data_as(C.POINTER(C.c_int)), _p(res))
        if rc:
            raise RuntimeError(f"oracle CN step {rc - 1} did not converge")
        return psi, its, res

    def pade22_step(self, psi, h, sE, rtol=1e-12, maxit=1000, restart=30, pc=1):
        psi = self._c(psi).copy()
        sE = np.ascontiguousarray(sE, dtype=np.float64)
        its = np.zeros(2, dtype=np.int32)
        rc = lib().orc_pade22_step(self.h, _p(psi.view(np.float64)), h, _p(sE), rtol, maxit, restart, pc,
                                   its.ctypes.data_as(C.POINTER(C.c_int)))
        if rc:
            raise RuntimeError("oracle Pade stage did not converge")
        return psi, its

    def observables(self, psi):
        psi = self._c(psi)
        out = np.zeros(6)
        lib().orc_observables(self.h, _p(psi.view(np.float64)), _p(out))
        return out

    def project_gaussian(self, centre, width, momentum=(0, 0, 0)):
        c = np.ascontiguousarray(list(centre) + [0.0] * (3 - len(centre)), dtype=np.float64)
        w = np.ascontiguousarray(list(width) + [1.0] * (3 - len(width)), dtype=np.float64)
        k = np.ascontiguousarray(list(momentum) + [0.0] * (3 - len(momentum)), dtype=np.float64)
        out = np.zeros(self.N, dtype=np.complex128)
        lib().orc_project_gaussian(self.h, _p(c), _p(w), _p(k), _p(out.view(np.float64)))
        return out

"""Restart files and observables traces (HDF5, libtdz_io.so) and the CN run driver
that writes them.

The file layouts follow the reference: checkpoints as written by
TDSEZCheckpointer (src/checkpoint.cpp:101-216: root attributes version, step,
time, time_step, dofs, propagator; dataset psi [dofs][2]; optional Pade/Magnus
guess history), traces as written by the observables writer
(src/manager.cpp:51-201, columns of src/monitor.cpp:1281-1330: dipoles,
energies, autocorrelation).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
IO_PATH = os.path.join(HERE, "libtdz_io.so")
PROPAGATORS = {"cn": 0, "pade22": 1, "magnus4": 2}

_io = None


class IOError_(RuntimeError):
    pass


def iolib():
    """Load libtdz_io.so; raises when it has not been built (no HDF5 at build time)."""
    global _io
    if _io is None:
        if not os.path.exists(IO_PATH):
            raise RuntimeError(f"HDF5 IO library missing: {IO_PATH} (python tdsez_amd/build.py)")
        L = C.CDLL(IO_PATH)
        dp, ll, d, i = C.POINTER(C.c_double), C.c_longlong, C.c_double, C.c_int
        L.tdz_io_last_error.restype = C.c_char_p
        L.tdz_io_checkpoint_write.argtypes = [C.c_char_p, ll, d, d, i, ll, dp, i, i, C.POINTER(C.c_int),
                                              C.POINTER(dp)]
        L.tdz_io_checkpoint_read.argtypes = [C.c_char_p, ll, d, i, C.POINTER(ll), C.POINTER(d), dp, i, i,
                                             C.POINTER(C.c_int), C.POINTER(dp)]
        L.tdz_io_trace_open.argtypes = [C.c_char_p, i, i, C.POINTER(C.c_void_p)]
        L.tdz_io_trace_append.argtypes = [C.c_void_p, ll, dp, dp, dp]
        L.tdz_io_trace_close.argtypes = [C.c_void_p]
        L.tdz_io_trace_read.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(ll), C.POINTER(i), dp]
        _io = L
    return _io


def _chk(rc):
    if rc:
        raise IOError_(iolib().tdz_io_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ------------------------------------------------------------------ checkpoints
def write_checkpoint_host(path, psi, step, time, dt, propagator=0, guess_counts=None, guess=None):
    """psi: complex (dofs,); guess: complex (slots, order, dofs) with guess_counts (slots,)."""
    psi = np.ascontiguousarray(psi, np.complex128)
    nslots = order = 0
    cnt_p = None
    ptrs = None
    if guess is not None:
        guess = np.ascontiguousarray(guess, np.complex128)
        nslots, order = guess.shape[0], guess.shape[1]
        cnt = np.ascontiguousarray(guess_counts, np.int32)
        cnt_p = cnt.ctypes.data_as(C.POINTER(C.c_int))
        ptrs = (C.POINTER(C.c_double) * (nslots * order))(
            *[_dp(guess[s, j].view(np.float64)) for s in range(nslots) for j in range(order)])
    _chk(iolib().tdz_io_checkpoint_write(os.fsencode(path), step, time, dt, propagator, psi.size,
                                         _dp(psi.view(np.float64)), nslots, order, cnt_p, ptrs))


def read_checkpoint_host(path, dofs, dt, propagator=0, nslots=0, order=0):
    """Returns (psi, step, time, guess_counts, guess) (the last two None without slots)."""
    psi = np.zeros(dofs, np.complex128)
    step, t = C.c_longlong(0), C.c_double(0)
    cnt = guess = ptrs = cnt_p = None
    if nslots:
        cnt = np.zeros(nslots, np.int32)
        guess = np.zeros((nslots, order, dofs), np.complex128)
        cnt_p = cnt.ctypes.data_as(C.POINTER(C.c_int))
        ptrs = (C.POINTER(C.c_double) * (nslots * order))(
            *[_dp(guess[s, j].view(np.float64)) for s in range(nslots) for j in range(order)])
    _chk(iolib().tdz_io_checkpoint_read(os.fsencode(path), dofs, dt, propagator, C.byref(step), C.byref(t),
                                        _dp(psi.view(np.float64)), nslots, order, cnt_p, ptrs))
    return psi, step.value, t.value, cnt, guess


def write_checkpoint(path, ctx, psi, step, time, dt, propagator="cn"):
    """Checkpoint a device state (and, for pade22/magnus4, the stage guess history)."""
    code = PROPAGATORS[propagator]
    if code == 0:
        write_checkpoint_host(path, psi.download(), step, time, dt, 0)
    else:
        cnt, hist = ctx.pade_history()
        write_checkpoint_host(path, psi.download(), step, time, dt, code, cnt, hist)


def read_checkpoint(path, ctx, psi, dt, propagator="cn"):
    """Restore a device state from a checkpoint; returns (step, time)."""
    code = PROPAGATORS[propagator]
    if code == 0:
        x, step, t, _, _ = read_checkpoint_host(path, ctx.N, dt, 0)
    else:
        x, step, t, cnt, hist = read_checkpoint_host(path, ctx.N, dt, code, 4, ctx.guess_order)
        ctx.pade_history(cnt, hist)
    psi.upload(x)
    return step, t


# ------------------------------------------------------------------ traces
class TraceWriter:
    """Observables file: rows of dipoles / energies / autocorrelation."""

    def __init__(self, path, dim, append=False):
        h = C.c_void_p()
        _chk(iolib().tdz_io_trace_open(os.fsencode(path), dim, 1 if append else 0, C.byref(h)))
        self.h = h
        self.dim = dim

    def append(self, dipoles, energies, autocorr):
        dip = np.ascontiguousarray(dipoles, np.float64).reshape(-1, 1 + 3 * self.dim)
        en = np.ascontiguousarray(energies, np.float64).reshape(-1, 7)
        ac = np.ascontiguousarray(autocorr, np.float64).reshape(-1, 3)
        _chk(iolib().tdz_io_trace_append(self.h, dip.shape[0], _dp(dip), _dp(en), _dp(ac)))

    def close(self):
        if self.h:
            _chk(iolib().tdz_io_trace_close(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_trace(path, dataset):
    rows, cols = C.c_longlong(0), C.c_int(0)
    _chk(iolib().tdz_io_trace_read(os.fsencode(path), dataset.encode(), C.byref(rows), C.byref(cols), None))
    out = np.zeros((rows.value, cols.value))
    _chk(iolib().tdz_io_trace_read(os.fsencode(path), dataset.encode(), C.byref(rows), C.byref(cols), _dp(out)))
    return out


# ------------------------------------------------------------------ run driver
def run_cn(ctx, psi, dt, nsteps, pulse=None, rtol=1e-10, trace=None, output_stride=1, checkpoint=None,
           checkpoint_stride=0, restart=False, psi0=None):
    """CN run in the manner of the reference's Evolve (src/propagator.cpp:78-251): steps
    in blocks of output_stride (each block one chained evolve call), one trace row per
    block, a checkpoint every checkpoint_stride steps. With restart the state, step and
    time come from the checkpoint and the trace file is extended. Returns (step, time,
    Krylov iterations per step)."""
    import tdsez_amd as T  # noqa: F401
    step, t = 0, 0.0
    if restart:
        step, t = read_checkpoint(checkpoint, ctx, psi, dt, "cn")
    writer = TraceWriter(trace, ctx.desc["dim"], append=restart) if trace else None
    its_all = []

    def field(tt):
        if pulse is None:
            return np.zeros(3)
        s = -ctx.desc.get("charge", -1.0)   # the coupling is s E with s = -charge
        return ctx.coupling(pulse, tt) / s if s != 0 else np.zeros(3)

    try:
        if writer is not None and not restart:
            writer.append(*ctx.trace_row(psi, t, field(t), psi0))
        while step < nsteps:
            m = min(output_stride, nsteps - step)
            its, _ = ctx.evolve_cn(psi, t, dt, m, pulse, rtol=rtol)
            its_all.extend(its.tolist())
            step += m
            t = step * dt
            if writer is not None:
                writer.append(*ctx.trace_row(psi, t, field(t), psi0))
            if checkpoint and checkpoint_stride and step % checkpoint_stride == 0 and step < nsteps:
                write_checkpoint(checkpoint, ctx, psi, step, t, dt, "cn")
    finally:
        if writer is not None:
            writer.close()
    return step, t, np.array(its_all, dtype=np.int32)

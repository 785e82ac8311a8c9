"""HDF5 restart files and observables traces (host library, no GPU): layouts of the
reference's checkpoint (src/checkpoint.cpp:101-216) and observables file
(src/manager.cpp:51-201), round trips and the restart checks."""
import os
import shutil
import subprocess

import numpy as np
import pytest

io = pytest.importorskip("tdsez_amd.io")
if not os.path.exists(io.IO_PATH):
    pytest.skip("libtdz_io.so not built (no HDF5)", allow_module_level=True)


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def test_checkpoint_round_trip_and_layout(tmp_path):
    p = str(tmp_path / "checkpoint.h5")
    psi = _rand(1000, 1)
    io.write_checkpoint_host(p, psi, 123, 2.46, 0.02, 0)
    assert not os.path.exists(p + ".tmp")
    x, step, t, _, _ = io.read_checkpoint_host(p, 1000, 0.02, 0)
    assert step == 123 and t == 2.46 and np.array_equal(x, psi)
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if os.path.exists(h5dump):
        out = subprocess.run([h5dump, "-H", p], capture_output=True, text=True).stdout
        for name in ("version", "step", "time", "time_step", "dofs", "propagator", "psi", "complex"):
            assert f'"{name}"' in out, name
        assert "( 1000, 2 )" in out


def test_checkpoint_guess_history(tmp_path):
    p = str(tmp_path / "ck.h5")
    psi = _rand(64, 2)
    guess = np.stack([np.stack([_rand(64, 10 + 3 * s + j) for j in range(3)]) for s in range(4)])
    counts = np.array([3, 2, 1, 0], np.int32)
    io.write_checkpoint_host(p, psi, 7, 0.7, 0.1, 1, counts, guess)
    x, step, t, cnt, g = io.read_checkpoint_host(p, 64, 0.1, 1, 4, 3)
    assert np.array_equal(cnt, counts) and np.array_equal(x, psi)
    for s in range(4):
        for j in range(counts[s]):
            assert np.array_equal(g[s, j], guess[s, j])


@pytest.mark.parametrize("kw,msg", [(dict(dofs=999), "DOFs"), (dict(dt=0.01), "TimeStep"),
                                    (dict(propagator=1), "Propagator")])
def test_checkpoint_mismatch_is_an_error(tmp_path, kw, msg):
    p = str(tmp_path / "ck.h5")
    io.write_checkpoint_host(p, _rand(1000, 3), 1, 0.02, 0.02, 0)
    args = dict(dofs=1000, dt=0.02, propagator=0)
    args.update(kw)
    with pytest.raises(io.IOError_, match=msg):
        io.read_checkpoint_host(p, args["dofs"], args["dt"], args["propagator"])
    with pytest.raises(io.IOError_, match="no checkpoint file"):
        io.read_checkpoint_host(str(tmp_path / "missing.h5"), 1000, 0.02, 0)


def test_trace_append_and_reopen(tmp_path):
    p = str(tmp_path / "obs.h5")
    rng = np.random.default_rng(5)
    dip, en, ac = rng.standard_normal((3, 10)), rng.standard_normal((3, 7)), rng.standard_normal((3, 3))
    with io.TraceWriter(p, 3) as w:
        w.append(dip[:2], en[:2], ac[:2])
    with io.TraceWriter(p, 3, append=True) as w:   # a restarted run extends the datasets
        w.append(dip[2:], en[2:], ac[2:])
    assert np.array_equal(io.read_trace(p, "dipoles"), dip)
    assert np.array_equal(io.read_trace(p, "energies"), en)
    assert np.array_equal(io.read_trace(p, "autocorrelation"), ac)
    with pytest.raises(io.IOError_):
        with io.TraceWriter(p, 2, append=True) as w:   # 1D/2D/3D dipole widths differ: 4, 7, 10 columns
            w.append(np.zeros((1, 7)), np.zeros((1, 7)), np.zeros((1, 3)))

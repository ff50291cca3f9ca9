"""Plain C programs calling the library through its C ABI, as the
reference's C callers do (tests/c_callers/*.c, built by
__graft_entry__.build()): their results equal the same runs driven through
the Python mirror, bit for bit."""
import os
import subprocess

import numpy as np
import pytest

import gpu_bpf
import smc2_rbpf

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c_callers", "_build")


def say(*a):
    print("[case]", *a, flush=True)


def _bin(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} is missing: run __graft_entry__.build()")
    return path


def test_rbpf_learner_from_c(orc, tmp_path):
    """smc2/test_smc2_rbpf.cu's learning flow from C."""
    T, nt, ni, seed, lag = 300, 64, 64, 7, 50
    y, _, _ = orc.rb_dgp(T, 0.95, 0.10, -1.0, 0.15, seed=5)
    yp = tmp_path / "y.f32"
    y.astype(np.float32).tofile(yp)
    say("c caller rbpf", T, nt, ni)
    r = subprocess.run([_bin("rbpf_learn"), str(yp), str(T), str(nt), str(ni), str(seed), str(lag)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line0, line1 = r.stdout.strip().splitlines()
    v = np.float32([float(t) for t in line0.split()])
    cnt = [int(t) for t in line1.split()]
    f = smc2_rbpf.Smc2Rbpf(nt, ni, seed=seed, noise_capacity=T + 128, fixed_lag=lag)
    f.run(y)
    m, s, c = f.theta_mean(), f.theta_std(), f.counters()
    f.close()
    assert np.array_equal(v[0::2], m) and np.array_equal(v[1::2], s)
    assert cnt == [c["n_resamples"], c["n_rejuv_accepts"], c["n_rejuv_total"], c["t_current"],
                   c["t_checkpoint"]]


def test_plain_filter_from_c(orc, tmp_path):
    """Per-tick gpu_bpf_step from C equals the Python step loop."""
    T, n, seed = 200, 100_000, 42
    _, y = orc.dgp(T, 0.97, 0.15, -1.0, seed=42)
    y = y.astype(np.float32)
    yp, op = tmp_path / "y.f32", tmp_path / "out.f32"
    y.tofile(yp)
    say("c caller bpf", T, n)
    r = subprocess.run([_bin("bpf_step"), str(yp), str(T), str(n), str(seed), str(op)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(op, np.float32).reshape(T, 2)
    f = gpu_bpf.GpuBpf(n, 0.97, 0.15, -1.0, seed=seed)
    ref = np.array([f.step(float(v)) for v in y], np.float32)
    cnt = f.resample_count()
    f.close()
    assert np.array_equal(out, ref)
    assert int(r.stdout.strip()) == cnt > 0

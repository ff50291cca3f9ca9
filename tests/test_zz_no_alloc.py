"""The allocation-free entry point gpu_bpf_run_device_noalloc against the
entry point it is a variant of (gpu_bpf_create + gpu_bpf_set_ess_threshold +
gpu_bpf_run_device), through the C ABI: final particle states and every
tick's statistics bit for bit, on consecutive calls that share the
library's scratch array (the second with fewer particles and ticks, so stale
contents of the larger first call must not leak into it), at the maximum
particle count, and without a stats output.  Sizes beyond the maximum are
refused before anything runs."""
import numpy as np
import pytest

import gpu_bpf
import sv_data

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _series(T, mu, nu, seed):
    return sv_data.sv_series(T, 0.97, 0.15, mu, nu_obs=nu, seed=seed)[1].astype(np.float32)


def _noalloc(n, mu, nu, seed, thr, y_dev, T, with_stats=True):
    L = gpu_bpf.load()
    x = gpu_bpf.DeviceArray.zeros(n)
    st = gpu_bpf.DeviceArray.zeros(16 * T) if with_stats else None   # BpfTickStats = 64 B
    rc = L.gpu_bpf_run_device_noalloc(n, 0.97, 0.15, mu, nu, seed, thr, y_dev.ptr, T, x.ptr,
                                      st.ptr if st else None, None)
    if rc < 0:
        raise gpu_bpf.GpuBpfError(gpu_bpf.last_error())
    xs = x.download()                       # synchronous copy after the default stream's work
    ss = st.download().view(gpu_bpf.STATS_DTYPE) if st else None
    x.close()
    if st:
        st.close()
    return xs, ss


def _reference(n, mu, nu, seed, thr, y_dev, T):
    f = gpu_bpf.GpuBpf(n, 0.97, 0.15, mu, nu_obs=nu, seed=seed, ess_threshold=thr)
    f.set_stats_capacity(T)
    f.run_device(y_dev.ptr, T)
    x = f.particles()
    st = f.stats_tail(T)
    f.close()
    return x, st


@pytest.mark.parametrize("case", [
    # (n, T, mu, nu_obs, seed, threshold): a large call, then a smaller one
    (1_000_003, 300, -1.0, 5.0, 7, 0.5),
    (70_000, 120, -0.6, 0.0, 11, 0.0),
    (gpu_bpf.NOALLOC_MAX_N, 24, -1.0, 0.0, 3, 0.5),
    (4097, 60, -1.2, 0.0, 5, 0.5),
])
def test_noalloc_equals_run_device(case):
    n, T, mu, nu, seed, thr = case
    say("noalloc", case)
    y = gpu_bpf.DeviceArray(_series(T, mu, nu, seed))
    xa, sa = _noalloc(n, mu, nu, seed, thr, y, T)
    xb, sb = _reference(n, mu, nu, seed, thr, y, T)
    y.close()
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_array_equal(sa, sb)
    if thr > 0:
        assert 0 < int(sa["resampled"].sum()) < T


def test_noalloc_twice_in_a_row_and_without_stats():
    say("noalloc twice")
    ya = gpu_bpf.DeviceArray(_series(200, -1.0, 0.0, 1))
    yb = gpu_bpf.DeviceArray(_series(50, -1.0, 0.0, 2))
    first = _noalloc(300_000, -1.0, 0.0, 1, 0.5, ya, 200)
    second = _noalloc(20_000, -1.0, 0.0, 2, 0.5, yb, 50, with_stats=False)
    ref_a = _reference(300_000, -1.0, 0.0, 1, 0.5, ya, 200)
    ref_b = _reference(20_000, -1.0, 0.0, 2, 0.5, yb, 50)
    ya.close()
    yb.close()
    np.testing.assert_array_equal(first[0], ref_a[0])
    np.testing.assert_array_equal(first[1], ref_a[1])
    np.testing.assert_array_equal(second[0], ref_b[0])
    assert second[1] is None


def test_noalloc_refuses_too_many_particles():
    L = gpu_bpf.load()
    rc = L.gpu_bpf_run_device_noalloc(gpu_bpf.NOALLOC_MAX_N + 1, 0.97, 0.15, -1.0, 0.0, 1, 0.5, None, 0,
                                      None, None, None)
    assert rc == -1 and "particles" in gpu_bpf.last_error()

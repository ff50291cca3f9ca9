"""GPU tests of the build-owned host entry points around the tick:
gpu_bpf_restart (tick 0 again, no allocation), the statistics ring (growing
keeps the records it holds, set_stats_capacity, the bound on what can be
read back), and the SMC^2 outer-resampling history.  The checks are exact:
a restarted filter repeats a fresh filter bit for bit, ring records equal
the ones read before the ring grew."""
import numpy as np
import pytest

import gpu_bpf
import gpu_bpf_batch
import sv_data

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _y(T, seed=42):
    return sv_data.sv_series(T, 0.97, 0.15, -1.0, seed=seed)[1].astype(np.float32)


@pytest.mark.parametrize("thr", [0.0, 0.5])
def test_restart_repeats_a_fresh_filter(thr):
    y = _y(120)
    say("restart", thr)
    a = gpu_bpf.GpuBpf(20000, 0.97, 0.15, -1.0, seed=9, ess_threshold=thr)
    a.run(y[:70])
    a.set_mu(-0.7)                       # restart keeps the parameters in force
    a.restart()
    assert a.timestep() == 0 and a.resample_count() == 0
    ra, sa = a.run(y)
    xa = a.particles()
    a.close()
    b = gpu_bpf.GpuBpf(20000, 0.97, 0.15, -0.7, seed=9, ess_threshold=thr)
    rb, sb = b.run(y)
    xb = b.particles()
    b.close()
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_array_equal(sa, sb)
    np.testing.assert_array_equal(ra, rb)


def test_ring_regrow_keeps_records():
    y = _y(2000)
    say("ring regrow")
    f = gpu_bpf.GpuBpf(4096, 0.97, 0.15, -1.0, seed=3, ess_threshold=0.5)
    _, s1 = f.run(y[:500])                # ring 1024 holds ticks 0..499
    _, s2 = f.run(y[500:2000])            # grows the ring to 2048
    allst = f.stats_tail(2000)
    np.testing.assert_array_equal(allst[:500], s1)
    np.testing.assert_array_equal(allst[500:], s2)
    with pytest.raises(gpu_bpf.GpuBpfError):
        f.stats_tail(2001)                # more than the filter has run
    f.close()
    # a ring that grows after wrapping keeps only its last old-capacity records
    f = gpu_bpf.GpuBpf(4096, 0.97, 0.15, -1.0, seed=3, ess_threshold=0.5)
    f.run(y[:1000])
    f.run(y[1000:1500])                   # no growth: ring 1024 now holds ticks 476..1499
    tail = f.stats_tail(1024)
    assert np.array_equal(tail["tick"], np.arange(476, 1500))
    f.set_stats_capacity(4096)
    np.testing.assert_array_equal(f.stats_tail(1024), tail)
    with pytest.raises(gpu_bpf.GpuBpfError):
        f.stats_tail(1025)                # tick 475 was overwritten before the growth
    f.run(y[1500:2000])
    st = f.stats_tail(1524)
    np.testing.assert_array_equal(st[:1024], tail)
    assert np.array_equal(st["tick"], np.arange(476, 2000))
    f.close()


def test_stats_capacity_covers_a_device_run():
    y = _y(3000)
    say("stats capacity")
    f = gpu_bpf.GpuBpf(8192, 0.97, 0.15, -1.0, seed=4, ess_threshold=0.5)
    f.set_stats_capacity(3000)
    yd = gpu_bpf.DeviceArray(y)
    f.run_device(yd.ptr, 3000)
    st = f.stats_tail(3000)
    f.close()
    yd.close()
    g = gpu_bpf.GpuBpf(8192, 0.97, 0.15, -1.0, seed=4, ess_threshold=0.5)
    _, ref = g.run(y)
    g.close()
    np.testing.assert_array_equal(st, ref)


def test_smc2_history_accounts_for_forward_particle_ticks():
    y = sv_data.sv_series(400, 0.97, 0.15, -1.0, seed=5)[1].astype(np.float32)
    say("smc2 history")
    s = gpu_bpf_batch.GpuSmc2(64, 128, 1024, seed=3, capacity=400, moves=2, outer_threshold=0.8)
    n0 = int(s.theta()[2].sum())
    s.run(y)
    ticks, sum_n = s.history()
    c = s.counters()
    s.close()
    assert c["outer_resamples"] == len(ticks) >= 1
    assert np.all(np.diff(ticks) > 0) and ticks[-1] <= 400
    seg = np.diff(np.concatenate([[0], ticks, [400]]))
    want = int(np.dot(np.concatenate([[n0], sum_n]), seg))
    assert c["pt_forward"] == want
    assert c["replay_ticks"] == 2 * int(ticks.sum())
    assert c["pt_replay"] == 2 * int(np.dot(sum_n, ticks))


@pytest.mark.parametrize("thr", [0.0, 0.5])
def test_captured_ticks_equal_tick_by_tick(thr):
    """gpu_bpf_run / gpu_bpf_run_device replay a captured 32-tick graph whose
    kernels take the tick's inputs from the control block; gpu_bpf_step
    launches tick by tick.  Both give the same particles and statistics, also
    across a set_mu (new arguments: a new capture), a second observation
    array, and a run that mixes the two."""
    y = _y(150, seed=6)
    say("captured ticks", thr)
    a = gpu_bpf.GpuBpf(30000, 0.97, 0.15, -1.0, seed=2, ess_threshold=thr)
    _, s1 = a.run(y[:70])                 # 2 graph replays + 6 launches
    a.set_mu(-0.6)
    yd = gpu_bpf.DeviceArray(y[70:150])
    a.set_stats_capacity(200)
    a.run_device(yd.ptr, 40)              # recaptured (mu changed), other array
    a.step(float(y[110]))                 # one tick launched on its own
    a.run_device(yd.ptr + 4 * 41, 39)     # below one graph: launched tick by tick
    st_a = a.stats_tail(150)
    xa = a.particles()
    a.close()
    yd.close()
    b = gpu_bpf.GpuBpf(30000, 0.97, 0.15, -1.0, seed=2, ess_threshold=thr)
    for t in range(150):
        if t == 70:
            b.set_mu(-0.6)
        b.step(float(y[t]))
    b.set_stats_capacity(200)
    st_b = b.stats_tail(150)
    xb = b.particles()
    b.close()
    np.testing.assert_array_equal(st_a[:70], s1)
    np.testing.assert_array_equal(st_a, st_b)
    np.testing.assert_array_equal(xa, xb)

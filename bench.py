#!/usr/bin/env python3
"""Benchmark of the SV bootstrap particle filter on one MI355X.

Workload (the large single-filter configuration): N = 51,200,000 FP32
particles, T = 5,000 ticks of matched-DGP data (phi = 0.97, sigma_eta = 0.15,
PCG32 DGP seed 42) whose mu switches -1.5 / -0.5 / +0.5 every 1,250 ticks (the
fourth block repeats -1.5); the filter is told each switch through
gpu_bpf_set_mu; resampling when ESS < 0.5 N; FP64 CDF scan and cumulative
log-likelihood.  A step is one tick.  The observations sit in HBM before the
timed region starts.

Prints one JSON line: particle-ticks/s, the HBM roofline of the dominant
kernel (propagate+weight), and the CPU oracle's throughput on a bounded
sample of the same workload.  Under "workloads" the same line carries the
other configurations (unless --no-extra): the always-resample N=65,536
filter with its log-likelihood error against the FP64 oracle, the
differentiable N=2^20 filter (forward + adjoint score), and the batched SMC^2
runs (CPMMH, and MALA with Student-t observations) next to the plain filter
looped over the same parameter sets.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bootstrap-particle-filter-in-ptx_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
METRIC = "particle-ticks/s + %HBM roofline; plain vs batched/adjoint at fixed loglik err"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def mu_schedule(T, block=1250, levels=(-1.5, -0.5, 0.5)):
    return np.array([levels[(t // block) % 3] for t in range(T)], np.float32)


def segments(mu_t):
    cuts = [0] + [t for t in range(1, len(mu_t)) if mu_t[t] != mu_t[t - 1]] + [len(mu_t)]
    return list(zip(cuts[:-1], cuts[1:]))


def nanmean_or_none(v):
    v = v[np.isfinite(v)]
    return float(v.mean()) if v.size else None


def enqueue_run(f, L, y_dev_ptr, mu_t, t0, t1):
    """Enqueue ticks [t0, t1) with the mu schedule; observations already in HBM."""
    for a, b in segments(mu_t[t0:t1]):
        f.set_mu(float(mu_t[t0 + a]))
        rc = L.gpu_bpf_run_device(f.h, y_dev_ptr + 4 * (t0 + a), b - a)
        if rc < 0:
            raise RuntimeError(f.L.gpu_bpf_last_error().decode())


def cpu_baseline(y, mu_t, n, ticks, threads):
    import oracle
    oracle.lib()
    t0 = time.perf_counter()
    oracle.run_filter(y[:0], n, rho=0.97, sigma_z=0.15, mu=float(mu_t[0]),
                      ess_threshold=0.5, tile=4096, threads=threads)
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.run_filter(y[:ticks], n, rho=0.97, sigma_z=0.15, mu=float(mu_t[0]),
                      ess_threshold=0.5, tile=4096, threads=threads, mu_t=mu_t[:ticks])
    dt = time.perf_counter() - t0 - t_init
    return n * ticks / dt, dt


def log_uniform_counts(rng, M, lo, hi):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), M)).round().astype(np.int32)


def prior_draws(rng, M):
    mu = rng.uniform(-2.0, 0.0, M).astype(np.float32)
    phi = rng.uniform(0.90, 0.995, M).astype(np.float32)
    sig = rng.uniform(0.05, 0.40, M).astype(np.float32)
    seeds = rng.integers(1, 1 << 30, M).astype(np.int32)
    return mu, phi, sig, seeds


def looped_plain(gpu_bpf, n, mu, phi, sig, seeds, y, ticks, nu=0.0, score=False):
    """The plain filter looped over parameter sets (the baseline the batch
    replaces): particle-ticks/s over `ticks` ticks of each filter."""
    tot = 0.0
    pt = 0
    for m in range(len(n)):
        f = gpu_bpf.GpuBpf(int(n[m]), float(phi[m]), float(sig[m]), float(mu[m]), nu_obs=nu,
                           seed=int(seeds[m]), ess_threshold=0.0)
        if score:
            t0 = time.perf_counter()
            f.score_forward(y[:ticks], 64)
            f.score_backward(True)
            tot += time.perf_counter() - t0
        else:
            f.run(y[:2])
            f.synchronize()
            t0 = time.perf_counter()
            f.run(y[2:2 + ticks])
            f.synchronize()
            tot += time.perf_counter() - t0
        f.close()
        pt += int(n[m]) * ticks
    return pt / tot


def extra_workloads(args, gpu_bpf, oracle):
    import gpu_bpf_batch
    out = {}
    # always-resample N = 65,536, T = 1,000 (the reference's CPU-runnable case)
    _, y0 = oracle.dgp(1000, 0.97, 0.15, -1.0, seed=42)
    y0 = y0.astype(np.float32)
    f = gpu_bpf.GpuBpf(65536, 0.97, 0.15, -1.0, seed=42, ess_threshold=0.0)
    tile, _ = f.layout()
    t0 = time.perf_counter()
    _, st = f.run(y0)
    wall = time.perf_counter() - t0
    f.close()
    ref = oracle.run_filter(y0, 65536, ess_threshold=0.0, tile=tile, mode=oracle.MODE_F64_COUPLED)[0]
    t0 = time.perf_counter()
    oracle.run_filter(y0[:200], 65536, ess_threshold=0.0, tile=tile, threads=args.cpu_threads)
    cpu = 65536 * 200 / (time.perf_counter() - t0)
    out["plain_n65536_t1000"] = {
        "particle_ticks_per_s": 65536 * 1000 / wall,
        "ms_per_tick": wall,
        "ll_fp64": float(st["ll_cum"][-1]),
        "ll_rel_err_vs_fp64_oracle": abs(st["ll_cum"][-1] - ref["ll_cum"][-1]) / abs(ref["ll_cum"][-1]),
        "cpu_oracle_particle_ticks_per_s": cpu,
        "cpu_threads": args.cpu_threads,
    }
    log("C0", out["plain_n65536_t1000"])
    # differentiable filter: N = 2^20, T = 2,000, checkpoint every 64 ticks
    hs = np.array(oracle.host_stream(42, 3), np.uint64)
    u = (hs >> np.uint64(9)).astype(np.float64) / 8388608.0
    th = (-2.0 + 2.0 * u[0], 0.90 + 0.095 * u[1], 0.05 + 0.35 * u[2])
    _, y2 = oracle.dgp(2000, th[1], th[2], th[0], seed=42)
    y2 = y2.astype(np.float32)
    n2 = 1 << 20
    f = gpu_bpf.GpuBpf(n2, th[1], th[2], th[0], seed=42)
    t0 = time.perf_counter()
    ll2 = f.score_forward(y2, 64)
    t1 = time.perf_counter()
    g2 = f.score_backward(True)
    t2 = time.perf_counter()
    f.close()
    out["adjoint_n1M_t2000"] = {
        "theta": [float(v) for v in th],
        "loglik": ll2,
        "score": [float(v) for v in g2],
        "forward_s": t1 - t0,
        "backward_s": t2 - t1,
        "particle_ticks_per_s": n2 * 2000 / (t2 - t0),
        "checkpoint_every": 64,
    }
    log("C2", out["adjoint_n1M_t2000"])
    # batched SMC^2: CPMMH, M = 512, N log-uniform in [256, 8192], T = 2,000
    T3 = 2000
    mu3 = mu_schedule(T3)
    _, y3 = oracle.dgp_regimes(mu3.astype(np.float64), rho=0.97, sigma_z=0.15, seed=42)
    y3 = y3.astype(np.float32)
    rng = np.random.default_rng(42)
    M = 512
    n3 = log_uniform_counts(rng, M, 256, 8192)
    mu, phi, sig, seeds = prior_draws(rng, M)
    b = gpu_bpf_batch.GpuBpfBatch(n3, mu, phi, sig, seeds)
    blk = b.autotune(y3, 64)
    b.close()
    b = gpu_bpf_batch.GpuBpfBatch(n3, mu, phi, sig, seeds, block=blk)
    b.run(y3[:32], stats=False)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(y3[32:T3], stats=False)
    b.synchronize()
    batch_pt = float(n3.sum()) * (T3 - 32) / (time.perf_counter() - t0)
    b.close()
    loop_pt = looped_plain(gpu_bpf, n3[:64], mu, phi, sig, seeds, y3, 40)
    t0 = time.perf_counter()
    s = gpu_bpf_batch.GpuSmc2(M, 256, 8192, seed=42, capacity=T3, block=blk)
    s.run(y3)
    wall = time.perf_counter() - t0
    c = s.counters()
    tm, tsd = s.theta_mean_std()
    s.close()
    out["smc2_cpmmh_m512_t2000"] = {
        "inner_batch_particle_ticks_per_s": batch_pt,
        "looped_plain_particle_ticks_per_s": loop_pt,
        "batch_over_looped": batch_pt / loop_pt,
        "block": blk,
        "sum_n": int(n3.sum()),
        "smc2_wall_s": wall,
        "smc2_useful_particle_ticks_per_s": c["pt_forward"] / wall,
        "smc2_total_particle_ticks_per_s": (c["pt_forward"] + c["pt_replay"]) / wall,
        "counters": c,
        "theta_mean": [float(v) for v in tm],
        "theta_sd": [float(v) for v in tsd],
    }
    log("C3", out["smc2_cpmmh_m512_t2000"])
    # batched SMC^2: MALA from per-filter scores, M = 1024, N log-uniform in
    # [512, 16384], Student-t (nu = 5) observations, T = 5,000
    T4 = 5000
    mu4 = mu_schedule(T4)
    _, y4 = oracle.dgp_regimes(mu4.astype(np.float64), rho=0.97, sigma_z=0.15, nu_obs=5.0, seed=42)
    y4 = y4.astype(np.float32)
    M = 1024
    n4 = log_uniform_counts(rng, M, 512, 16384)
    mu, phi, sig, seeds = prior_draws(rng, M)
    b = gpu_bpf_batch.GpuBpfBatch(n4, mu, phi, sig, seeds, nu_obs=5.0)
    blk4 = b.autotune(y4, 64)
    b.close()
    b = gpu_bpf_batch.GpuBpfBatch(n4, mu, phi, sig, seeds, nu_obs=5.0, block=blk4)
    b.enable_score()
    b.run(y4[:32], stats=False)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(y4[32:1056], stats=False)
    b.synchronize()
    batch4 = float(n4.sum()) * 1024 / (time.perf_counter() - t0)
    b.close()
    loop4 = looped_plain(gpu_bpf, n4[:16], mu, phi, sig, seeds, y4, 100, nu=5.0, score=True)
    t0 = time.perf_counter()
    s = gpu_bpf_batch.GpuSmc2(M, 512, 16384, nu_obs=5.0, seed=42, kind="mala", moves=3,
                              mala_step=0.5, block=blk4)
    s.run(y4)
    wall = time.perf_counter() - t0
    c = s.counters()
    tm, tsd = s.theta_mean_std()
    s.close()
    out["smc2_mala_m1024_t5000_student_t"] = {
        "inner_batch_with_score_particle_ticks_per_s": batch4,
        "looped_plain_score_particle_ticks_per_s": loop4,
        "batch_over_looped": batch4 / loop4,
        "block": blk4,
        "sum_n": int(n4.sum()),
        "smc2_wall_s": wall,
        "smc2_useful_particle_ticks_per_s": c["pt_forward"] / wall,
        "smc2_total_particle_ticks_per_s": (c["pt_forward"] + c["pt_replay"]) / wall,
        "counters": c,
        "theta_mean": [float(v) for v in tm],
        "theta_sd": [float(v) for v in tsd],
    }
    log("C4", out["smc2_mala_m1024_t5000_student_t"])
    out["smc2_rbpf_ntheta1024_t1000"] = rbpf_workload(oracle)
    log("RBPF", out["smc2_rbpf_ntheta1024_t1000"])
    return out


def rbpf_workload(oracle):
    """SMC^2 with the RBPF / OCSN inner filter, the reference's throughput and
    latency cases (smc2/test_smc2_rbpf.cu:412-470,600-680): N_theta = 1024,
    N_inner in {64, 128, 256, 512}, T = 1000 observations of the
    regime-switching SV model, fixed lag 200, seed 42."""
    import smc2_rbpf
    T = 1000
    y, _, _ = oracle.rb_dgp(T, 0.95, 0.10, -1.0, 0.15, seed=12345)
    rows = {}
    for ni in (64, 128, 256, 512):
        s = smc2_rbpf.Smc2Rbpf(1024, ni, seed=42, fixed_lag=200)
        s.run(y[:50])
        s.init_from_prior()
        lat = np.zeros(T)
        t0 = time.perf_counter()
        for t in range(T):
            a = time.perf_counter()
            s.update(y[t])
            lat[t] = time.perf_counter() - a
        wall = time.perf_counter() - t0
        c = s.counters()
        m, sd = s.theta_mean(), s.theta_std()
        s.close()
        med = float(np.median(lat))
        rows[f"n_inner_{ni}"] = {
            "ms_per_obs": 1e3 * wall / T,
            "forward_inner_particle_ticks_per_s": 1024 * ni * T / wall,
            "resamples": c["n_resamples"],
            "accept_rate": c["n_rejuv_accepts"] / max(c["n_rejuv_total"], 1),
            "update_us_p50": 1e6 * med,
            "update_us_p99": 1e6 * float(np.percentile(lat, 99)),
            "update_us_max": 1e6 * float(lat.max()),
            "updates_over_3x_median": int((lat > 3 * med).sum()),
            "theta_mean": [float(v) for v in m],
            "theta_sd": [float(v) for v in sd],
        }
        log("RBPF", ni, rows[f"n_inner_{ni}"])
    # the reference's learning case (smc2/test_smc2_rbpf.cu:250-335): 1024 x
    # 512, T = 1200, fixed lag 200; its published time is 300 ms per 1200
    # ticks on an RTX 5080 (smc2/multi_scale_smc2_architecture.md:34)
    truth = (0.95, 0.10, -1.0, 0.15)
    yl, _, _ = oracle.rb_dgp(1200, *truth, seed=12345)
    s = smc2_rbpf.Smc2Rbpf(1024, 512, seed=12345, noise_capacity=1328, fixed_lag=200)
    t0 = time.perf_counter()
    s.run(yl)
    wall = time.perf_counter() - t0
    m, sd, c = s.theta_mean(), s.theta_std(), s.counters()
    s.close()
    rows["learn_1024x512_t1200"] = {
        "wall_ms": 1e3 * wall,
        "reference_wall_ms_rtx5080": 300.0,
        "speedup_vs_reference": 300.0 / (1e3 * wall),
        "within_2sd": int(sum(abs(m[q] - truth[q]) <= 2 * sd[q] for q in range(4))),
        "accept_rate": c["n_rejuv_accepts"] / max(c["n_rejuv_total"], 1),
        "resamples": c["n_resamples"],
        "theta_mean": [float(v) for v in m],
        "theta_sd": [float(v) for v in sd],
    }
    log("RBPF learn", rows["learn_1024x512_t1200"])
    s = smc2_rbpf.Smc2Rbpf(1024, 512, seed=12345, noise_capacity=1328, fixed_lag=200, noise_bits=16)
    t0 = time.perf_counter()
    s.run(yl)
    rows["learn_1024x512_t1200_fp16_noise"] = {"wall_ms": 1e3 * (time.perf_counter() - t0),
                                                "theta_mean": [float(v) for v in s.theta_mean()]}
    s.close()
    log("RBPF learn fp16 noise", rows["learn_1024x512_t1200_fp16_noise"])
    # fixed-lag against full-history replays at T = 5000
    # (smc2/test_smc2_rbpf.cu:343-410)
    y5, _, _ = oracle.rb_dgp(5000, *truth, seed=12345)
    for lag in (0, 200):
        s = smc2_rbpf.Smc2Rbpf(1024, 512, seed=12345, noise_capacity=5128, fixed_lag=lag)
        t0 = time.perf_counter()
        s.run(y5)
        wall = time.perf_counter() - t0
        m, sd, c = s.theta_mean(), s.theta_std(), s.counters()
        s.close()
        rows[f"lag_{lag}_1024x512_t5000"] = {
            "wall_ms": 1e3 * wall,
            "resamples": c["n_resamples"],
            "accept_rate": c["n_rejuv_accepts"] / max(c["n_rejuv_total"], 1),
            "theta_mean": [float(v) for v in m],
            "theta_sd": [float(v) for v in sd],
        }
        log("RBPF lag", lag, rows[f"lag_{lag}_1024x512_t5000"])
    # CPU oracle (OpenMP over theta-particles) on a bounded sample
    cfg = oracle.rb_cfg(64, 512, seed=42, lag=200, T=T)
    t0 = time.perf_counter()
    oracle.run_rb_smc2(cfg, y[:40])
    rows["cpu_oracle_forward_inner_particle_ticks_per_s"] = 64 * 512 * 40 / (time.perf_counter() - t0)
    rows["cpu_sample"] = "oracle, N_theta=64, N_inner=512, first 40 observations"
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000, help="timed ticks")
    ap.add_argument("--warmup", type=int, default=100, help="untimed ticks")
    ap.add_argument("--particles", type=int, default=51_200_000)
    ap.add_argument("--ticks", type=int, default=5000, help="length of the observation series")
    ap.add_argument("--items", type=int, default=0, help="particles per lane (0 = auto)")
    ap.add_argument("--cpu-ticks", type=int, default=100)
    ap.add_argument("--timing-stride", type=int, default=10,
                    help="HIP-event timing of every n-th tick of the timed region")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the other configurations")
    ap.add_argument("--rbpf-only", action="store_true",
                    help="run only the RBPF SMC^2 workload and print its JSON line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 5 device-to-device copies of BYTES (the reference kernel "
                         "for scaling HBM traffic counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 5}))
        return
    if args.rbpf_only:
        import oracle
        print(json.dumps({"smc2_rbpf_ntheta1024_t1000": rbpf_workload(oracle)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(this build runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import torch
    import gpu_bpf
    import oracle

    n = args.particles
    T = max(args.ticks, args.steps)
    mu_t = mu_schedule(T)
    log("generating observations", T)
    _, y64 = oracle.dgp_regimes(mu_t.astype(np.float64), rho=0.97, sigma_z=0.15, seed=42)
    y = y64.astype(np.float32)
    dev = torch.device("cuda:0")
    y_dev = torch.from_numpy(y).to(dev)
    torch.cuda.synchronize()
    L = gpu_bpf.load()

    # warm-up on a separate filter instance
    if args.warmup > 0:
        log("warmup", args.warmup)
        w = gpu_bpf.GpuBpf(n, 0.97, 0.15, float(mu_t[0]), seed=42, ess_threshold=0.5,
                           items_per_thread=args.items or None)
        enqueue_run(w, L, y_dev.data_ptr(), mu_t, 0, args.warmup)
        w.synchronize()
        w.close()

    f = gpu_bpf.GpuBpf(n, 0.97, 0.15, float(mu_t[0]), seed=42, ess_threshold=0.5,
                       items_per_thread=args.items or None)
    tile, n_tiles = f.layout()
    # HIP events around the three kernels of every timing_stride-th tick
    # (each event adds ~2 us to the stream, so not every tick is timed)
    f.enable_kernel_timing(True, stride=args.timing_stride)
    K = args.steps
    log("timed", K, "ticks, N =", n, "tile", tile)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    enqueue_run(f, L, y_dev.data_ptr(), mu_t, 0, K)
    f.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tick_ms = f.kernel_tick_times(K)
    ms, cnt = f.kernel_times()
    st = np.zeros(K, gpu_bpf.STATS_DTYPE)
    ring = min(K, 1024)
    rc = L.gpu_bpf_copy_stats(f.h, st[K - ring:].ctypes.data, ring)
    rs_mask = st["resampled"][K - ring:] == 1
    tm = tick_ms[K - ring:]
    rcount = f.resample_count()
    f.close()
    if rc < 0:
        raise RuntimeError(gpu_bpf.last_error())
    log(f"wall {wall:.3f}s  kernel ms {ms.tolist()}  launches {cnt.tolist()}  resamples {rcount}")

    pt = float(n) * K
    value = pt / wall
    # algorithmic bytes of the propagate kernel: read x, write x', write l,
    # plus read l on ticks that carry weights (every tick after a tick that
    # did not resample); tick 0 reads x and writes l only
    n_prop = K - 1
    last_rs = int(st["resampled"][-1])
    n_carry = (K - 1) - (rcount - last_rs)     # ticks 1..K-1 whose predecessor kept weights
    bytes_prop = n * (8.0 + 12.0 * n_prop + 4.0 * n_carry)
    # the timed launches are propagating ticks (tick 0 is not sampled): their
    # mean algorithmic bytes are those of ticks 1..K-1
    avg_prop_s = ms[0] / 1e3 / max(cnt[0], 1)
    bytes_per_launch = n * (12.0 * n_prop + 4.0 * n_carry) / n_prop
    achieved = bytes_per_launch / avg_prop_s
    # whole-tick algorithmic traffic incl. resampling (read x', l; write x)
    bytes_total = bytes_prop + 12.0 * n * rcount
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-ticks/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "large single SV bootstrap filter: N=51.2M FP32 particles, T=5000 ticks, "
                        "mu regime shifts -1.5/-0.5/+0.5 every 1250 ticks, ESS<0.5N resampling, "
                        "FP64 CDF and log-likelihood",
            "n_particles": n,
            "ticks": K,
            "ess_threshold": 0.5,
            "tile": tile,
            "resamples": rcount,
            "ll_cum_fp64": float(st["ll_cum"][-1]),
            "whole_tick_hbm_frac": bytes_total / wall / HBM_PEAK,
            "kernel_ms_per_launch": {
                "propagate": ms[0] / max(cnt[0], 1),
                "finalize": ms[1] / max(cnt[1], 1),
                "resample": ms[2] / max(cnt[2], 1),
                "resample_on_resample_ticks": nanmean_or_none(tm[rs_mask, 2]),
                "resample_on_other_ticks": nanmean_or_none(tm[~rs_mask, 2]),
                "timed_ticks": int(cnt[0]),
            },
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_propagate",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": None,
        },
    }
    # measured HBM bytes per launch of the same kernel (rocprofv3 FETCH_SIZE /
    # WRITE_SIZE, calibrated on a copy of known size), recorded for this workload
    tpath = os.path.join(ROOT, "profiles", "k_propagate_sv51M_traffic.json")
    if n == 51_200_000 and tile == 4096 and os.path.exists(tpath):
        tr = json.load(open(tpath))
        res["roofline"]["traffic"] = tr["traffic_bytes_per_launch_mean"]
        res["roofline"]["traffic_carry_tick"] = tr["traffic_bytes_per_launch_carry_tick"]
        res["roofline"]["algorithmic_bytes_per_launch"] = bytes_per_launch
        res["roofline"]["traffic_source"] = tr["source"]
    # a plain device copy moving the same bytes as one propagate launch (half
    # read, half written), timed the same way: the streaming rate this access
    # pattern can reach on this GPU, beside the 8 TB/s peak
    half = int(bytes_per_launch // 8)
    src = torch.ones(half, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(20):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    copy_bw = 8.0 * half * 20 / (e0.elapsed_time(e1) * 1e-3)
    del src, dst
    res["roofline"]["copy_same_bytes_GBps"] = copy_bw / 1e9
    res["roofline"]["frac_of_copy"] = achieved / copy_bw
    log("copy of the same bytes", copy_bw / 1e9, "GB/s")
    if not args.no_cpu:
        log("cpu baseline", args.cpu_ticks, "ticks x", n, "threads", args.cpu_threads)
        v, dt = cpu_baseline(y, mu_t, n, args.cpu_ticks, args.cpu_threads)
        res["cpu_baseline"] = {
            "value": v,
            "unit": "particle-ticks/s",
            "cores": args.cpu_threads,
            "kind": "port",
            "sample": f"oracle FP32 filter (OpenMP), N={n}, first {args.cpu_ticks} ticks of the "
                      f"same workload, {dt:.1f} s",
        }
    if not args.no_extra:
        res["workloads"] = extra_workloads(args, gpu_bpf, oracle)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

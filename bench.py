#!/usr/bin/env python3
"""Benchmark of the SV bootstrap particle filter on one MI355X.

Workload (the large single-filter configuration): N = 51,200,000 FP32
particles, T = 5,000 ticks of matched-DGP data (phi = 0.97, sigma_eta = 0.15,
PCG32 DGP seed 42) whose mu switches -1.5 / -0.5 / +0.5 every 1,250 ticks (the
fourth block repeats -1.5); the filter is told each switch through
gpu_bpf_set_mu; resampling when ESS < 0.5 N; FP64 CDF scan and cumulative
log-likelihood.

A step is one batch of --ticks-per-step observations (default 250), so the
default 20 timed steps are the workload's T = 5,000 ticks from tick 0.  The
warm-up steps run on the same filter instance, which is then restarted at
tick 0 (gpu_bpf_restart, no allocation) before the timed region.  The
observations sit in HBM before the timed region starts.

Prints one JSON line: particle-ticks/s, the HBM roofline of the dominant
kernel (propagate+weight, HIP events on every tenth tick of the timed region),
and the CPU oracle's throughput on a bounded sample of the same workload.
Under "workloads" the same line carries the other configurations (unless
--no-extra): the always-resample N=65,536 filter, the differentiable N=2^20
filter (forward + adjoint score), the batched SMC^2 runs (CPMMH, and MALA with
Student-t observations) next to the plain filter looped over the same
parameter sets, each with its log-likelihood error against the FP64 oracle on
a sample the oracle covers, and the RBPF SMC^2 learner.

The observation series come from sv_data (the reference's data generator);
the CPU oracle is used only by the CPU-baseline and error-check legs
(cpu_* functions below).
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

import sv_data  # noqa: E402

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
METRIC = "particle-ticks/s + %HBM roofline; plain vs batched/adjoint at fixed loglik err"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(y, mu_t, n, ticks, ticks_1, threads):
    """The oracle's FP32 filter (the reference's kernel sequence restated in
    C, OpenMP over particles) on the first ticks of the same workload, on
    `threads` threads and on one; returns (rate, seconds, rate 1, seconds 1)."""
    import oracle
    oracle.lib()
    out = []
    for thr, tk in ((threads, ticks), (1, ticks_1)):
        t0 = time.perf_counter()
        oracle.run_filter(y[:0], n, rho=0.97, sigma_z=0.15, mu=float(mu_t[0]),
                          ess_threshold=0.5, tile=4096, threads=thr)
        t_init = time.perf_counter() - t0
        t0 = time.perf_counter()
        oracle.run_filter(y[:tk], n, rho=0.97, sigma_z=0.15, mu=float(mu_t[0]),
                          ess_threshold=0.5, tile=4096, threads=thr, mu_t=mu_t[:tk])
        dt = time.perf_counter() - t0 - t_init
        out += [n * tk / dt, dt]
    return tuple(out)


def cpu_ll_error(y, ll_gpu, n, tile, threads, **kw):
    """Relative error of a GPU cumulative log-likelihood against the FP64
    oracle run on the FP32 restatement's genealogy (same streams)."""
    import oracle
    ref = oracle.run_filter(y, n, tile=tile, mode=oracle.MODE_F64_COUPLED, threads=threads, **kw)[0]
    return abs(ll_gpu - ref["ll_cum"][-1]) / abs(ref["ll_cum"][-1])


def host_uniforms(seed, k):
    """First k uniforms (r >> 9) / 2^23 of the reference's host PCG32 stream:
    state seed * 67890 + 12345, increment 1442695040888963407, XSH-RR of the
    old state (gpu_bpf/gpu_bpf_ptx_full.cu:241-251,353)."""
    mask = (1 << 64) - 1
    st = (seed * 67890 + 12345) & mask
    out = []
    for _ in range(k):
        old = st
        st = (old * 6364136223846793005 + 1442695040888963407) & mask
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        r = ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF
        out.append((r >> 9) / 8388608.0)
    return out


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
    replaces), one filter after another: particle-ticks/s over `ticks` ticks
    of each filter.  With score, each filter runs the adjoint pair."""
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
            # untimed: the first 34 ticks, whose run captures the filter's
            # 32-tick graph; the timed ticks replay it
            f.run(y[:34])
            f.synchronize()
            t0 = time.perf_counter()
            f.run(y[34:34 + ticks], stats=False)
            f.synchronize()
            tot += time.perf_counter() - t0
        f.close()
        pt += int(n[m]) * ticks
    return pt / tot


def looped_plain_streams(gpu_bpf, n, mu, phi, sig, seeds, y, ticks, nu=0.0):
    """The plain filters over parameter sets on one stream each, all enqueued
    before any is waited on (the reference's IMM loop, gpu_bpf/gpu_bpf.cu:
    1033-1038): each filter's ticks replay its captured tick graphs, and the
    filters' streams overlap on the GPU."""
    y_dev = gpu_bpf.DeviceArray(y[:2 + ticks])
    fs = []
    for m in range(len(n)):
        f = gpu_bpf.GpuBpf(int(n[m]), float(phi[m]), float(sig[m]), float(mu[m]), nu_obs=nu,
                           seed=int(seeds[m]), ess_threshold=0.0)
        f.run_device(y_dev.ptr, 2)
        fs.append(f)
    for f in fs:
        f.synchronize()
    t0 = time.perf_counter()
    for f in fs:
        f.run_device(y_dev.ptr + 8, ticks)
    for f in fs:
        f.synchronize()
    dt = time.perf_counter() - t0
    for f in fs:
        f.close()
    y_dev.close()
    return float(np.sum(n[:len(fs)])) * ticks / dt


def adjoint_workload(args, gpu_bpf, thr):
    """C2: the differentiable filter's forward pass and pathwise score."""
    # differentiable filter: N = 2^20, T = 2,000, checkpoint every 64 ticks;
    # theta from the host stream of seed 42 (uniform prior box)
    u = host_uniforms(42, 3)
    th = (-2.0 + 2.0 * u[0], 0.90 + 0.095 * u[1], 0.05 + 0.35 * u[2])
    _, y2 = sv_data.sv_series(2000, th[1], th[2], th[0], seed=42)
    y2 = y2.astype(np.float32)
    n2 = 1 << 20
    # one untimed run on the timed instance (it allocates the checkpoints and
    # replay buffers, which the timed run after the restart keeps)
    f = gpu_bpf.GpuBpf(n2, th[1], th[2], th[0], seed=42)
    t0 = time.perf_counter()
    f.score_forward(y2, 64)
    g_first = f.score_backward(True)
    first_s = time.perf_counter() - t0
    f.restart()
    t0 = time.perf_counter()
    ll2 = f.score_forward(y2, 64)
    t1 = time.perf_counter()
    g2 = f.score_backward(True)
    t2 = time.perf_counter()
    f.close()
    row = {
        "theta": [float(v) for v in th],
        "loglik": ll2,
        "score": [float(v) for v in g2],
        "forward_s": t1 - t0,
        "backward_s": t2 - t1,
        "particle_ticks_per_s": n2 * 2000 / (t2 - t0),
        "checkpoint_every": 64,
        "first_run_s": first_s,
        "first_run_particle_ticks_per_s": n2 * 2000 / first_s,
        "first_run_score_equal": bool(np.array_equal(g_first, g2)),
    }
    if not args.no_cpu:
        # the forward pass is the always-resample filter: its log-likelihood
        # over the first 200 ticks against the FP64 oracle
        f = gpu_bpf.GpuBpf(n2, th[1], th[2], th[0], seed=42, ess_threshold=0.0)
        tile2, _ = f.layout()
        _, s2 = f.run(y2[:200])
        f.close()
        row["ll_rel_err_vs_fp64_oracle"] = cpu_ll_error(y2[:200], s2["ll_cum"][-1], n2, tile2, thr,
                                                        rho=th[1], sigma_z=th[2], mu=th[0],
                                                        ess_threshold=0.0)
        row["ll_err_sample"] = "N=2^20, first 200 ticks"
    return row


def extra_workloads(args, gpu_bpf, torch):
    import gpu_bpf_batch
    out = {}
    thr = args.cpu_threads
    # always-resample N = 65,536, T = 1,000 (the reference's CPU-runnable case)
    _, y0 = sv_data.sv_series(1000, 0.97, 0.15, -1.0, seed=42)
    y0 = y0.astype(np.float32)
    f = gpu_bpf.GpuBpf(65536, 0.97, 0.15, -1.0, seed=42, ess_threshold=0.0)
    tile, _ = f.layout()
    t0 = time.perf_counter()
    _, st = f.run(y0)
    cold = time.perf_counter() - t0
    # steady state: the same run again on the same instance after
    # gpu_bpf_restart (its captured tick graph is reused)
    f.restart()
    t0 = time.perf_counter()
    _, st2 = f.run(y0)
    wall = time.perf_counter() - t0
    f.close()
    if float(st2["ll_cum"][-1]) != float(st["ll_cum"][-1]):
        raise RuntimeError("C0 rerun after restart differs from the first run")
    row = {
        "particle_ticks_per_s": 65536 * 1000 / wall,
        "ms_per_tick": wall,
        "first_run_ms_per_tick": cold,
        "ll_fp64": float(st["ll_cum"][-1]),
    }
    if not args.no_cpu:
        import oracle
        row["ll_rel_err_vs_fp64_oracle"] = cpu_ll_error(y0, st["ll_cum"][-1], 65536, tile, thr,
                                                        ess_threshold=0.0)
        t0 = time.perf_counter()
        oracle.run_filter(y0[:200], 65536, ess_threshold=0.0, tile=tile, threads=thr)
        row["cpu_oracle_particle_ticks_per_s"] = 65536 * 200 / (time.perf_counter() - t0)
        row["cpu_threads"] = thr
    out["plain_n65536_t1000"] = row
    log("C0", row)
    out["adjoint_n1M_t2000"] = adjoint_workload(args, gpu_bpf, thr)
    log("C2", row)
    # batched SMC^2: CPMMH, M = 512, N log-uniform in [256, 8192], T = 2,000
    T3 = 2000
    mu3 = sv_data.mu_schedule(T3)
    _, y3 = sv_data.sv_series(T3, 0.97, 0.15, mu3.astype(np.float64), seed=42)
    y3 = y3.astype(np.float32)
    rng = np.random.default_rng(42)
    M = 512
    n3 = log_uniform_counts(rng, M, 256, 8192)
    mu, phi, sig, seeds = prior_draws(rng, M)
    out["smc2_cpmmh_m512_t2000"] = batch_and_smc2(args, gpu_bpf, gpu_bpf_batch, torch, y3, n3, mu,
                                                  phi, sig, seeds, 256, 8192, nu=0.0, kind="cpmmh",
                                                  moves=5, ticks_batch=T3 - 32)
    log("C3", out["smc2_cpmmh_m512_t2000"])
    # batched SMC^2: MALA from per-filter scores, M = 1024, N log-uniform in
    # [512, 16384], Student-t (nu = 5) observations, T = 5,000
    T4 = 5000
    mu4 = sv_data.mu_schedule(T4)
    _, y4 = sv_data.sv_series(T4, 0.97, 0.15, mu4.astype(np.float64), nu_obs=5.0, seed=42)
    y4 = y4.astype(np.float32)
    M = 1024
    n4 = log_uniform_counts(rng, M, 512, 16384)
    mu, phi, sig, seeds = prior_draws(rng, M)
    out["smc2_mala_m1024_t5000_student_t"] = batch_and_smc2(args, gpu_bpf, gpu_bpf_batch, torch, y4,
                                                            n4, mu, phi, sig, seeds, 512, 16384,
                                                            nu=5.0, kind="mala", moves=3,
                                                            ticks_batch=1024)
    log("C4", out["smc2_mala_m1024_t5000_student_t"])
    out["smc2_rbpf_ntheta1024_t1000"] = rbpf_workload(args)
    log("RBPF", out["smc2_rbpf_ntheta1024_t1000"])
    return out


def batch_only(args):
    """The two batched inner-filter workloads alone, as extra_workloads builds
    them (same seeded particle counts, priors and observations): autotuned
    block, 32 untimed ticks, then args.batch_ticks timed ticks each.  The
    command for kernel traces and counters of k_batch_tick."""
    import gpu_bpf_batch
    rng = np.random.default_rng(42)
    rows = {}
    for name, M, lo, hi, T, nu, score in (
            ("smc2_cpmmh_m512_t2000", 512, 256, 8192, 2000, 0.0, False),
            ("smc2_mala_m1024_t5000_student_t", 1024, 512, 16384, 5000, 5.0, True)):
        mus = sv_data.mu_schedule(T)
        _, yy = sv_data.sv_series(T, 0.97, 0.15, mus.astype(np.float64), nu_obs=nu, seed=42)
        yy = yy.astype(np.float32)
        n = log_uniform_counts(rng, M, lo, hi)
        mu, phi, sig, seeds = prior_draws(rng, M)
        b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, nu_obs=nu, stats=False)
        if score:
            b.enable_score()         # the autotuner times the kernel that will run
        blk = b.autotune(yy, 64)
        tile = b.tile
        b.close()
        b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, nu_obs=nu, block=blk, tile=tile, stats=False)
        if score:
            b.enable_score()
        b.run(yy[:32], stats=False)
        b.synchronize()
        k = min(args.batch_ticks, T - 32)
        t0 = time.perf_counter()
        b.run(yy[32:32 + k], stats=False)
        b.synchronize()
        dt = (time.perf_counter() - t0) / k
        b.close()
        rows[name] = {"block": blk, "tile": tile, "sum_n": int(n.sum()), "ticks": k, "us_per_tick": dt * 1e6,
                      "particle_ticks_per_s": float(n.sum()) / dt, "score": score}
        log(name, rows[name])
    return rows


def batch_and_smc2(args, gpu_bpf, gpu_bpf_batch, torch, y, n, mu, phi, sig, seeds, nmin, nmax, nu,
                   kind, moves, ticks_batch):
    """Inner-batch throughput (autotuned), the looped plain filters, the
    SMC^2 run, and the inner filters' log-likelihood error on a sample."""
    score = kind == "mala"
    M = len(n)
    # the inner filters as SMC^2 runs them: no per-tick mean, variance or ESS
    b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, nu_obs=nu, stats=False)
    if score:
        b.enable_score()             # the autotuner times the kernel that will run
    blk = b.autotune(y, 64)
    tile = b.tile
    b.close()

    def batch_rate(stats):
        b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, nu_obs=nu, block=blk, tile=tile, stats=stats)
        if score:
            b.enable_score()
        b.run(y[:32], stats=False)
        b.synchronize()
        t0 = time.perf_counter()
        b.run(y[32:32 + ticks_batch], stats=False)
        b.synchronize()
        rate = float(n.sum()) * ticks_batch / (time.perf_counter() - t0)
        b.close()
        return rate

    batch_pt = batch_rate(False)
    batch_stats_pt = batch_rate(True)
    k = 16 if score else 64
    loop_pt = looped_plain(gpu_bpf, n[:k], mu, phi, sig, seeds, y, 100 if score else 64, nu=nu,
                           score=score)
    row = {
        ("inner_batch_with_score_particle_ticks_per_s" if score else
         "inner_batch_particle_ticks_per_s"): batch_pt,
        ("looped_plain_score_particle_ticks_per_s" if score else
         "looped_plain_particle_ticks_per_s"): loop_pt,
        "batch_over_looped": batch_pt / loop_pt,
        ("inner_batch_with_score_and_stats_particle_ticks_per_s" if score else
         "inner_batch_with_stats_particle_ticks_per_s"): batch_stats_pt,
    }
    if not score:
        strm = looped_plain_streams(gpu_bpf, n[:M], mu, phi, sig, seeds, y, 200, nu=nu)
        row["looped_plain_streams_particle_ticks_per_s"] = strm
        row["batch_over_looped_streams"] = batch_pt / strm
        row["looped_streams_sample"] = f"all {M} filters, one stream each, 200 ticks"
    row.update({"block": blk, "tile": tile, "sum_n": int(n.sum())})
    T = len(y)
    t0 = time.perf_counter()
    s = gpu_bpf_batch.GpuSmc2(M, nmin, nmax, nu_obs=nu, seed=42, kind=kind, moves=moves,
                              mala_step=0.5, capacity=T, block=blk, tile=tile)
    s.run(y)
    wall = time.perf_counter() - t0
    c = s.counters()
    tm, tsd = s.theta_mean_std()
    s.close()
    row.update({
        "smc2_wall_s": wall,
        "smc2_useful_particle_ticks_per_s": c["pt_forward"] / wall,
        "smc2_total_particle_ticks_per_s": (c["pt_forward"] + c["pt_replay"]) / wall,
        "counters": c,
        "theta_mean": [float(v) for v in tm],
        "theta_sd": [float(v) for v in tsd],
    })
    if not args.no_cpu:
        # 8 inner filters spread over the particle-count range, 500 ticks,
        # against the FP64 oracle on the same streams
        pick = np.argsort(n)[np.linspace(0, M - 1, 8).astype(int)]
        errs = []
        bs = gpu_bpf_batch.GpuBpfBatch(n[pick], mu[pick], phi[pick], sig[pick], seeds[pick], nu_obs=nu,
                                       block=blk, tile=tile)
        bs.run(y[:500], stats=False)
        ll = bs.loglik()
        bs.close()
        for q, m in enumerate(pick):
            errs.append(cpu_ll_error(y[:500], ll[q], int(n[m]), tile, args.cpu_threads,
                                     rho=float(phi[m]), sigma_z=float(sig[m]), mu=float(mu[m]),
                                     nu_obs=nu, ess_threshold=0.0, seed=int(seeds[m])))
        row["ll_rel_err_vs_fp64_oracle_max"] = float(max(errs))
        row["ll_err_sample"] = "8 inner filters across the N range, first 500 ticks"
    return row


def rbpf_workload(args):
    """SMC^2 with the RBPF / OCSN inner filter, the reference's throughput and
    latency cases (smc2/test_smc2_rbpf.cu:412-470,600-680): N_theta = 1024,
    N_inner in {64, 128, 256, 512}, T = 1000 observations of the
    regime-switching SV model, fixed lag 200, seed 42."""
    import smc2_rbpf
    T = 1000
    y, _, _ = sv_data.rbpf_series(T, 0.95, 0.10, -1.0, 0.15, seed=12345)
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
    yl, _, _ = sv_data.rbpf_series(1200, *truth, seed=12345)
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
    y5, _, _ = sv_data.rbpf_series(5000, *truth, seed=12345)
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
    if not args.no_cpu:
        # CPU oracle (OpenMP over theta-particles) on a bounded sample
        import oracle
        cfg = oracle.rb_cfg(64, 512, seed=42, lag=200, T=T)
        t0 = time.perf_counter()
        oracle.run_rb_smc2(cfg, y[:40])
        rows["cpu_oracle_forward_inner_particle_ticks_per_s"] = 64 * 512 * 40 / (time.perf_counter() - t0)
        rows["cpu_sample"] = "oracle, N_theta=64, N_inner=512, first 40 observations"
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps (same instance)")
    ap.add_argument("--ticks-per-step", type=int, default=250,
                    help="observations per step; 20 x 250 = the workload's T = 5000")
    ap.add_argument("--particles", type=int, default=51_200_000)
    ap.add_argument("--items", type=int, default=0, help="particles per lane (0 = auto)")
    ap.add_argument("--cpu-ticks", type=int, default=100)
    ap.add_argument("--cpu-ticks-1", type=int, default=8, help="ticks of the one-thread CPU run")
    ap.add_argument("--timing-stride", type=int, default=10,
                    help="HIP-event timing of every n-th tick of the timed region")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the other configurations")
    ap.add_argument("--rbpf-only", action="store_true",
                    help="run only the RBPF SMC^2 workload and print its JSON line")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the differentiable filter's forward pass and score (profiling)")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the two batched inner-filter workloads and print their JSON line")
    ap.add_argument("--batch-ticks", type=int, default=256, help="timed ticks of --batch-only")
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
    if args.adjoint_only:
        import gpu_bpf
        print(json.dumps({"adjoint_n1M_t2000": adjoint_workload(args, gpu_bpf, args.cpu_threads)}),
              flush=True)
        return
    if args.batch_only:
        print(json.dumps({"inner_batches": batch_only(args)}), flush=True)
        return
    if args.rbpf_only:
        print(json.dumps({"smc2_rbpf_ntheta1024_t1000": rbpf_workload(args)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(this build runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import torch
    import gpu_bpf

    n = args.particles
    tps = args.ticks_per_step
    K = args.steps
    T = K * tps                      # timed ticks, from tick 0 of the series
    W = args.warmup * tps            # warm-up ticks on the same instance
    n_series = max(T, W)
    mu_t = sv_data.mu_schedule(n_series)
    log("generating observations", n_series)
    _, y64 = sv_data.sv_series(n_series, 0.97, 0.15, mu_t.astype(np.float64), seed=42)
    y = y64.astype(np.float32)
    dev = torch.device("cuda:0")
    y_dev = torch.from_numpy(y).to(dev)
    torch.cuda.synchronize()
    L = gpu_bpf.load()

    f = gpu_bpf.GpuBpf(n, 0.97, 0.15, float(mu_t[0]), seed=42, ess_threshold=0.5,
                       items_per_thread=args.items or None)
    f.set_stats_capacity(T)          # every timed tick's statistics stay on the device
    tile, n_tiles = f.layout()
    if W > 0:
        log("warmup", W, "ticks on the timed instance")
        enqueue_run(f, L, y_dev.data_ptr(), mu_t, 0, W)
        f.synchronize()
        f.restart()                  # back to tick 0, no allocation
        f.set_mu(float(mu_t[0]))
    # HIP events around the three kernels of every timing_stride-th tick
    # (each event adds ~2 us to the stream, so not every tick is timed)
    f.enable_kernel_timing(True, stride=args.timing_stride)
    log("timed", K, "steps =", T, "ticks, N =", n, "tile", tile)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    enqueue_run(f, L, y_dev.data_ptr(), mu_t, 0, T)
    f.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tick_ms = f.kernel_tick_times(T)
    st = f.stats_tail(T)
    rcount = f.resample_count()
    f.close()
    if len(tick_ms) != T or not np.array_equal(st["tick"], np.arange(T)):
        raise RuntimeError("timed ticks were not all recorded")
    rs = st["resampled"] == 1
    log(f"wall {wall:.3f}s  resamples {rcount}")

    pt = float(n) * T
    value = pt / wall
    # Algorithmic bytes per tick (4 B per read or write of a per-particle FP32
    # array): propagate reads x and writes x' and l (12 B), plus reads l on a
    # tick whose predecessor kept its weights (16 B); tick 0 reads x and
    # writes l (8 B); a resampling tick adds read x', l and write x (12 B).
    carry = np.zeros(T, bool)
    carry[1:] = ~rs[:-1]
    prop_bytes = np.where(carry, 16.0, 12.0) * n
    prop_bytes[0] = 8.0 * n
    tick_bytes = prop_bytes + np.where(rs, 12.0 * n, 0.0)
    timed = np.isfinite(tick_ms[:, 0])
    timed[0] = False                 # tick 0 propagates nothing
    prop_us = tick_ms[timed, 0] * 1e3
    avg_prop_s = float(np.mean(tick_ms[timed, 0])) / 1e3
    bytes_per_launch = float(np.mean(prop_bytes[timed]))
    achieved = bytes_per_launch / avg_prop_s
    fin_ms = tick_ms[np.isfinite(tick_ms[:, 1]), 1]
    rs_t = timed & rs
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
            "workload": f"large single SV bootstrap filter: N={n} FP32 particles, {T} ticks from "
                        f"tick 0 ({K} steps of {tps} observations; warm-up {W} ticks on the same "
                        f"instance, then gpu_bpf_restart), mu -1.5/-0.5/+0.5 switching every 1250 "
                        f"ticks via gpu_bpf_set_mu, ESS<0.5N resampling ({rcount} resampling "
                        f"ticks), FP64 CDF and log-likelihood",
            "n_particles": n,
            "ticks": T,
            "ticks_per_step": tps,
            "warmup_ticks": W,
            "ess_threshold": 0.5,
            "tile": tile,
            "resamples": rcount,
            "ll_cum_fp64": float(st["ll_cum"][-1]),
            "whole_tick_hbm_frac": float(tick_bytes.sum()) / wall / HBM_PEAK,
            "algorithmic_bytes_per_tick_mean": float(tick_bytes.mean()),
            "kernel_us": {
                "propagate_mean": float(prop_us.mean()),
                "propagate_p5_p50_p95": [float(v) for v in np.percentile(prop_us, [5, 50, 95])],
                "propagate_min_max": [float(prop_us.min()), float(prop_us.max())],
                # (small filters finalise inside the propagate launch: no samples)
                "finalize_mean": float(fin_ms.mean()) * 1e3 if fin_ms.size else None,
                "resample_on_resample_ticks": (float(np.mean(tick_ms[rs_t, 2])) * 1e3
                                               if rs_t.any() else None),
                "resample_on_other_ticks": float(np.mean(tick_ms[timed & ~rs, 2])) * 1e3,
                "timed_ticks": int(timed.sum()),
                "timing_stride": args.timing_stride,
                "propagate_per_timed_tick": [round(float(v), 1) for v in prop_us],
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
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
    }
    # measured HBM bytes per launch of the same kernel (rocprofv3 FETCH_SIZE /
    # WRITE_SIZE, calibrated on a copy of known size), recorded for this workload
    tpath = os.path.join(ROOT, "profiles", "k_propagate_sv51M_traffic.json")
    if n == 51_200_000 and tile == 4096 and os.path.exists(tpath):
        tr = json.load(open(tpath))
        res["roofline"]["traffic"] = tr["traffic_bytes_per_launch_mean"]
        res["roofline"]["traffic_carry_tick"] = tr["traffic_bytes_per_launch_carry_tick"]
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
        v, dt, v1, dt1 = cpu_baseline(y, mu_t, n, args.cpu_ticks, args.cpu_ticks_1, args.cpu_threads)
        try:
            aff = len(os.sched_getaffinity(0))
        except AttributeError:
            aff = None
        res["cpu_baseline"] = {
            "value": v,
            "unit": "particle-ticks/s",
            "cores": args.cpu_threads,
            "kind": "port",
            "sample": f"oracle FP32 filter (the reference's kernel sequence in C, OpenMP over "
                      f"particles), N={n}, first {args.cpu_ticks} ticks of the same workload on "
                      f"{args.cpu_threads} threads ({dt:.1f} s) and first {args.cpu_ticks_1} ticks "
                      f"on 1 thread ({dt1:.1f} s)",
            "single_core_value": v1,
            "cpu_model": cpu_model(),
            "nproc": os.cpu_count(),
            "cpus_allowed": aff,
        }
    if not args.no_extra:
        res["workloads"] = extra_workloads(args, gpu_bpf, torch)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

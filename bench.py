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
sample of the same workload.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000, help="timed ticks")
    ap.add_argument("--warmup", type=int, default=100, help="untimed ticks")
    ap.add_argument("--particles", type=int, default=51_200_000)
    ap.add_argument("--ticks", type=int, default=5000, help="length of the observation series")
    ap.add_argument("--items", type=int, default=0, help="particles per lane (0 = auto)")
    ap.add_argument("--cpu-ticks", type=int, default=12)
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
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
    f.enable_kernel_timing(True)
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
    avg_prop_s = ms[0] / 1e3 / max(cnt[0], 1)
    bytes_per_launch = bytes_prop / K
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
                "resample_on_resample_ticks": float(tm[rs_mask, 2].mean()) if rs_mask.any() else None,
                "resample_on_other_ticks": float(tm[~rs_mask, 2].mean()) if (~rs_mask).any() else None,
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
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

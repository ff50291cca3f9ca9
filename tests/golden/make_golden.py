"""Writes the committed fixtures of tests/golden/ from the CPU oracle.

Run:  python tests/golden/make_golden.py

The reference repository holds no golden vectors (no CSVs, no asserted
values; its only known-answer material is the printed self-check of
tests/test_rng_diag.cu:76-96).  These fixtures are produced by the oracle and
are pinned independently in tests/test_oracle.py (pure-Python restatement of
the reference seeding formula, SciPy's ndtri, brute-force resampling).
"""
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def f32_hex(v):
    return struct.pack(">f", v).hex()


def main():
    # PCG32 per-particle streams, seed 42, 16 particles (the shape of
    # tests/test_rng_diag.cu:21-96)
    pcg = {"seed": 42, "particles": []}
    for i in range(16):
        s0 = O.pcg_state(42, i, 0)
        r0 = O.pcg_output(s0)
        u0 = O.uniform(r0)
        pcg["particles"].append({
            "i": i,
            "state_after_init": f"{s0:016x}",
            "output_draw0": f"{r0:08x}",
            "u_draw0": f32_hex(u0),
            "z_draw0": f32_hex(O.icdf_f32(u0)),
            "state_after_1_draw": f"{O.pcg_state(42, i, 1):016x}",
            "state_after_1000_draws": f"{O.pcg_state(42, i, 1000):016x}",
            "output_draw1000": f"{O.pcg_output(O.pcg_state(42, i, 1000)):08x}",
        })
    pcg["host_stream_seed42"] = [f"{v:08x}" for v in O.host_stream(42, 16)]
    pcg["host_stream_seed7"] = [f"{v:08x}" for v in O.host_stream(7, 4)]
    pcg["host_stream_seed_minus3"] = [f"{v:08x}" for v in O.host_stream(-3, 4)]
    with open(os.path.join(HERE, "pcg32_seed42.json"), "w") as f:
        json.dump(pcg, f, indent=1)

    # ICDF table: u -> z (FP32 bits) over the clamp range, both tails and the
    # central/tail boundary
    us = np.concatenate([
        np.linspace(1e-6, 0.99993598, 257, dtype=np.float32),
        np.float32([1e-6, 0.02425, 0.0242499, 0.5, 0.52425, 0.97575, 0.975751, 0.99993598]),
    ]).astype(np.float32)
    tab = [[f32_hex(float(u)), f32_hex(O.icdf_f32(float(u)))] for u in us]
    with open(os.path.join(HERE, "icdf_f32.json"), "w") as f:
        json.dump({"u_z_hex": tab}, f)

    # Matched-DGP series of the plain-filter workload: mu=-1, phi=0.97,
    # sigma=0.15, Gaussian noise, PCG32 DGP seed 42, T=1000
    h, y = O.dgp(1000, rho=0.97, sigma_z=0.15, mu=-1.0, seed=42)
    np.save(os.path.join(HERE, "dgp_sv_seed42_T1000.npy"), np.stack([h, y]))

    # Resampling known answers: small CDFs and offsets -> ancestors
    rng = np.random.default_rng(3)
    cases = []
    for n in (1, 2, 3, 5, 8, 17, 64):
        for _ in range(3):
            w = rng.exponential(size=n)
            w[rng.random(n) < 0.3] = 0.0
            if w.sum() == 0:
                w[0] = 1.0
            cdf = np.cumsum(w) / w.sum()
            u = float(np.float32(rng.random() / n))
            anc = O.resample_ancestors(cdf, u)
            cases.append({"cdf": cdf.tolist(), "u": u, "ancestors": anc.tolist()})
    # exact ties between targets and CDF values (u = 0 is a possible draw):
    # lower_bound and upper_bound differ only here
    for n, w in ((8, np.ones(8)), (16, np.ones(16)), (8, np.array([1, 0, 0, 1, 1, 0, 1, 0], float)),
                 (4, np.array([0, 1, 0, 1], float))):
        cdf = np.cumsum(w) / w.sum()
        for u in (0.0, float(np.float32(0.5 / n))):
            anc = O.resample_ancestors(cdf, u)
            cases.append({"cdf": cdf.tolist(), "u": u, "ancestors": anc.tolist()})
    with open(os.path.join(HERE, "resample_kat.json"), "w") as f:
        json.dump(cases, f)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()

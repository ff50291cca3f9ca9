"""CPU tests: pin the oracle against the reference's formulas, the committed
golden fixtures and what the mathematics fixes.  No GPU needed.

Each pinning check is chosen so that a plausible mistake fails it; the
test_planted_* cases plant such mistakes in the oracle and show they are
caught.
"""
import json
import math
import os
import struct

import numpy as np
import pytest
from scipy.special import ndtri

M64 = (1 << 64) - 1
PCG_MULT = 6364136223846793005
PCG_HOST_INC = 1442695040888963407


def f32_from_hex(h):
    return struct.unpack(">f", bytes.fromhex(h))[0]


def py_pcg_output(old):
    xsh = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
    rot = old >> 59
    return ((xsh >> rot) | (xsh << ((-rot) & 31))) & 0xFFFFFFFF


def py_uniform(r):
    u = np.float32(r >> 9) * np.float32(1.0 / 8388608.0)
    u = max(u, np.float32(f32_from_hex("358637bd")))
    return min(u, np.float32(f32_from_hex("3f7ffbce")))


# ---------------------------------------------------------------------------
# PCG32 streams
# ---------------------------------------------------------------------------

def test_pcg_seeding_matches_reference_selfcheck(golden_dir, orc):
    """tests/test_rng_diag.cu:76-96 restated in Python integers."""
    g = json.load(open(os.path.join(golden_dir, "pcg32_seed42.json")))
    for p in g["particles"]:
        i = p["i"]
        inc = ((2 * i + 1) | 1) & M64
        s = (42 * PCG_MULT + inc) & M64
        s = (s * PCG_MULT + inc) & M64
        assert f"{s:016x}" == p["state_after_init"]
        out = py_pcg_output(s)
        assert f"{out:08x}" == p["output_draw0"]
        u = py_uniform(out)
        assert struct.pack(">f", u).hex() == p["u_draw0"]
        s1 = (s * PCG_MULT + inc) & M64
        assert f"{s1:016x}" == p["state_after_1_draw"]
        sk = s
        for _ in range(1000):
            sk = (sk * PCG_MULT + inc) & M64
        assert f"{sk:016x}" == p["state_after_1000_draws"]
        assert f"{py_pcg_output(sk):08x}" == p["output_draw1000"]
        # the oracle's closed-form jump-ahead reproduces the fixture
        assert orc.pcg_state(42, i, 0) == s
        assert orc.pcg_state(42, i, 1000) == sk
        assert orc.pcg_output(s) == out


@pytest.mark.parametrize("seed", [0, 1, 42, -3, 2**31 - 1, -(2**31)])
def test_pcg_jump_ahead_equals_stepping(orc, seed):
    for i in (0, 1, 5, 65535, 51_199_999):
        for k in (0, 1, 2, 3, 64, 999, 5000):
            assert orc.pcg_state(seed, i, k) == orc.pcg_state_naive(seed, i, k)


def test_host_stream(golden_dir, orc):
    """gpu_bpf/gpu_bpf_ptx_full.cu:241-251,353 restated in Python integers."""
    g = json.load(open(os.path.join(golden_dir, "pcg32_seed42.json")))
    for seed, key in ((42, "host_stream_seed42"), (7, "host_stream_seed7"),
                      (-3, "host_stream_seed_minus3")):
        st = (seed * 67890 + 12345) & M64
        vals = []
        for _ in range(len(g[key])):
            vals.append(py_pcg_output(st))
            st = (st * PCG_MULT + PCG_HOST_INC) & M64
        assert [f"{v:08x}" for v in vals] == g[key]
        assert orc.host_stream(seed, len(vals)) == vals


def test_uniform_clamp(orc):
    assert orc.uniform(0) == np.float32(f32_from_hex("358637bd"))
    assert orc.uniform(0xFFFFFFFF) == np.float32(f32_from_hex("3f7ffbce"))
    for r in (1 << 9, 12345 << 9, 0x80000000, 0x7FFFFE00):
        assert orc.uniform(r) == py_uniform(r)


# ---------------------------------------------------------------------------
# Scalar math
# ---------------------------------------------------------------------------

def test_expf_accuracy(orc):
    xs = np.concatenate([np.linspace(-86, 88, 40001, dtype=np.float32),
                         np.float32([0.0, -0.0, 1e-8, -1e-8, 0.5, -0.5, 1.0])])
    got = np.array([orc.expf(float(v)) for v in xs], np.float64)
    ref = np.exp(xs.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 1.5
    assert orc.expf(0.0) == 1.0
    assert orc.expf(-86.5) == 0.0 and orc.expf(float("-inf")) == 0.0
    assert orc.expf(88.5) == float("inf")
    assert math.isnan(orc.expf(float("nan")))


def test_logf_accuracy(orc):
    xs = np.concatenate([np.exp(np.linspace(-100, 88, 40001)),
                         np.float64([1.0, 2.0, 0.5, 1e-6, 0.02425, 1e-40, 1e-44])]).astype(np.float32)
    got = np.array([orc.logf(float(v)) for v in xs], np.float64)
    ref = np.log(xs.astype(np.float64))
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 1.0
    assert orc.logf(1.0) == 0.0
    assert orc.logf(0.0) == float("-inf")
    assert math.isnan(orc.logf(-1.0)) and math.isnan(orc.logf(float("nan")))
    assert orc.logf(float("inf")) == float("inf")


def test_icdf_f64_is_acklam(orc):
    """Acklam's algorithm: relative error below 1.15e-9 against the true inverse."""
    us = np.concatenate([np.linspace(1e-6, 1 - 1e-6, 20001), [0.02425, 0.97575, 0.5]])
    got = np.array([orc.icdf_f64(u) for u in us])
    ref = ndtri(us)
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    assert np.max(rel[us != 0.5]) < 1.2e-9
    assert orc.icdf_f64(0.5) == 0.0


def test_icdf_f32_fixture_and_bounds(golden_dir, orc):
    tab = json.load(open(os.path.join(golden_dir, "icdf_f32.json")))["u_z_hex"]
    for uh, zh in tab:
        u = f32_from_hex(uh)
        z = orc.icdf_f32(u)
        assert struct.pack(">f", z).hex() == zh
        z64 = ndtri(u)
        q = abs(u - 0.5)
        # FP32 evaluation of the reference's rational: tight away from the
        # central/tail boundary, ~2e-4 just inside it (cancellation in the
        # denominator, present in the reference's own FP32 code)
        bound = 3e-4
        for edge, b in ((0.2, 2e-7), (0.3, 6e-7), (0.35, 2e-6), (0.40, 8e-6), (0.45, 6e-5)):
            if q < edge:
                bound = b
                break
        if q >= 0.47575:
            bound = 2e-6    # tails: sqrt(-2 log p) rational is well conditioned
        assert abs(z - z64) <= bound, (u, z, z64)
    assert orc.icdf_f32(0.5) == 0.0


def test_icdf_f32_symmetry_and_monotone(orc):
    us = np.linspace(1e-6, 0.5, 5001, dtype=np.float32)
    z = np.array([orc.icdf_f32(float(u)) for u in us])
    assert np.all(np.diff(z) >= -3e-4)      # monotone up to the boundary glitch
    # u and 1-u are both exact in FP32 for u >= 2^-24 multiples near 1/2 only;
    # use pairs that are exactly representable
    for u in (0.25, 0.125, 0.0625, 0.4375, 0.0234375):
        assert orc.icdf_f32(u) == -orc.icdf_f32(1.0 - u)


# ---------------------------------------------------------------------------
# DGP (tests/test_bpf_matched_dgp.cu:59-175) restated in pure Python
# ---------------------------------------------------------------------------

def py_dgp(n, rho, sigma_z, mu, seed):
    class R:
        pass
    g = R()

    def nxt():
        old = g.state
        g.state = (old * PCG_MULT + g.inc) & M64
        return py_pcg_output(old)

    g.state = 0
    g.inc = ((seed * 12345) | 1) & M64
    nxt()
    g.state = (g.state + seed * 67890) & M64
    nxt()
    nxt()

    def dbl():
        return (nxt() >> 11) * (1.0 / 2097152.0)

    def gauss():
        u1 = dbl()
        while u1 < 1e-10:
            u1 = dbl()
        u2 = dbl()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * 3.14159265358979323846 * u2)

    var = sigma_z * sigma_z / max(1.0 - rho * rho, 1e-6)
    h = mu + math.sqrt(var) * gauss()
    hs, ys = [], []
    for t in range(n):
        if t > 0:
            h = mu + rho * (h - mu) + sigma_z * gauss()
        h = min(max(h, -12.0), 4.0)
        hs.append(h)
        ys.append(math.exp(h / 2.0) * gauss())
    return np.array(hs), np.array(ys)


def test_dgp_fixture(golden_dir, orc):
    hy = np.load(os.path.join(golden_dir, "dgp_sv_seed42_T1000.npy"))
    h, y = orc.dgp(1000, 0.97, 0.15, -1.0, seed=42)
    np.testing.assert_array_equal(h, hy[0])
    np.testing.assert_array_equal(y, hy[1])
    ph, py = py_dgp(200, 0.97, 0.15, -1.0, 42)
    np.testing.assert_allclose(ph, hy[0][:200], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(py, hy[1][:200], rtol=1e-12, atol=1e-13)


# ---------------------------------------------------------------------------
# Resampling (gpu_bpf/gpu_bpf.cu:297-310, ptx:816-830)
# ---------------------------------------------------------------------------

def brute_ancestors(cdf, u):
    n = len(cdf)
    out = []
    for i in range(n):
        t = np.float32(u) + np.float32(i) / np.float32(n)
        if t >= np.float32(1.0):
            t = t - np.float32(1.0)
        j = next((j for j in range(n) if cdf[j] >= float(t)), n - 1)
        out.append(min(j, n - 1))
    return out


def test_resample_fixture_brute_force(golden_dir, orc):
    cases = json.load(open(os.path.join(golden_dir, "resample_kat.json")))
    for c in cases:
        exp = brute_ancestors(np.array(c["cdf"]), c["u"])
        assert exp == c["ancestors"]
        assert orc.resample_ancestors(c["cdf"], c["u"]).tolist() == exp


def test_resample_offspring_counts(orc):
    """Systematic resampling gives each particle floor(Nw) or ceil(Nw) copies."""
    rng = np.random.default_rng(11)
    for n in (7, 64, 1000, 4096):
        w = rng.exponential(size=n) ** 3
        w /= w.sum()
        cdf = np.cumsum(w)
        cdf[-1] = 1.0
        u = float(np.float32(rng.random() / n))
        anc = orc.resample_ancestors(cdf, u)
        assert np.all(np.diff(anc) >= 0) or n > 1000   # monotone unless a wrap occurs
        cnt = np.bincount(anc, minlength=n)
        nw = n * w
        assert np.all(cnt >= np.floor(nw) - 1) and np.all(cnt <= np.ceil(nw) + 1)
        assert cnt.sum() == n


def test_resample_wrap_quirk(orc):
    """Targets that round up to >= 1 wrap to the front (ptx:812-814)."""
    n = 51_200_000 // 4096
    cdf = np.linspace(1.0 / n, 1.0, n)
    u = float(np.float32(0.999999 / n))
    anc = orc.resample_ancestors(cdf, u)
    assert anc[0] == 0 and anc.max() == n - 1


# ---------------------------------------------------------------------------
# Whole-filter invariants
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def sv_data(orc):
    h, y = orc.dgp(300, 0.97, 0.15, -1.0, seed=42)
    return h, y


def test_always_resample_semantics(orc, sv_data):
    _, y = sv_data
    out, cnt, x, _ = orc.run_filter(y[:100], 4096, ess_threshold=0.0, want_x=True)
    assert cnt == 100 and np.all(out["resampled"] == 1)
    # without carry the reference log_lik equals the incremental estimate
    np.testing.assert_array_equal(out["ll_ref"], out["ll_incr"])
    assert np.all(out["ess"] >= 1.0) and np.all(out["ess"] <= 4096 * (1 + 1e-12))
    np.testing.assert_allclose(np.cumsum(out["ll_incr"]), out["ll_cum"], rtol=1e-12)


def test_ess_threshold_semantics(orc, sv_data):
    _, y = sv_data
    out, cnt, _, _ = orc.run_filter(y, 4096, ess_threshold=0.5)
    rs = out["resampled"] == 1
    assert cnt == rs.sum() and 0 < cnt < len(y)
    assert np.all(out["ess"][rs] / 4096 < 0.5)
    assert np.all(out["ess"][~rs] / 4096 >= 0.5)


def test_fp32_vs_fp64_oracle(orc, sv_data):
    """Arithmetic error of the FP32 pipeline: FP64 oracle on the same genealogy.
    Loglik error <= 1e-4 relative (the floating-point tolerance of the filter's
    log-likelihood); measured ~1e-7."""
    _, y = sv_data
    for thr in (0.0, 0.5):
        o32 = orc.run_filter(y, 16384, ess_threshold=thr, threads=8)[0]
        o64 = orc.run_filter(y, 16384, ess_threshold=thr, mode=orc.MODE_F64_COUPLED)[0]
        np.testing.assert_array_equal(o32["resampled"], o64["resampled"])
        rel = abs(o32["ll_cum"][-1] - o64["ll_cum"][-1]) / abs(o64["ll_cum"][-1])
        assert rel <= 1e-5, rel
        np.testing.assert_allclose(o32["ll_incr"], o64["ll_incr"], rtol=0, atol=2e-4)
        np.testing.assert_allclose(o32["mean"], o64["mean"], rtol=0, atol=2e-4)
        np.testing.assert_allclose(o32["var"], o64["var"], rtol=0, atol=2e-4)


def test_independent_fp64_filter_agrees_statistically(orc, sv_data):
    """The uncoupled FP64 filter differs only by Monte-Carlo divergence after
    ancestor flips: same exact-likelihood check as the FP32 filter."""
    _, y = sv_data
    g, _ = orc.grid_loglik(y, 0.97, 0.15, -1.0, n_grid=2048)
    v = np.array([orc.run_filter(y, 16384, seed=s, mode=orc.MODE_F64)[0]["ll_cum"][-1]
                  for s in (1, 2, 3)])
    assert abs(v.mean() - g) < 4 * 0.06 + 0.05


def _grid_check(orc, y, thr, n=16384, seeds=range(1, 7)):
    g, _ = orc.grid_loglik(y, 0.97, 0.15, -1.0, n_grid=2048)
    vals = np.array([orc.run_filter(y, n, ess_threshold=thr, seed=s, threads=8)[0]["ll_cum"][-1]
                     for s in seeds])
    return abs(vals.mean() - g), vals.std() / math.sqrt(len(vals))


def test_filter_matches_exact_likelihood(orc, sv_data):
    _, y = sv_data
    for thr in (0.0, 0.5):
        d, se = _grid_check(orc, y, thr)
        assert d < 4 * se + 0.05, (thr, d, se)


def test_threaded_oracle_matches_sequential(orc, sv_data):
    _, y = sv_data
    a, _, xa, _ = orc.run_filter(y[:50], 8192, ess_threshold=0.5, want_x=True, threads=1)
    b, _, xb, _ = orc.run_filter(y[:50], 8192, ess_threshold=0.5, want_x=True, threads=8)
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_allclose(a["ll_cum"], b["ll_cum"], rtol=1e-12)


# ---------------------------------------------------------------------------
# Planted mistakes: each must be caught by the checks above
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("fault", [1, 2])
def test_planted_model_faults_are_caught(orc, sv_data, fault):
    _, y = sv_data
    with orc.planted_fault(fault):
        d, se = _grid_check(orc, y, 0.0, seeds=range(1, 4))
    assert d > 4 * se + 0.05


def test_planted_carry_fault_is_caught(orc, sv_data):
    _, y = sv_data
    with orc.planted_fault(4):
        d, se = _grid_check(orc, y, 0.5, seeds=range(1, 4))
    assert d > 4 * se + 0.05


def test_planted_resample_fault_is_caught(golden_dir, orc):
    cases = json.load(open(os.path.join(golden_dir, "resample_kat.json")))
    bad = 0
    with orc.planted_fault(3):
        for c in cases:
            if orc.resample_ancestors(c["cdf"], c["u"]).tolist() != brute_ancestors(np.array(c["cdf"]), c["u"]):
                bad += 1
    assert bad > 0


@pytest.mark.parametrize("nu,tile", [(0.0, 256), (5.0, 4096)])
def test_tick_by_tick_stepping_repeats_the_whole_run(orc, nu, tile):
    """orc_bpf_tick from its own state, tick after tick (carried weights,
    host stream advanced per resample, a mu switch), repeats orc_bpf_run:
    particles bit for bit, statistics to 1e-12; its FP64 increments agree
    with the FP64-coupled run's to 1e-6 (same genealogy, FP32 vs FP64 state)."""
    import sv_data

    T, n = 80, 9000
    mu_t = np.where(np.arange(T) < 40, -1.0, -0.4).astype(np.float32)
    y = sv_data.sv_series(T, 0.97, 0.15, mu_t.astype(np.float64), nu_obs=nu, seed=3)[1].astype(np.float32)
    ref, _, xo, _ = orc.run_filter(y, n, nu_obs=nu, ess_threshold=0.5, tile=tile, mu_t=mu_t, seed=9,
                                   want_x=True)
    ref64 = orc.run_filter(y, n, nu_obs=nu, ess_threshold=0.5, tile=tile, mu_t=mu_t, seed=9,
                           mode=orc.MODE_F64_COUPLED, threads=4)[0]
    # x_0 from the oracle itself (zero ticks)
    _, _, x, _ = orc.run_filter(y[:0], n, ess_threshold=0.5, tile=tile, mu=float(mu_t[0]), seed=9, want_x=True)
    l = np.zeros(n, np.float32)
    h = int(orc.lib().orc_pcg_host_seed(9))
    carry = False
    for t in range(T):
        rec, rs, _, _, ll64, h = orc.filter_tick(x, l, t, y[t], carry, h, n, float(mu_t[t]), nu_obs=nu,
                                                 ess_threshold=0.5, seed=9, tile=tile, threads=4)
        assert rs == bool(ref["resampled"][t]), t
        for k in ("mean", "var", "ess", "ll_incr"):
            assert abs(rec[k] - ref[k][t]) <= 1e-12 * max(1.0, abs(ref[k][t])), (t, k)
        assert abs(ll64 - ref64["ll_incr"][t]) <= 1e-6 * max(1.0, abs(ref64["ll_incr"][t])), t
        carry = not rs
    np.testing.assert_array_equal(x, xo)


def test_threaded_score_oracle_matches_sequential(orc):
    _, y = orc.dgp(60, 0.97, 0.15, -1.0, seed=42)
    y = y.astype(np.float32)
    a = orc.score(y, 5000, tile=256)
    b = orc.score(y, 5000, tile=256, threads=4)
    assert abs(a[0] - b[0]) <= 1e-12 * abs(a[0])
    np.testing.assert_allclose(b[1], a[1], rtol=1e-10)

"""Oracle pins of the sigma_z scale bands (CPU).

Reference: gpu_bpf/gpu_bpf_ptx_full.cu:257-334 (band builder, adaptive regime
machine), ptx_assembly/bpf_kernels_full.ptx:297-324 (per-particle scale).  The
reference holds no vectors for them, so:

* one band of scale 1 leaves the filter unchanged, bit for bit;
* with several bands, the normal each particle drew is recovered from its
  transition (tick 1, no resampling) by dividing by sigma_z * k of the band
  the particle index falls in, and must equal the stream's normal -- this pins
  the rounded boundaries (fractions of n_particles) and the scales;
* the adaptive regimes equal a NumPy float32 evaluation of the reference's
  rule (EMA 0.3, reset in calm, thresholds) on the oracle's own surprise
  series |y_t| exp(-h_t / 2) of the previous tick.
"""
import numpy as np
import pytest


def _y(orc, T, seed=42):
    _, y = orc.dgp(T, 0.97, 0.15, -1.0, seed=seed)
    return y.astype(np.float32)


def test_unit_band_is_the_plain_filter(orc):
    y = _y(orc, 30)
    ref, _, x_ref, _ = orc.run_filter(y, 3000, ess_threshold=0.5, want_x=True)
    t, reg, x = orc.run_filter_bands(y, 3000, ([1.0], [1.0]), ess_threshold=0.5, want_x=True)
    assert np.array_equal(x, x_ref) and np.array_equal(t["ll_cum"], ref["ll_cum"])
    assert np.all(reg == 0)


def _band_scale(i, n, fracs, scales):
    if len(scales) <= 1:
        return np.full(i.shape, scales[0], np.float32)
    bnd, cum = [], 0
    for b in range(len(scales)):
        cnt = int(np.float32(fracs[b]) * np.float32(n) + np.float32(0.5)) if b < len(scales) - 1 else n - cum
        cum += cnt
        bnd.append(cum)
    k = np.full(i.shape, scales[-1], np.float32)
    for b in reversed(range(len(scales) - 1)):
        k = np.where(i < bnd[b], np.float32(scales[b]), k)
    return k


@pytest.mark.parametrize("fracs, scales", [
    ([0.9], [1.0, 4.0]),
    ([0.7, 0.15, 0.10], [1.0, 2.0, 5.0, 12.0]),
    ([0.5], [3.0]),                       # one band: every particle takes k[0]
])
def test_band_scales_recovered_from_the_transition(orc, fracs, scales):
    n, phi, sig, mu = 1000, 0.97, 0.15, -1.0
    y = _y(orc, 2)
    thr = 1e-30                           # no resampling
    _, _, x0 = orc.run_filter_bands(y[:1], n, (fracs, scales), ess_threshold=thr, want_x=True)
    _, _, x1 = orc.run_filter_bands(y, n, (fracs, scales), ess_threshold=thr, want_x=True)
    i = np.arange(n)
    k = _band_scale(i, n, fracs, scales).astype(np.float64)
    z = np.array([orc.normal(42, j, 1) for j in range(n)], np.float64)
    pred = mu + phi * (x0.astype(np.float64) - mu)
    z_rec = (x1.astype(np.float64) - pred) / (sig * k)
    np.testing.assert_allclose(z_rec, z, atol=2e-4)
    if len(scales) > 1:
        # the wrong band at a boundary would be off by the ratio of the scales
        wrong = (x1.astype(np.float64) - pred) / (sig * np.roll(k, 1))
        assert not np.allclose(wrong, z, atol=2e-4)


CALM = ([0.99], [1.0, 4.0])
ALERT = ([0.90, 0.05, 0.03], [1.0, 2.0, 4.0, 8.0])
PANIC = ([0.70, 0.15, 0.10], [1.0, 2.0, 5.0, 12.0])


def test_adaptive_regimes_follow_the_reference_rule(orc):
    T, n = 300, 4096
    y = _y(orc, T, seed=11)
    ta, tp = 0.5, 1.0                     # low thresholds: all three regimes occur
    t, reg, _ = orc.run_filter_bands(y, n, (CALM, ALERT, PANIC), adaptive=True, thr_alert=ta,
                                     thr_panic=tp, ess_threshold=0.5)
    f32 = np.float32
    ema, r, want = f32(0.0), 0, np.zeros(T, np.int32)
    surprise = f32(0.0)
    for k in range(T):
        if k > 0:
            ema = f32(f32(0.3) * surprise) + f32(f32(f32(1.0) - f32(0.3)) * ema)
            if surprise < f32(ta):
                r, ema = 0, surprise
            elif ema >= f32(tp):
                r = 2
            elif ema >= f32(ta):
                r = 1
            else:
                r = 0
        want[k] = r
        surprise = f32(np.abs(y[k]) * np.exp(f32(-t["h_mean"][k]) * f32(0.5), dtype=np.float32))
    assert set(np.unique(reg)) == {0, 1, 2}
    np.testing.assert_array_equal(reg, want)


def test_adaptive_with_equal_tables_is_static(orc):
    y = _y(orc, 60)
    t1, _, x1 = orc.run_filter_bands(y, 2048, (PANIC, PANIC, PANIC), adaptive=True, thr_alert=0.5,
                                     thr_panic=1.0, want_x=True)
    t2, _, x2 = orc.run_filter_bands(y, 2048, PANIC, want_x=True)
    assert np.array_equal(x1, x2) and np.array_equal(t1["ll_cum"], t2["ll_cum"])

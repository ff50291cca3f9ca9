"""The device's per-particle FP32 arithmetic against the oracle, bit for bit.

The inverse CDF is checked on every uniform the PCG32 mapping can produce
((r >> 9) * 2^-23 clamped, ptx_assembly/bpf_kernels_full.ptx:159-164: 2^23
values), which also pins the device's division of the central rational to
IEEE division.  exp and log are checked on dense samples of their working
ranges and on the special values.
"""
import numpy as np
import pytest

import gpu_bpf

pytestmark = pytest.mark.gpu


def all_uniforms():
    k = np.arange(1 << 23, dtype=np.float32)
    u = k * np.float32(1.0 / 8388608.0)
    lo = np.array([0x358637BD], np.uint32).view(np.float32)[0]
    hi = np.array([0x3F7FFBCE], np.uint32).view(np.float32)[0]
    return np.minimum(np.maximum(u, lo), hi)


def assert_bits_equal(a, b, what):
    ua, ub = a.view(np.uint32), b.view(np.uint32)
    both_nan = np.isnan(a) & np.isnan(b)
    bad = (ua != ub) & ~both_nan
    assert not bad.any(), (what, int(bad.sum()), a[bad][:5], b[bad][:5])


def test_icdf_every_uniform(orc):
    u = all_uniforms()
    print("[case] icdf over", u.size, "uniforms", flush=True)
    z = gpu_bpf.eval_device_math(gpu_bpf.MATH_ICDF, u)
    assert_bits_equal(z, orc.math_array(orc.MATH_ICDF, u), "icdf")


def _exp_inputs(nonpos):
    rng = np.random.default_rng(7)
    hi = 0.0 if nonpos else 90.0
    x = rng.uniform(-90.0, hi, 1 << 22).astype(np.float32)
    edges = np.float32([-np.inf, -100.0, -86.0, np.nextafter(np.float32(-86.0), np.float32(0)),
                        np.nextafter(np.float32(-86.0), np.float32(-100)), -1e-30, -0.0, 0.0,
                        np.nan])
    if not nonpos:
        edges = np.concatenate([edges, np.float32([1e-30, 88.0, np.nextafter(np.float32(88.0), 
                                np.float32(100)), 88.7, 100.0, 1e30, np.inf])])
    # every float in [-1, 0] at a stride, where the weights of a tile live
    fine = -np.arange(0, 0x3F800000, 97, dtype=np.uint32).view(np.float32)
    return np.concatenate([x, edges, fine]).astype(np.float32)


@pytest.mark.parametrize("fn", ["exp", "exp_nonpos"])
def test_exp(orc, fn):
    nonpos = fn == "exp_nonpos"
    x = _exp_inputs(nonpos)
    print("[case]", fn, x.size, flush=True)
    code = gpu_bpf.MATH_EXP_NONPOS if nonpos else gpu_bpf.MATH_EXP
    assert_bits_equal(gpu_bpf.eval_device_math(code, x), orc.math_array(orc.MATH_EXP, x), fn)


def test_log(orc):
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(0.0, 1.0, 1 << 21), np.exp(rng.uniform(-80, 80, 1 << 21)),
                        [0.0, -0.0, -1.0, np.inf, np.nan, 1e-45, 1e-40, 1.0]]).astype(np.float32)
    print("[case] log", x.size, flush=True)
    assert_bits_equal(gpu_bpf.eval_device_math(gpu_bpf.MATH_LOG, x),
                      orc.math_array(orc.MATH_LOG, x), "log")

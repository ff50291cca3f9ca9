"""CPU pins of the bit identities the device's normal draw relies on, over
every one of the 2^23 uniforms the generator can produce.

The device forms q = u - 1/2 as float(0x3F800000 | (r >> 9)) - 1.5 (one
funnel shift and one subtraction) instead of converting, scaling, clamping
and subtracting, evaluates every item as central, and re-forms the clamped
uniform from q only for the tail items (the reference's uniform of
ptx_assembly/bpf_kernels_full.ptx:159-169).  These checks are exact float32
arithmetic in numpy, the same IEEE operations the device performs."""
import numpy as np

CENTRAL = np.uint32(0x3EF39581).view(np.float32)       # 0.47575
LO = np.uint32(0x358637BD).view(np.float32)            # 1e-6
HI = np.uint32(0x3F7FFBCE).view(np.float32)            # 0.999936 (the reference's constant)


def all_draws():
    k = np.arange(1 << 23, dtype=np.uint32)              # r >> 9 for every r
    u_raw = k.astype(np.float32) * np.float32(1.0 / 8388608.0)
    u = np.minimum(np.maximum(u_raw, LO), HI)           # the reference's clamp
    q = (k | np.uint32(0x3F800000)).view(np.float32) - np.float32(1.5)
    return k, u_raw, u, q


def test_half_is_u_minus_half_for_every_draw():
    _, u_raw, _, q = all_draws()
    np.testing.assert_array_equal(q, u_raw - np.float32(0.5))
    np.testing.assert_array_equal(q + np.float32(0.5), u_raw)   # the tail pass re-forms u exactly


def test_clamp_moves_only_tail_draws():
    _, u_raw, u, q = all_draws()
    moved = u != u_raw
    assert moved.sum() == 9 + 536                        # u < 1e-6 and u > 0x3F7FFBCE
    central = np.abs(u - np.float32(0.5)) < CENTRAL
    assert not np.any(moved & central)
    # centrality is the same on q and on the clamped uniform, so every central
    # draw sees the q the reference's u - 1/2 gives
    np.testing.assert_array_equal(np.abs(q) < CENTRAL, central)
    np.testing.assert_array_equal(q[central], u[central] - np.float32(0.5))


def test_tail_uniform_is_the_clamped_uniform():
    _, _, u, q = all_draws()
    tail = ~(np.abs(q) < CENTRAL)
    assert 0.048 < tail.mean() < 0.049                   # 4.85 % of draws
    re = np.minimum(np.maximum(q[tail] + np.float32(0.5), LO), HI)
    np.testing.assert_array_equal(re, u[tail])

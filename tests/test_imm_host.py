"""Host logic of the IMM layer that needs no GPU: the model-grid builder
(gpu_bpf/gpu_bpf.cu:1100-1122, rho outermost, then sigma_z, then mu)."""
import numpy as np

import gpu_bpf


def test_build_grid_order_and_values():
    g = gpu_bpf.imm_build_grid([0.9, 0.97], [0.1, 0.2, 0.3], [-1.0, 0.5], nu_state=0.0, nu_obs=5.0)
    assert len(g) == 12
    want = [(r, s, m) for r in (0.9, 0.97) for s in (0.1, 0.2, 0.3) for m in (-1.0, 0.5)]
    got = [(a[0], a[1], a[2]) for a in g]
    np.testing.assert_array_equal(np.float32(got), np.float32(want))
    assert all(a[3] == 0.0 and a[4] == 5.0 for a in g)

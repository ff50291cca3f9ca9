"""Per-kernel event timing of the single filter (every stride-th tick)."""
import numpy as np
import pytest

import gpu_bpf

pytestmark = pytest.mark.gpu


def test_timing_stride_samples_every_nth_tick(orc):
    _, y = orc.dgp(50, 0.97, 0.15, -1.0, seed=3)
    f = gpu_bpf.GpuBpf(100_000, 0.97, 0.15, -1.0, seed=3)
    f.enable_kernel_timing(True, stride=5)
    f.run(y.astype(np.float32), stats=False)
    ms, cnt = f.kernel_times()
    per_tick = f.kernel_tick_times(50)
    f.close()
    assert cnt.tolist() == [10, 10, 10]
    assert np.all(ms > 0)
    timed = np.isfinite(per_tick).all(1)
    assert timed.tolist() == [(t + 1) % 5 == 0 for t in range(50)]
    assert np.isnan(per_tick[~timed]).all()
    np.testing.assert_allclose(per_tick[timed].sum(0), ms, rtol=1e-5)

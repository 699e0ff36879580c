"""The linear-grid bench workload and the bench's shift-group count (CPU only)."""
import numpy as np

from soap_gpu_amd import synth


def test_linear_series_shape():
    w = synth.linear_series(n_phase=25)
    assert w.n == 100_000 and w.grid_kind == 0 and w.meta["step"] == 0.005
    assert synth.is_linear(w.wave) and w.wave[0] == 5000.0
    np.testing.assert_allclose(np.diff(w.wave), 0.005, rtol=0, atol=1e-9)
    assert len(w.lines["center"]) == 1250 and len(w.mask_l) == 625  # 8000 lines per 3200 A; every other in the mask
    assert w.n_mu == 20 and w.star["grid"] == 1000 and w.star["w_ratio"] == 0.0
    assert len(w.phases) == 25 and len(w.vgrid) == 161
    # lines clear of the edges by more than the velocity window (the duplicated join point)
    assert w.lines["center"].min() >= w.wave[0] + 5.0 and w.lines["center"].max() <= w.wave[-1] - 5.0


def test_shift_group_count():
    import bench
    delta = np.array([1.0, 1.0, 2.0, 2.0, 2.0, 3.0, 1.0])
    region = np.array([0, 0, 0, 1, 1, 1, 1])
    order = np.arange(7)
    # runs of equal (region, delta) in list order: [0,1] [2] [3,4] [5] [6]
    assert bench._shift_groups((region, delta), order) == 5
    assert bench._shift_groups((region, delta), order[:0]) == 0
    # a planted mistake: ignoring the bin merges [2] with [3,4]
    assert bench._shift_groups((delta,), order) == 4

"""Synthetic workload generators (CPU): seeded, shaped and scaled as the
benchmark configurations state."""
import numpy as np

from opentm_amd import workloads as wl


def test_plain_density_seeded_and_clamped():
    a = wl.plain_density(16, 3)
    b = wl.plain_density(16, 3)
    np.testing.assert_array_equal(a, b)
    assert a.dtype == np.float32 and a.shape == (16, 16, 16)
    assert a.min() >= 1e-3 and a.max() <= 1.0
    assert abs(float(a.mean()) - 0.3) < 0.05


def test_grf_density_volume_and_periodicity():
    for vol in (0.1, 0.3, 0.5):
        r = wl.grf_density(24, 5, 6.0, vol)
        assert r.shape == (24, 24, 24) and r.dtype == np.float32
        assert abs(float(r.mean()) - vol) < 2e-3
        assert r.min() >= 1e-3 and r.max() <= 1.0
    # a periodic field has no jump across the cell faces beyond its interior variation
    r = wl.grf_density(32, 7, 8.0, 0.3).astype(np.float64)
    interior = np.abs(np.diff(r, axis=2)).max()
    seam = np.abs(r[:, :, 0] - r[:, :, -1]).max()
    assert seam <= interior + 1e-6
    np.testing.assert_array_equal(wl.grf_density(16, 9, 5.0, 0.2), wl.grf_density(16, 9, 5.0, 0.2))


def test_batch_cell_draws():
    import bench
    cells = bench.batch_cells(96, 0)
    assert len(cells) == 96
    assert {c[0] for c in cells} <= {32, 48, 64, 96, 128}
    assert all(0.15 <= c[1] <= 0.5 for c in cells)
    assert cells == bench.batch_cells(96, 0)

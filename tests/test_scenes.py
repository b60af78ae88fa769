"""The synthetic workload generators (CPU): the closed-form hinge renderer used for the
100-frame sequence agrees with sphere tracing of the same implicit scene."""
import numpy as np

from killingfusion_amd import scenes


def test_hinge_closed_form_matches_sphere_tracing():
    for t in (0.0, 0.6):
        a = scenes.hinge_frame_exact(t, noise=False)
        b = scenes.hinge_frame(t, noise=False)
        both = (a > 0) & (b > 0)
        assert both.sum() > 20000
        assert ((a > 0) != (b > 0)).sum() <= 10  # silhouette pixels only
        assert np.abs(a[both] - b[both]).max() <= 1e-6


def test_hinge_noise_is_the_seeded_stream():
    a = scenes.hinge_frame_exact(0.5, frame=3)
    b = scenes.hinge_frame_exact(0.5, frame=3, noise=False)
    assert np.array_equal(a > 0, b > 0)
    assert np.array_equal(a, scenes.hinge_frame_exact(0.5, frame=3))
    assert not np.array_equal(a, scenes.hinge_frame_exact(0.5, frame=4))
    # N(0, (1 mm z^2)^2): standardised residuals have mean ~0 and sd ~1
    z = b[b > 0].astype(np.float64)
    d = (a[b > 0].astype(np.float64) - z) / (0.001 * z * z)
    assert abs(d.mean()) < 0.05 and 0.95 < d.std() < 1.05


def test_batch_edges_in_range_and_reproducible():
    e = scenes.batch_edges(16)
    assert e == scenes.batch_edges(16) and len(e) == 16
    assert all(48 <= x <= 160 for x in e) and len(set(e)) > 8

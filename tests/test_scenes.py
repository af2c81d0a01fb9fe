"""Synthetic rendered scenes (training and evaluation data for the
proposal network): determinism and the ground-truth relations the scene
model fixes (make_dataset.py:219-227)."""
import numpy as np

from conftest import package


def _scenes():
    return package().scenes


def test_render_pair_is_deterministic():
    s = _scenes()
    a = s.render_pair(10007, 64, 48)
    b = s.render_pair(10007, 64, 48)
    for k in ("left", "right", "disp", "valid"):
        np.testing.assert_array_equal(a[k], b[k])
    c = s.render_pair(10008, 64, 48)
    assert not np.array_equal(a["left"], c["left"])


def test_peak_disparity_is_the_target():
    # the baseline is set from the nearest depth: max disparity == target
    s = _scenes()
    for seed in (10000, 10001, 90003):
        p = s.render_pair(seed, 96, 64, 22.0, 40.0)
        d = p["disp"][p["valid"]]
        assert 22.0 <= p["peak"] <= 40.0
        np.testing.assert_allclose(d.max(), p["peak"], rtol=1e-12)
        assert d.min() > 0.0
        assert p["left"].dtype == np.uint8 and p["left"].shape == (64, 96)


def test_views_differ_by_the_baseline_only():
    # a zero target peak means a zero baseline: both views render the same
    s = _scenes()
    p = s.render_pair(10005, 64, 48, 0.0, 0.0)
    np.testing.assert_array_equal(p["left"], p["right"])
    assert np.all(p["disp"][p["valid"]] == 0.0)


def test_value_noise_range_and_interpolation():
    s = _scenes()
    rng = np.random.default_rng(3)
    t = s.value_noise(rng, 128, 4)
    assert t.shape == (128, 128) and t.min() >= 0.0 and t.max() <= 1.0
    # Catmull-Rom weights sum to one and reproduce a constant lattice
    m = s._cubic_matrix(9, 100)
    np.testing.assert_allclose(m.sum(1), 1.0, atol=1e-12)
    np.testing.assert_allclose(m @ np.full(9, 0.3), 0.3, atol=1e-12)
    # and a linear ramp away from the clamped edges
    ramp = np.arange(9, dtype=float)
    pos = (np.arange(100) + 0.5) * (9 / 100.0) - 0.5
    inner = (pos >= 1.0) & (pos <= 7.0)
    np.testing.assert_allclose((m @ ramp)[inner], pos[inner], atol=1e-12)

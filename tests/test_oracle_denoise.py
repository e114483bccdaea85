"""The oracle's EATWT denoiser (denoiseSpectralRadianceEATWT,
cpu_denoise.ipp:10-187) pinned against what does not depend on it:

* the literal restatement (the reference's image(y, x) = data[y + x*W]
  indexing) equals the generalised one (first pass along j, then along i) bit
  for bit on square images, including pixels without a hit (infinite depth);
* on flat geometry the filter is the plain a-trous B3-spline transform, which
  scipy's correlate1d (clamped edges, dilated 1 4 6 4 1 / 16 kernel) computes
  independently;
* the normal and depth edge-stopping weights follow their closed forms on
  images that vary along one axis only;
* pixels without a hit keep their value, and hit pixels ignore their values;
* albedo demodulation: a constant albedo factors out; zero levels leave
  (v / max(a, eps)) * a."""
import numpy as np
import pytest
from scipy.ndimage import correlate1d

import pyoracle as O

NB = 8
K5 = np.array([1, 4, 6, 4, 1], np.float64) / 16.0


def _random_scene(H, W, seed, misses=True):
    rng = np.random.default_rng(seed)
    direct = rng.random((H, W, NB)).astype(np.float32)
    indirect = (0.2 * rng.random((H, W, NB))).astype(np.float32)
    albedo = (0.05 + rng.random((H, W, NB))).astype(np.float32)
    depth = (100.0 + 40.0 * rng.random((H, W))).astype(np.float32)
    n = rng.normal(size=(H, W, 3))
    n[..., 2] = np.abs(n[..., 2]) + 1.0
    normal = (n / np.linalg.norm(n, axis=2, keepdims=True)).astype(np.float32)
    if misses:
        m = rng.random((H, W)) < 0.05
        depth[m] = np.inf
        normal[m] = 0.0
        direct[m] = 0.0
        indirect[m] = 0.0
        albedo[m] = 0.0
    return direct, indirect, albedo, depth, normal


@pytest.mark.parametrize("levels", [(5, 5), (2, 4), (0, 3)])
def test_literal_equals_generalised_on_square_images(levels):
    args = _random_scene(48, 48, 1)
    a = O.denoise(*args, levels=levels, literal=True)
    b = O.denoise(*args, levels=levels, literal=False)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def _atrous(img, levels):
    out = img.astype(np.float64)
    for level in range(levels):
        s = 1 << level
        k = np.zeros(4 * s + 1)
        k[::s] = K5
        out = correlate1d(out, k, axis=0, mode="nearest")  # along j first
        out = correlate1d(out, k, axis=1, mode="nearest")
    return out


@pytest.mark.parametrize("levels", [1, 3, 5])
def test_flat_geometry_is_the_atrous_b3_transform(levels):
    H, W = 40, 56
    rng = np.random.default_rng(2)
    direct = rng.random((H, W, NB)).astype(np.float32)
    indirect = rng.random((H, W, NB)).astype(np.float32)
    ones = np.ones((H, W, NB), np.float32)
    depth = np.full((H, W), 50.0, np.float32)
    normal = np.zeros((H, W, 3), np.float32)
    normal[..., 2] = 1.0
    d, i = O.denoise(direct, indirect, ones, depth, normal, levels=(levels, levels))
    for got, src in ((d, direct), (i, indirect)):
        want = _atrous(src, levels)
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


def _one_axis_case(vary):
    """Rows identical; along i the normal (or depth) changes.  One level: the
    pass along j averages equal values, the pass along i follows the closed
    form out(i) = sum_k w_k v(i_k) / sum_k w_k."""
    H, W = 8, 24
    rng = np.random.default_rng(3)
    v_row = rng.random((W, NB)).astype(np.float32)
    direct = np.repeat(v_row[None], H, 0)
    ones = np.ones((H, W, NB), np.float32)
    ang = np.radians(np.linspace(0.0, 170.0, W)) if vary == "normal" else np.zeros(W)
    n_row = np.stack([np.sin(ang), np.zeros(W), np.cos(ang)], 1).astype(np.float32)
    d_row = (np.linspace(10.0, 40.0, W) * (1.0 + 0.5 * (np.arange(W) % 3 == 0))).astype(np.float32) \
        if vary == "depth" else np.full(W, 10.0, np.float32)
    normal = np.repeat(n_row[None], H, 0)
    depth = np.repeat(d_row[None], H, 0)
    d, _ = O.denoise(direct, direct, ones, depth, normal, levels=(1, 1))
    want = np.zeros((W, NB))
    for i in range(W):
        num, den = np.zeros(NB), 0.0
        for t in range(-2, 3):
            s = min(max(i + t, 0), W - 1)
            cd, sd = float(d_row[i]), float(d_row[s])
            ratio = max(cd, sd) / (min(cd, sd) + 1e-6)
            dw = 1.0 if ratio < 1.2 else np.exp(-(ratio - 1.0) / 0.2)
            sim = max(0.0, float(np.dot(n_row[i].astype(np.float64), n_row[s].astype(np.float64))))
            nw = max(0.3, np.exp(-((1.0 - sim) ** 2) / 0.5))
            w = K5[t + 2] * dw * nw
            num += w * v_row[s]
            den += w
        want[i] = num / den
    return d, want


@pytest.mark.parametrize("vary", ["normal", "depth"])
def test_edge_stopping_weights_closed_form(vary):
    d, want = _one_axis_case(vary)
    for row in d:
        assert np.abs(row - want).max() <= 2e-6 * np.abs(want).max()


def test_misses_keep_their_value_and_are_ignored_by_hits():
    direct, indirect, albedo, depth, normal = _random_scene(32, 32, 4)
    miss = np.isinf(depth)
    assert miss.any()
    direct[miss] = 3.0
    albedo[miss] = 1.0
    a, _ = O.denoise(direct, indirect, albedo, depth, normal)
    assert np.array_equal(a[miss], direct[miss])
    direct2 = direct.copy()
    direct2[miss] = 1e6
    b, _ = O.denoise(direct2, indirect, albedo, depth, normal)
    assert np.array_equal(a[~miss], b[~miss])


def test_albedo_demodulation():
    direct, indirect, albedo, depth, normal = _random_scene(24, 24, 5, misses=False)
    c = np.full_like(albedo, 0.37)
    ones = np.ones_like(albedo)
    d1, _ = O.denoise(direct, indirect, c, depth, normal, levels=(3, 3))
    d2, _ = O.denoise((direct / c).astype(np.float32), indirect, ones, depth, normal, levels=(3, 3))
    assert np.abs(d1 - 0.37 * d2).max() <= 1e-6 * np.abs(d1).max()
    d0, i0 = O.denoise(direct, indirect, albedo, depth, normal, levels=(0, 0))
    eps = np.float32(1e-6)
    assert np.array_equal(d0, (direct / np.maximum(albedo, eps)) * albedo)
    assert np.array_equal(i0, (indirect / np.maximum(albedo, eps)) * albedo)

"""EATWT denoiser on the GPU (k_eatwt_pass) against the oracle's restatement
of denoiseSpectralRadianceEATWT (cpu_denoise.ipp:10-187).

Bar: 1e-5 relative per bin (host and device expf differ by an ulp or so; the
weighted averages carry that through the levels), with pixels without a hit
bit-exact.  Cases: the passes of a real render (square, as the reference
defines the filter), a non-square image with misses and different level
counts, zero levels, and render(denoise=True) against render + denoise."""
import numpy as np
import pytest

import pyoracle as O
from vira_amd import scene as S
from vira_amd import workloads as W

pytestmark = pytest.mark.gpu


def _to_dev(*arrays):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def _passes_from(arrays, with_total=True):
    direct, indirect, albedo, depth, normal = arrays
    H, Wd = depth.shape
    p = S.RenderPasses(Wd, H, which=("radiance", "direct", "indirect", "albedo", "normal", "depth")
                       if with_total else ("direct", "indirect", "albedo", "normal", "depth"))
    for t, a in ((p.direct_radiance, direct), (p.indirect_radiance, indirect), (p.albedo, albedo),
                 (p.depth, depth), (p.normal_global, normal)):
        t.copy_(_to_dev(a)[0])
    return p


def _check(got, want, miss=None):
    got = got.astype(np.float64)
    want = want.astype(np.float64)
    fin = np.isfinite(want)
    assert np.array_equal(np.isfinite(got), fin)
    scale = np.abs(want[fin]).max()
    lit = fin & (np.abs(want) > 1e-6 * scale)
    rel = np.abs(got - want)[lit] / np.abs(want[lit])
    assert rel.max() <= 1e-5, np.percentile(rel, [50, 99, 100])
    if miss is not None:
        assert np.array_equal(got[miss], want[miss])


def _random(H, Wd, seed):
    rng = np.random.default_rng(seed)
    direct = rng.random((H, Wd, 8)).astype(np.float32)
    indirect = (0.3 * rng.random((H, Wd, 8))).astype(np.float32)
    albedo = (0.05 + rng.random((H, Wd, 8))).astype(np.float32)
    depth = (50.0 + 30.0 * rng.random((H, Wd))).astype(np.float32)
    n = rng.normal(size=(H, Wd, 3))
    n[..., 2] = np.abs(n[..., 2]) + 0.5
    normal = (n / np.linalg.norm(n, axis=2, keepdims=True)).astype(np.float32)
    miss = rng.random((H, Wd)) < 0.05
    depth[miss] = np.inf
    normal[miss] = 0.0
    for a in (direct, indirect, albedo):
        a[miss] = 0.0
    return [direct, indirect, albedo, depth, normal], miss


def test_denoise_render_passes(gpu_ctx):
    import torch
    b = W.build(gpu_ctx, W.scaled(W.DEM257, image=256))
    passes = S.RenderPasses(256, 256)
    gpu_ctx.render(b.camera, S.PathTracerOptions(samples=4, bounces=2, seed=7), passes)
    torch.cuda.synchronize()
    host = [t.cpu().numpy() for t in (passes.direct_radiance, passes.indirect_radiance, passes.albedo,
                                      passes.depth, passes.normal_global)]
    gpu_ctx.denoise(passes)
    torch.cuda.synchronize()
    d, i = O.denoise(*host)
    miss = np.isinf(host[3])
    _check(passes.direct_radiance.cpu().numpy(), d, miss)
    _check(passes.indirect_radiance.cpu().numpy(), i, miss)
    tot = passes.total_radiance.cpu().numpy()
    assert np.array_equal(tot, passes.direct_radiance.cpu().numpy() + passes.indirect_radiance.cpu().numpy())
    # the denoiser lowers the pixel-to-pixel variation of the indirect pass
    before = np.abs(np.diff(host[1], axis=1)).mean()
    after = np.abs(np.diff(passes.indirect_radiance.cpu().numpy(), axis=1)).mean()
    print(f"indirect: mean |dI/dx| {before:.3e} -> {after:.3e}")
    assert after < before


@pytest.mark.parametrize("levels", [(3, 5), (5, 2), (0, 2), (0, 0), (1, 1)])
def test_denoise_non_square_levels(gpu_ctx, levels):
    import torch
    arrays, miss = _random(72, 136, 21)
    p = _passes_from(arrays)
    gpu_ctx.denoise(p, S.DenoiseOptions(max_levels_direct=levels[0], max_levels_indirect=levels[1]))
    torch.cuda.synchronize()
    d, i = O.denoise(*arrays, levels=levels)
    _check(p.direct_radiance.cpu().numpy(), d, miss)
    _check(p.indirect_radiance.cpu().numpy(), i, miss)
    assert np.array_equal(p.total_radiance.cpu().numpy(),
                          p.direct_radiance.cpu().numpy() + p.indirect_radiance.cpu().numpy())


def test_render_with_denoise_option(gpu_ctx):
    import torch
    b = W.build(gpu_ctx, W.scaled(W.DEM257, image=128))
    opts = S.PathTracerOptions(samples=2, bounces=1, seed=3)
    p1 = S.RenderPasses(128, 128)
    gpu_ctx.render(b.camera, opts, p1)
    gpu_ctx.denoise(p1, S.DenoiseOptions(max_levels_direct=3))
    p2 = S.RenderPasses(128, 128)
    opts.denoise = True
    opts.denoiser = S.DenoiseOptions(max_levels_direct=3)
    gpu_ctx.render(b.camera, opts, p2)
    torch.cuda.synchronize()
    for a, c in ((p1.total_radiance, p2.total_radiance), (p1.direct_radiance, p2.direct_radiance),
                 (p1.indirect_radiance, p2.indirect_radiance)):
        assert torch.equal(a, c)
    with pytest.raises(ValueError):
        gpu_ctx.render(b.camera, opts, S.RenderPasses(128, 128, which=("radiance",)))

"""vira_gpu_render_noalloc (scratch in the library's fixed per-GPU array) against
vira_gpu_render with the same per-pixel integrator: every pass bit for bit, on
two calls in a row, the second with a smaller image, so state the shared
scratch keeps from the first call cannot leak into the second.  Called through
the C-ABI (vira_amd.scene.Context wraps it with ctypes)."""
import numpy as np
import pytest

from vira_amd import ViraGPUError
from vira_amd import scene as S
from vira_amd import workloads as W

pytestmark = pytest.mark.gpu

PASSES = ("total_radiance", "direct_radiance", "indirect_radiance", "albedo", "normal_global", "depth", "alpha",
          "triangle_id", "instance_id", "normal_camera", "mesh_id", "material_id")


def _both(ctx, built, image):
    built.options.integrator = "fused"
    a = S.RenderPasses(image, image)
    sa = ctx.render(built.camera, built.options, a)
    b = S.RenderPasses(image, image)
    sb = ctx.render_noalloc(built.camera, built.options, b)
    return a, sa, b, sb


def _same(a, b):
    for name in PASSES:
        x, y = getattr(a, name), getattr(b, name)
        if x is None:
            continue
        x, y = x.cpu().numpy(), y.cpu().numpy()
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name


def test_noalloc_matches_render_two_calls(gpu_ctx):
    # call 1: 256^2 (1024 tiles), sphere sun, texture, 2 bounces + RR
    sc1 = W.scaled(W.DEM4097, dem=129, image=256, spp=3, texture=64)
    b1 = W.build(gpu_ctx, sc1)
    a, sa, b, sb = _both(gpu_ctx, b1, sc1.image)
    _same(a, b)
    assert (a.triangle_id.cpu().numpy().view(np.uint32) != 0xFFFFFFFF).mean() > 0.5
    assert (sa["primary_rays"], sa["shadow_rays"], sa["bounce_rays"]) == \
           (sb["primary_rays"], sb["shadow_rays"], sb["bounce_rays"])
    assert sb["integrator"] == 1 and sb["total_ms"] > 0
    # call 2: a smaller image (72^2: a ragged 9 x 9 tiles) of another scene, point sun, no bounces
    sc2 = W.scaled(W.DEM257, dem=65, image=72, spp=2)
    b2 = W.build(gpu_ctx, sc2)
    a, sa, b, sb = _both(gpu_ctx, b2, sc2.image)
    _same(a, b)
    assert sa["primary_rays"] == sb["primary_rays"] == 72 * 72 * 2


def test_noalloc_instanced_matches_render(gpu_ctx):
    isc = W.build_instanced(gpu_ctx, dem=65, image=64, spp=2, bounces=1, sphere=(21, 32), body=(21, 32))
    isc.options.integrator = "fused"
    a = S.RenderPasses(64, 64)
    gpu_ctx.render(isc.camera, isc.options, a)
    b = S.RenderPasses(64, 64)
    gpu_ctx.render_noalloc(isc.camera, isc.options, b)
    _same(a, b)


def test_noalloc_limits(gpu_ctx):
    sc = W.scaled(W.DEM257, dem=65, image=520, spp=1)  # 65 x 65 = 4225 tiles > 4096
    built = W.build(gpu_ctx, sc)
    with pytest.raises(ViraGPUError) as e:
        gpu_ctx.render_noalloc(built.camera, built.options, S.RenderPasses(520, 520))
    assert e.value.status == 1 and "4096" in str(e.value)
    built.options.integrator = "wavefront"
    with pytest.raises(ViraGPUError) as e:
        gpu_ctx.render_noalloc(built.camera, built.options, S.RenderPasses(64, 64))
    assert e.value.status == 5
    built.options.integrator = "auto"


def test_noalloc_several_lights_matches_render(gpu_ctx):
    """The several-light variant of the per-pixel renderer through the fixed scratch."""
    sc = W.scaled(W.DEM257, dem=65, image=80, spp=2, bounces=1)
    built = W.build(gpu_ctx, sc)
    gpu_ctx.lights_set(built.lights + [S.directional_sun(50.0, 200.0, 1.5, distance=1e6)])
    try:
        a, sa, b, sb = _both(gpu_ctx, built, sc.image)
        _same(a, b)
        assert sa["shadow_rays"] == sb["shadow_rays"] > 0
    finally:
        gpu_ctx.lights_set(built.lights)

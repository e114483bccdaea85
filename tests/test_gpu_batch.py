"""Batched rendering (the 64-request mixed batch and scaled copies): one persistent,
graph-captured launch over many requests must write, for every request, the
passes the per-request render writes -- bit for bit -- and the small-body mesh
must render as the oracle does."""
from dataclasses import replace

import numpy as np
import pytest

import pyoracle as O
from vira_amd import scene as S
from vira_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_meshes(gpu_ctx):
    ids, scenes, keep = [], [], []
    for sc, dem in ((W.DEM257, 65), (W.DEM4097, 129), (W.DEM1025_DI, 65)):
        s2 = W.scaled(sc, dem=dem, texture=64 if sc.texture else None)
        b = W.build(gpu_ctx, s2)
        ids.append(b.mesh)
        scenes.append(s2)
        keep += [b.heights, b.texture, b.vertex_albedo]
    pos, idx = W.small_body_mesh(nlat=41, nlon=64)
    mid = gpu_ctx.mesh_upload(pos, idx)
    gpu_ctx.blas_build(mid)
    gpu_ctx.material_set(mid, albedo=0.12)
    ids.append(mid)
    scenes.append(None)
    return W.BatchMeshes(ids, scenes, keep)


def _passes_equal(a: S.RenderPasses, b: S.RenderPasses):
    for name in ("total_radiance", "direct_radiance", "indirect_radiance", "albedo", "normal_global", "depth",
                 "alpha", "triangle_id"):
        x, y = getattr(a, name), getattr(b, name)
        if x is None:
            continue
        xa, ya = x.cpu().numpy(), y.cpu().numpy()
        if not np.array_equal(xa.view(np.uint32), ya.view(np.uint32)):
            return name
    return None


def _clone(reqs):
    out = []
    for r in reqs:
        out.append(replace(r, passes=S.RenderPasses(r.passes.width, r.passes.height)))
    return out


def test_batch_equals_per_request_renders(gpu_ctx, small_meshes):
    reqs = W.make_batch(gpu_ctx, small_meshes, n=16, seed=5, bounces=2, passes=S.RenderPasses.ALL,
                        sides=(16, 24, 40, 64), spps=(1, 2, 5))
    plain = _clone(reqs)
    hits = 0
    for r in plain:
        gpu_ctx.tlas_build([r.mesh])
        gpu_ctx.lights_set([] if r.light is None else [r.light])
        hits += gpu_ctx.render(r.camera, r.options, r.passes)["closest_hits"]
    st = gpu_ctx.batch_submit(reqs)
    assert st["closest_hits"] == hits > 0
    assert st["graph_reused"] == 0
    for i, (a, b) in enumerate(zip(reqs, plain)):
        bad = _passes_equal(a.passes, b.passes)
        assert bad is None, f"request {i}: pass {bad} differs"
    prim = sum(r.passes.width * r.passes.height * r.options.samples for r in reqs)
    assert st["primary_rays"] == prim
    # the same layout again replays the captured graph and writes the same passes
    again = _clone(reqs)
    st2 = gpu_ctx.batch_submit(again)
    assert st2["graph_reused"] == 1
    for a, b in zip(again, plain):
        assert _passes_equal(a.passes, b.passes) is None
    # submission order and grab size do not change any pixel
    cfg = S.BatchConfig(bvh_width=4, waves_per_block=2, waves_per_cu=8, tiles_per_grab=4, order=1)
    third = _clone(reqs)
    gpu_ctx.batch_submit(third, cfg)
    for a, b in zip(third, plain):
        assert _passes_equal(a.passes, b.passes) is None


def test_batch_autotune_keeps_results(gpu_ctx, small_meshes):
    reqs = W.make_batch(gpu_ctx, small_meshes, n=12, seed=6, bounces=1, sides=(32, 64), spps=(1, 4))
    cfg, ms = gpu_ctx.batch_autotune(reqs, sample_stride=2)
    assert cfg.bvh_width in (2, 4) and cfg.waves_per_block in (1, 2) and ms > 0
    gpu_ctx.batch_submit(reqs)  # tuned configuration
    plain = _clone(reqs)
    W.render_plain(gpu_ctx, plain)
    for a, b in zip(reqs, plain):
        assert _passes_equal(a.passes, b.passes) is None
    # leave the meshes four-wide for the other tests
    for m in small_meshes.ids:
        gpu_ctx.blas_build(m, width=4)


def test_batch_errors(gpu_ctx, small_meshes):
    reqs = W.make_batch(gpu_ctx, small_meshes, n=2, seed=7, sides=(16,), spps=(1,))
    with pytest.raises(Exception) as e:
        gpu_ctx.batch_submit(reqs, S.BatchConfig(bvh_width=2))
    assert "width" in str(e.value)
    with pytest.raises(Exception):
        gpu_ctx.batch_submit(reqs, S.BatchConfig(waves_per_block=3))


def test_small_body_matches_oracle(gpu_ctx):
    import torch
    pos, idx = W.small_body_mesh(nlat=61, nlon=96)
    mid = gpu_ctx.mesh_upload(pos, idx)
    gpu_ctx.blas_build(mid)
    gpu_ctx.tlas_build([mid])
    gpu_ctx.material_set(mid, albedo=0.12)
    sun = S.directional_sun(35.0, 20.0, 1.0)
    gpu_ctx.lights_set([sun])
    cam = S.Camera(96, 96, 30.0, (0.0, -2000.0, 600.0))
    cam.look_in_direction((0.0, 2000.0, -600.0))
    passes = S.RenderPasses(96, 96)
    gpu_ctx.render(cam, S.PathTracerOptions(samples=1, bounces=0), passes)
    torch.cuda.synchronize()
    _, gidx, gnrm = gpu_ctx.mesh_download(mid)
    om = O.Mesh(pos, idx, O.vertex_normals(pos, idx), None, None)
    assert np.array_equal(gnrm.view(np.uint32), O.vertex_normals(pos, idx).view(np.uint32))
    out = om.render(O.camera_from(cam.to_c()), [O.light_from(sun.to_c())], O.material(0.12, None, False),
                    O.options(1, 0, False, 1, 0))
    tri = passes.triangle_id.cpu().numpy().reshape(-1).view(np.uint32)
    assert np.array_equal(tri, out["tri_id"])
    assert (tri != 0xFFFFFFFF).sum() > 1000
    rad = passes.total_radiance.cpu().numpy().reshape(-1, 8)
    assert np.array_equal(rad.view(np.uint32), out["radiance"].view(np.uint32))

"""Two-level instanced scenes (the 1 AU Moon scene and scaled copies): four bodies placed by
FP64 model matrices about 1.5e8 km from the origin, rendered camera-relative.

Bar: the oracle's FP32 twin (oracle render_scene, the same FP64 rebasing
rounded to FP32) gives the same first-hit triangle and instance ids and, for
a point or sphere sun, with or without bounces, the same radiance bit for bit
(sampling angles go through the shared sin/cos of fp_sincos.hpp).  The error
against the FP64 oracle (every matrix and position kept in FP64) is stated."""
import numpy as np
import pytest

import pyoracle as O
from vira_amd import scene as S
from vira_amd import workloads as W

pytestmark = pytest.mark.gpu


def _oracle_meshes(ctx, isc, dem):
    sc = W.scaled(W.DEM4097, dem=dem, texture=4 * (dem - 1)) if dem != W.DEM4097.dem else W.DEM4097
    heights = isc.keep[0].cpu().numpy()
    tex = isc.keep[1].cpu().numpy()
    p = W.make_projection(sc)
    oproj = O.make_projection(p.kind, p.spacing, p.radius, tuple(p.center), tuple(p.offset))
    tile = O.dem_mesh(heights, oproj, None)
    ms = [tile]
    for key in ("sphere", "body", "craft"):
        pos, idx = isc.local[key]
        ms.append(O.Mesh(pos, idx, O.vertex_normals(pos, idx), None, None))
    mats = [O.material(1.0, tex, False), O.material(0.11), O.material(0.12), O.material(0.5)]
    return ms, mats


def _render(ctx, isc, opts=None):
    import torch
    passes = S.RenderPasses(isc.camera.width, isc.camera.height)
    st = ctx.render(isc.camera, opts or isc.options, passes)
    torch.cuda.synchronize()
    return passes, st


@pytest.mark.parametrize("width", [4, 2])
def test_instanced_point_sun_bit_exact(gpu_ctx, width):
    isc = W.build_instanced(gpu_ctx, dem=129, image=128, spp=1, bounces=0, sphere=(41, 64), body=(41, 64),
                            sun="point", bvh_width=width)
    passes, st = _render(gpu_ctx, isc)
    ms, mats = _oracle_meshes(gpu_ctx, isc, 129)
    cam = O.camera_from(isc.camera.to_c())
    lights = [O.light_from(L.to_c()) for L in isc.lights]
    opt = O.options(1, 0, False, 1, 4)
    out = O.render_scene(ms, isc.transforms, mats, cam, lights, opt, precision=32)
    tri = passes.triangle_id.cpu().numpy().reshape(-1).view(np.uint32)
    inst = passes.instance_id.cpu().numpy().reshape(-1).view(np.uint32)
    assert np.array_equal(tri, out["tri_id"])
    assert np.array_equal(inst, out["instance_id"])
    seen = set(np.unique(inst[inst != 0xFFFFFFFF]).tolist())
    assert {0, 2, 3} <= seen, seen  # tile, small body and the spacecraft's boom are in view
    rad = passes.total_radiance.cpu().numpy().reshape(-1, 8)
    assert np.array_equal(rad.view(np.uint32), out["radiance"].view(np.uint32))
    assert (rad.max(1) > 0).sum() > 1000
    # FP64 oracle: every matrix and position in FP64
    o64 = O.render_scene(ms, isc.transforms, mats, cam, lights, opt, precision=64)
    same = np.mean(o64["tri_id"] == tri)
    both = (o64["tri_id"] == tri) & (tri != 0xFFFFFFFF)
    r64 = o64["radiance"][both]
    rel = np.abs(rad[both] - r64) / np.maximum(np.abs(r64), 1e-30)
    lit = r64 > 1e-3 * r64.max()
    print(f"BVH{width}: ids equal to FP64 oracle on {same:.4%} of pixels; radiance vs FP64 on shared hits: "
          f"median {np.median(rel[lit]):.2e}, 99th pct {np.percentile(rel[lit], 99):.2e}")
    assert same >= 0.99
    assert np.median(rel[lit]) <= 1e-6 and np.percentile(rel[lit], 99) <= 5e-5


def test_instanced_sphere_sun_bounces(gpu_ctx):
    isc = W.build_instanced(gpu_ctx, dem=129, image=96, spp=2, bounces=2, sphere=(41, 64), body=(41, 64))
    passes, st = _render(gpu_ctx, isc)
    assert st["bounce_rays"] > 0
    ms, mats = _oracle_meshes(gpu_ctx, isc, 129)
    out = O.render_scene(ms, isc.transforms, mats, O.camera_from(isc.camera.to_c()),
                         [O.light_from(L.to_c()) for L in isc.lights], O.options(2, 2, False, 1, 4), precision=32)
    tri = passes.triangle_id.cpu().numpy().reshape(-1).view(np.uint32)
    assert np.array_equal(tri, out["tri_id"])
    assert np.array_equal(passes.instance_id.cpu().numpy().reshape(-1).view(np.uint32), out["instance_id"])
    rad = passes.total_radiance.cpu().numpy().reshape(-1, 8)
    assert (out["radiance"].max(1) > 0).mean() > 0.1
    assert np.array_equal(rad.view(np.uint32), out["radiance"].view(np.uint32))


def test_tlas_validation(gpu_ctx):
    isc = W.build_instanced(gpu_ctx, dem=65, image=32, spp=1, bounces=0, sphere=(21, 32), body=(21, 32),
                            sun="point")
    bad = np.eye(4)
    bad[0, 0] = 0.0
    with pytest.raises(Exception) as e:
        gpu_ctx.tlas_build(isc.meshes[:2], [np.eye(4), bad])
    assert "singular" in str(e.value)
    proj = np.eye(4)
    proj[3, 0] = 1.0
    with pytest.raises(Exception) as e:
        gpu_ctx.tlas_build(isc.meshes[:1], [proj])
    assert "affine" in str(e.value)
    with pytest.raises(Exception) as e:
        gpu_ctx.tlas_build(isc.meshes[:1] * 65537, None)  # beyond the TLAS's 65536 instances
    assert "65536" in str(e.value)
    # the direct-illumination path takes one identity instance
    gpu_ctx.tlas_build(isc.meshes, isc.transforms)
    with pytest.raises(Exception) as e:
        gpu_ctx.di_render(isc.camera, S.DIOptions())
    assert "identity" in str(e.value)


def test_instanced_full_size_sample(gpu_ctx):
    """The instanced scene at full size (33.5M-triangle tile, 4096^2, 8 spp, 2 bounces):
    2,048 random pixels against the oracle's FP32 twin."""
    isc = W.build_instanced(gpu_ctx)
    passes, st = _render(gpu_ctx, isc)
    ms, mats = _oracle_meshes(gpu_ctx, isc, 4097)
    rng = np.random.default_rng(0)
    pix = rng.choice(4096 * 4096, 2048, replace=False).astype(np.int32)
    out = O.render_scene(ms, isc.transforms, mats, O.camera_from(isc.camera.to_c()),
                         [O.light_from(L.to_c()) for L in isc.lights], O.options(8, 2, False, 1, 4), precision=32,
                         pixels=pix)
    tri = passes.triangle_id.cpu().numpy().reshape(-1).view(np.uint32)[pix]
    inst = passes.instance_id.cpu().numpy().reshape(-1).view(np.uint32)[pix]
    assert np.array_equal(tri, out["tri_id"][pix]) and np.array_equal(inst, out["instance_id"][pix])
    rad = passes.total_radiance.cpu().numpy().reshape(-1, 8)[pix]
    ref = out["radiance"][pix]
    print(f"instanced scene: {st['primary_rays'] + st['shadow_rays']} primary+shadow rays in {st['total_ms']:.1f} ms")
    assert (ref.max(1) > 0).sum() > 100
    assert np.array_equal(rad.view(np.uint32), ref.view(np.uint32))


def test_tlas_256_instances_bit_exact(gpu_ctx):
    """A TLAS of 256 instances of two meshes (a small DEM and the small body), each
    rotated and placed on a jittered grid 1e7 m from the origin, seen camera-
    relative with a point sun and one bounce: triangle ids, instance ids, mesh ids
    and radiance equal the oracle's (which tests every instance in index order,
    ties to the lower instance), whatever order the TLAS visits them in."""
    import math
    rng = np.random.default_rng(8)
    sc = W.scaled(W.DEM257, dem=33, image=8)
    built = W.build(gpu_ctx, sc)
    pos, idx = W.small_body_mesh(3, nlat=21, nlon=32)
    body = gpu_ctx.mesh_upload(pos, idx)
    gpu_ctx.blas_build(body)
    gpu_ctx.material_set(body, albedo=0.2, material_id=7)
    meshes, transforms = [], []
    base = np.array([1.0e7, -2.0e6, 3.0e6])
    for i in range(256):
        a, b = rng.uniform(0, 2 * math.pi, 2)
        Rz = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
        Rx = np.array([[1, 0, 0], [0, math.cos(b), -math.sin(b)], [0, math.sin(b), math.cos(b)]])
        M = np.eye(4)
        M[:3, :3] = Rz @ Rx
        M[:3, 3] = base + np.array([(i % 16) * 2500.0, (i // 16) * 2500.0, 0.0]) + rng.uniform(-300, 300, 3)
        meshes.append(built.mesh if i % 3 else body)
        transforms.append(M)
    gpu_ctx.tlas_build(meshes, transforms)
    centre = base + np.array([7.5 * 2500, 7.5 * 2500, 0.0])
    cam = S.Camera(96, 96, 55.0, tuple(centre + np.array([0.0, -30000.0, 25000.0]))).look_in_direction(
        tuple(np.array([0.0, 30000.0, -25000.0])))
    sun = S.directional_sun(50.0, 30.0, 2.0, distance=1e10, origin=tuple(centre))
    gpu_ctx.lights_set([sun])
    opts = S.PathTracerOptions(samples=2, bounces=1, seed=5)
    passes = S.RenderPasses(96, 96)
    gpu_ctx.render(cam, opts, passes)
    heights = built.heights.cpu().numpy()
    p = W.make_projection(sc)
    tile = O.dem_mesh(heights, O.make_projection(p.kind, p.spacing, p.radius, tuple(p.center), tuple(p.offset)),
                      built.vertex_albedo.cpu().numpy() if built.vertex_albedo is not None else None)
    om_body = O.Mesh(pos, idx, O.vertex_normals(pos, idx), None, None)
    ms = [tile if m == built.mesh else om_body for m in meshes]
    tex = None if built.texture is None else built.texture.cpu().numpy()
    mats = [O.material(1.0, tex, sc.smooth_shading) if m == built.mesh else O.material(0.2, material_id=7)
            for m in meshes]
    out = O.render_scene(ms, transforms, mats, O.camera_from(cam.to_c()), [O.light_from(sun.to_c())],
                         O.options(2, 1, False, 1, 5), precision=32)
    inst = passes.instance_id.cpu().numpy().reshape(-1).view(np.uint32)
    hit = out["tri_id"] != 0xFFFFFFFF
    assert hit.mean() > 0.1 and len(np.unique(out["instance_id"][hit])) > 200  # 13 % of pixels, 224 instances
    assert np.array_equal(passes.triangle_id.cpu().numpy().reshape(-1).view(np.uint32), out["tri_id"])
    assert np.array_equal(inst, out["instance_id"])
    rad = passes.total_radiance.cpu().numpy().reshape(-1, 8)
    assert np.array_equal(rad.view(np.uint32), out["radiance"].view(np.uint32))
    mid = passes.mesh_id.cpu().numpy().reshape(-1).view(np.uint32)
    want = np.where(hit, np.array(meshes, np.uint32)[np.where(hit, out["instance_id"], 0)], 0xFFFFFFFF)
    assert np.array_equal(mid, want)
    matid = passes.material_id.cpu().numpy().reshape(-1).view(np.uint32)
    assert np.all(matid[hit & (mid == body)] == 7)

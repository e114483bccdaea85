"""Opt-in BVH visit counters of the per-pixel renderer (render option
count_visits), through the C-ABI.

Pinned by cases whose counts follow from the tree's shape:
* every ray misses every box below the root (camera looking away from the
  terrain): one internal node per camera ray, no leaf, no triangle, empty stack;
* two triangles 95 m apart in leaves of one triangle (the tree is a root and
  two leaves) seen from above within one of them: one node, one leaf and one
  triangle test per camera ray, nothing pushed;
and by invariants on a DEM: the counting kernel gives the same passes bit for
bit, every traversal fetches the root, every hit lies in an entered leaf, a
leaf holds one to fifteen triangles, pushes past the 8 LDS entries happen
exactly when the deepest stack passes 8. The oracle's BVH2 traversal counts of
the same render are printed beside the GPU's (its tree and pop rule differ, so
the totals are compared only for scale).
"""
from dataclasses import replace

import numpy as np
import pytest

import pyoracle as O
from vira_amd import ViraGPUError
from vira_amd import scene as S
from vira_amd import workloads as W

pytestmark = pytest.mark.gpu

PASSES = ("total_radiance", "direct_radiance", "indirect_radiance", "albedo", "normal_global", "depth", "alpha",
          "triangle_id")


@pytest.mark.parametrize("width", [4, 2])
def test_counts_leave_the_image_unchanged_and_hold_invariants(gpu_ctx, width):
    sc = W.scaled(W.DEM4097, dem=129, image=64, spp=4, texture=64)
    built = W.build(gpu_ctx, sc, bvh_width=width)
    rp0 = S.RenderPasses(64, 64)
    st0 = gpu_ctx.render(built.camera, built.options, rp0)
    rp1 = S.RenderPasses(64, 64)
    st1 = gpu_ctx.render(built.camera, replace(built.options, count_visits=True), rp1)
    for name in PASSES:
        a, b = getattr(rp0, name).cpu().numpy(), getattr(rp1, name).cpu().numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    assert st0["visits_counted"] == 0 and st1["visits_counted"] == 1
    for k in ("primary_rays", "shadow_rays", "bounce_rays", "closest_hits"):
        assert st0[k] == st1[k], k
    c, s = st1["closest_visits"], st1["shadow_visits"]
    n_closest = st1["primary_rays"] + st1["bounce_rays"]
    assert c["nodes"] >= n_closest and s["nodes"] >= st1["shadow_rays"]
    assert c["leaves"] >= st1["closest_hits"]
    for v in (c, s):
        assert v["leaves"] <= v["triangles"] <= 15 * v["leaves"]
        assert (v["stack_spills"] > 0) == (v["max_stack"] > 8)
        assert v["tlas_nodes"] == 0
    # the oracle's BVH2 traversal of the same render, for scale
    om = O.dem_mesh(built.heights.cpu().numpy(), O.make_projection(*_proj_args(sc)), None)
    o = om.render(O.camera_from(built.camera.to_c()), [O.light_from(L.to_c()) for L in built.lights],
                  O.material(1.0, built.texture.cpu().numpy(), sc.smooth_shading),
                  O.options(sc.spp, sc.bounces, sc.russian_roulette, 1, sc.render_seed))
    oc = o["counts"]
    g_int, o_int = c["nodes"] / n_closest, oc["closest_internal"] / n_closest
    print(f"BVH{width}: per closest ray {g_int:.2f} nodes, {c['leaves'] / n_closest:.2f} leaves, "
          f"{c['triangles'] / n_closest:.2f} triangles, max stack {c['max_stack']}, spills {c['stack_spills']}; "
          f"per shadow ray {s['nodes'] / st1['shadow_rays']:.2f} nodes, {s['triangles'] / st1['shadow_rays']:.2f} "
          f"triangles; oracle BVH2 {o_int:.2f} internal, {oc['closest_tris'] / n_closest:.2f} triangles per closest ray")
    if width == 2:
        assert 0.5 < g_int / o_int < 2.0


def _proj_args(sc):
    p = W.make_projection(sc)
    return p.kind, p.spacing, p.radius, tuple(p.center), tuple(p.offset)


def test_counts_when_every_box_is_missed(gpu_ctx):
    sc = W.scaled(W.DEM4097, dem=65, image=32, spp=2)
    built = W.build(gpu_ctx, sc)
    cam = S.Camera(32, 32, 30.0, built.camera.position).look_in_direction((1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0))
    rp = S.RenderPasses(32, 32)
    st = gpu_ctx.render(cam, replace(built.options, count_visits=True), rp)
    assert np.all(rp.triangle_id.cpu().numpy().view(np.uint32) == 0xFFFFFFFF)
    c = st["closest_visits"]
    assert st["primary_rays"] == 32 * 32 * 2 and st["shadow_rays"] == 0 and st["bounce_rays"] == 0
    assert c["nodes"] == st["primary_rays"]
    assert c["leaves"] == c["triangles"] == c["stack_spills"] == c["max_stack"] == 0


@pytest.mark.parametrize("width", [4, 2])
def test_counts_on_a_root_with_two_leaves(gpu_ctx, width):
    pos = np.array([[0, 0, 0], [10, 0, 0], [0, 10, 0], [100, 0, 0], [110, 0, 0], [100, 10, 0]], np.float32)
    idx = np.arange(6, dtype=np.uint32).reshape(2, 3)
    mid = gpu_ctx.mesh_upload(pos, idx)
    try:
        bst = gpu_ctx.blas_build(mid, max_leaf=1, width=width)
        assert bst["max_leaf_size"] == 1
        gpu_ctx.tlas_build([mid])
        gpu_ctx.lights_set([])
        cam = S.Camera(16, 16, 4.0, (3.0, 3.0, 50.0)).look_in_direction((0.0, 0.0, -1.0))
        rp = S.RenderPasses(16, 16)
        st = gpu_ctx.render(cam, S.PathTracerOptions(samples=3, count_visits=True), rp)
        assert np.all(rp.triangle_id.cpu().numpy().view(np.uint32) == 0)
        n = st["primary_rays"]
        c = st["closest_visits"]
        assert n == 16 * 16 * 3 and st["closest_hits"] == n
        assert c["nodes"] == n and c["leaves"] == n and c["triangles"] == n
        assert c["stack_spills"] == 0 and c["max_stack"] == 0
    finally:
        gpu_ctx.mesh_release(mid)


def test_counting_rejected_where_not_built(gpu_ctx):
    sc = W.scaled(W.DEM257, dem=33, image=16)
    built = W.build(gpu_ctx, sc)
    opt = replace(built.options, count_visits=True)
    with pytest.raises(ViraGPUError) as e:
        gpu_ctx.render(built.camera, replace(opt, integrator="wavefront"), S.RenderPasses(16, 16))
    assert e.value.status == 5
    gpu_ctx.lights_set(built.lights * 2)
    try:
        with pytest.raises(ViraGPUError) as e:
            gpu_ctx.render(built.camera, opt, S.RenderPasses(16, 16))
        assert e.value.status == 5
    finally:
        gpu_ctx.lights_set(built.lights)
    req = S.RenderRequest(built.mesh, built.camera, opt, built.lights[0], S.RenderPasses(16, 16))
    with pytest.raises(ViraGPUError) as e:
        gpu_ctx.batch_submit([req])
    assert e.value.status == 5

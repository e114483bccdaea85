"""The quantised node copy (blas_build compress=True: 64 B Node4Q with 8-bit
child boxes rounded outward) against the 128 B Node4 tree it is made from.

A decoded box always holds the child's box, so the traversal reaches every
triangle it reached before and the closest hit (ties to the lowest id) is the
same: t, id and barycentrics bit for bit, and the same occlusion, on the grid
tree of a DEM, the binned-SAH tree of the same DEM and a generic mesh; renders
(per-pixel integrator, plain and instanced) give the same passes; a refit
re-encodes the copy; width 2 refuses compression.
"""
from dataclasses import replace

import numpy as np
import pytest

import pyoracle as O
from vira_amd import ViraGPUError
from vira_amd import scene as S
from vira_amd import workloads as W

pytestmark = pytest.mark.gpu


def _trace(ctx, mid, rays6, any_hit=False, tmax=None):
    import torch
    r8 = np.zeros((rays6.shape[0], 8), np.float32)
    r8[:, 0:3], r8[:, 4:7] = rays6[:, :3], rays6[:, 3:]
    r8[:, 7] = np.inf if tmax is None else tmax
    t, ids, b1, b2, _ = ctx.trace_rays(mid, torch.from_numpy(r8).cuda(), any_hit=any_hit)
    return [x.cpu().numpy().view(np.uint32) for x in (t, ids, b1, b2)]


def _rays(n, lo, hi, z0, z1, rng):
    o = np.stack([rng.uniform(lo[0], hi[0], n), rng.uniform(lo[1], hi[1], n), rng.uniform(z0, z1, n)], 1)
    tgt = np.stack([rng.uniform(lo[0], hi[0], n), rng.uniform(lo[1], hi[1], n), np.full(n, lo[2])], 1)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


def _compare(ctx, mid, rays, method, oracle_mesh=None):
    plain = ctx.blas_build(mid, method=method)
    h0 = _trace(ctx, mid, rays)
    st = ctx.blas_build(mid, method=method, compress=True)
    assert st["compressed_bytes"] == 64 * st["num_wide_nodes"] and plain["compressed_bytes"] == 0
    h1 = _trace(ctx, mid, rays)
    for a, b in zip(h1, h0):
        assert np.array_equal(a, b)
    if oracle_mesh is not None:
        _, oid, _, _, _ = oracle_mesh.trace(rays)
        assert np.array_equal(h1[1], oid)
    t0 = h0[0].view(np.float32)
    tms = [np.where(np.isfinite(t0), t0 * f, 1e30).astype(np.float32) for f in (0.99, 1.01)]
    occ1 = [_trace(ctx, mid, rays, any_hit=True, tmax=tm)[1] != 0xFFFFFFFF for tm in tms]
    ctx.blas_build(mid, method=method)
    occ0 = [_trace(ctx, mid, rays, any_hit=True, tmax=tm)[1] != 0xFFFFFFFF for tm in tms]
    for a, b in zip(occ1, occ0):
        assert np.array_equal(a, b)
    assert occ0[1].sum() == (h0[1] != 0xFFFFFFFF).sum() and not occ0[0].any()
    return h1


@pytest.mark.parametrize("method", ["grid", "sah"])
def test_compressed_dem_trees_hit_like_the_full_ones(gpu_ctx, method):
    h = gpu_ctx.synth_fbm(129, 97, 6)
    h[30:36, 40:52] = float("nan")
    proj = S.projection(0, 30.0)
    mid = gpu_ctx.dem_mesh(h, proj)
    try:
        om = O.dem_mesh(h.cpu().numpy(), O.make_projection(0, 30.0), None)
        rays = _rays(80000, (-1920.0, -1440.0, -700.0), (1920.0, 1440.0, 0.0), 500.0, 9000.0, np.random.default_rng(2))
        hits = _compare(gpu_ctx, mid, rays, method, om)
        assert (hits[1] != 0xFFFFFFFF).mean() > 0.5
    finally:
        gpu_ctx.mesh_release(mid)


def test_compressed_generic_mesh(gpu_ctx):
    pos, idx = W.small_body_mesh(nlat=61, nlon=90)
    mid = gpu_ctx.mesh_upload(pos, idx)
    try:
        rng = np.random.default_rng(7)
        n = 60000
        o = rng.normal(size=(n, 3))
        o = o / np.linalg.norm(o, axis=1, keepdims=True) * 2000.0
        d = rng.normal(size=(n, 3)) * 150.0 - o
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = np.concatenate([o, d], 1).astype(np.float32)
        om = O.Mesh(pos, idx)
        hits = _compare(gpu_ctx, mid, rays, "auto", om)
        assert (hits[1] != 0xFFFFFFFF).mean() > 0.3
    finally:
        gpu_ctx.mesh_release(mid)


def test_compressed_render_and_counts(gpu_ctx):
    sc = W.scaled(W.DEM4097, dem=129, image=64, spp=4, texture=64)
    built = W.build(gpu_ctx, sc)
    rp0 = S.RenderPasses(64, 64)
    gpu_ctx.render(built.camera, built.options, rp0)
    gpu_ctx.blas_build(built.mesh, compress=True)
    try:
        rp1 = S.RenderPasses(64, 64)
        gpu_ctx.render(built.camera, built.options, rp1)
        for name in ("total_radiance", "direct_radiance", "albedo", "depth", "triangle_id", "normal_global"):
            a, b = getattr(rp0, name).cpu().numpy(), getattr(rp1, name).cpu().numpy()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
        st = gpu_ctx.render(built.camera, replace(built.options, count_visits=True), S.RenderPasses(64, 64))
        assert st["closest_visits"]["nodes"] >= st["primary_rays"] + st["bounce_rays"]
    finally:
        gpu_ctx.blas_build(built.mesh)


def test_compressed_instanced_render(gpu_ctx):
    isc = W.build_instanced(gpu_ctx, dem=129, image=96, spp=2, bounces=1, sphere=(41, 64), body=(41, 64),
                            sun="point")
    rp0 = S.RenderPasses(96, 96, which=("radiance", "depth", "tri_id", "instance_id"))
    gpu_ctx.render(isc.camera, isc.options, rp0)
    for m in isc.meshes:
        gpu_ctx.blas_build(m, compress=True)
    gpu_ctx.tlas_build(isc.meshes, isc.transforms)
    rp1 = S.RenderPasses(96, 96, which=("radiance", "depth", "tri_id", "instance_id"))
    gpu_ctx.render(isc.camera, isc.options, rp1)
    for name in ("total_radiance", "depth", "triangle_id", "instance_id"):
        a, b = getattr(rp0, name).cpu().numpy(), getattr(rp1, name).cpu().numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    assert (rp1.triangle_id.cpu().numpy().view(np.uint32) != 0xFFFFFFFF).mean() > 0.2


def test_compressed_refit_and_refusal(gpu_ctx):
    h = gpu_ctx.synth_fbm(65, 65, 3)
    proj = S.projection(0, 30.0)
    mid = gpu_ctx.dem_mesh(h, proj)
    try:
        gpu_ctx.blas_build(mid, compress=True)
        h2 = (h * 0.7 - 20.0).contiguous()
        gpu_ctx.dem_update(mid, h2, refit=True)  # the quantised copy is re-encoded from the new boxes
        rays = _rays(30000, (-960.0, -960.0, -800.0), (960.0, 960.0, 0.0), 800.0, 6000.0, np.random.default_rng(5))
        hr = _trace(gpu_ctx, mid, rays)
        om = O.dem_mesh(h2.cpu().numpy(), O.make_projection(0, 30.0), None)
        _, oid, _, _, _ = om.trace(rays)
        assert np.array_equal(hr[1], oid) and (oid != 0xFFFFFFFF).mean() > 0.5
        with pytest.raises(ViraGPUError):
            gpu_ctx.blas_build(mid, width=2, compress=True)
    finally:
        gpu_ctx.mesh_release(mid)

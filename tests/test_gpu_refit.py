"""BLAS refit (vira_gpu_blas_refit / vira_gpu_dem_update_refit): the tree of
the last build kept, the leaf triangles and every box recomputed from the
moved vertices.  The reference rebuilds (embree_blas.ipp:14-71); a refit must
give the same closest hits as a fresh build of the same geometry, because the
closest hit (ties on t to the lowest triangle id) does not depend on the tree.

Bar: bit-exact hits (t, id, barycentrics) against a rebuilt BLAS; any-hit
occlusion just beyond each closest hit on every ray that has one (a box
that no longer contains its triangles still lets a closest-hit search find
them late, but not an occlusion query bounded at the hit); bit-exact renders, for small perturbations (the height noise
of a fit), a large deformation that moves the surface far outside the old
boxes, a 2-wide BLAS (rebuilt instead) and an instanced scene whose instance
box comes from the refitted root."""
import numpy as np
import pytest

from vira_amd import scene as S
from vira_amd import workloads as W

pytestmark = pytest.mark.gpu


def _rays(sc, n, seed, z=3000.0):
    rng = np.random.default_rng(seed)
    ext = 0.45 * (sc.dem - 1) * sc.spacing
    o = np.stack([rng.uniform(-ext, ext, n), rng.uniform(-ext, ext, n), np.full(n, z)], 1)
    d = np.stack([rng.uniform(-0.4, 0.4, n), rng.uniform(-0.4, 0.4, n), -np.ones(n)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, np.zeros((n, 1)), d, np.full((n, 1), np.inf)], 1).astype(np.float32)


def _trace(ctx, mesh, rays, any_hit=False):
    import torch
    t, ids, b1, b2, _ = ctx.trace_rays(mesh, torch.from_numpy(rays).cuda(), any_hit=any_hit)
    return [a.cpu().numpy().view(np.uint32) for a in (t, ids, b1, b2)]


def _same_hits(ctx, mesh, rays, ref):
    got = _trace(ctx, mesh, rays)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
    return got


def _occlusion(ctx, mesh, rays, closest_t):
    r = rays.copy()
    r[:, 7] = np.where(np.isfinite(closest_t), closest_t * 1.001, 1e30).astype(np.float32)
    _, ids, _, _ = _trace(ctx, mesh, r, any_hit=True)
    return ids != 0xFFFFFFFF


@pytest.mark.parametrize("width", [4, 2])
def test_refit_hits_equal_rebuild(gpu_ctx, width):
    import torch
    sc = W.scaled(W.DEM257, image=128)
    b = W.build(gpu_ctx, sc, bvh_width=width)
    rays = _rays(sc, 50000, 3)
    g = torch.Generator(device="cuda").manual_seed(5)
    for step, sigma in enumerate((2.0, 2.0, 40.0)):
        h = (b.heights + sigma * torch.randn(b.heights.shape, generator=g, device="cuda")).contiguous()
        gpu_ctx.dem_update(b.mesh, h, refit=True)
        refit = _trace(gpu_ctx, b.mesh, rays)
        occ = _occlusion(gpu_ctx, b.mesh, rays, refit[0].view(np.float32))
        st = gpu_ctx.blas_build(b.mesh, width=width)  # a fresh tree of the same geometry
        assert st["width"] == width
        _same_hits(gpu_ctx, b.mesh, rays, refit)
        fin = np.isfinite(refit[0].view(np.float32))
        # every ray with a closest hit is occluded just beyond it, in the refitted tree as in the rebuilt one
        assert np.array_equal(occ, fin)
        assert np.array_equal(_occlusion(gpu_ctx, b.mesh, rays, refit[0].view(np.float32)), fin)
        assert (refit[1] != 0xFFFFFFFF).mean() > 0.9, step


def test_refit_render_and_large_deformation(gpu_ctx):
    """A render after a refit equals the render after a rebuild, bit for bit,
    also after the surface is lifted by 800 m and tilted (every old box
    misses it)."""
    import torch
    sc = W.scaled(W.DEM257, image=128)
    b = W.build(gpu_ctx, sc)
    yy, xx = torch.meshgrid(torch.arange(sc.dem, device="cuda"), torch.arange(sc.dem, device="cuda"), indexing="ij")
    lifted = (b.heights + 800.0 + 2.0 * (xx - yy).float()).contiguous()
    opts = S.PathTracerOptions(samples=2, bounces=1, seed=11)
    for h in (lifted, b.heights):
        gpu_ctx.dem_update(b.mesh, h, refit=True)
        p1 = S.RenderPasses(128, 128)
        gpu_ctx.render(b.camera, opts, p1)
        gpu_ctx.blas_build(b.mesh)
        p2 = S.RenderPasses(128, 128)
        gpu_ctx.render(b.camera, opts, p2)
        torch.cuda.synchronize()
        for name in ("total_radiance", "depth", "triangle_id", "normal_global"):
            assert torch.equal(getattr(p1, name), getattr(p2, name)), name
        assert (p1.triangle_id.cpu().numpy().view(np.uint32) != 0xFFFFFFFF).mean() > 0.2


def test_refit_updates_instance_box(gpu_ctx):
    """The BLAS box the instanced traversal culls with comes from the
    refitted root: after the DEM is lifted 3 km, a second instance placed to
    see only the lifted surface still hits it."""
    import torch
    sc = W.scaled(W.DEM257, image=64)
    b = W.build(gpu_ctx, sc)
    h = (b.heights + 3000.0).contiguous()
    gpu_ctx.dem_update(b.mesh, h, refit=True)
    shift = np.eye(4)
    shift[0, 3] = 1.0e6  # a second copy far away, so the scene is instanced
    gpu_ctx.tlas_build([b.mesh, b.mesh], transforms=[np.eye(4), shift])
    rays = _rays(sc, 20000, 9, z=6000.0)
    inst = _trace(gpu_ctx, b.mesh, rays)
    gpu_ctx.blas_build(b.mesh)
    ref = _trace(gpu_ctx, b.mesh, rays)
    for g, r in zip(inst, ref):
        assert np.array_equal(g, r)
    assert (ref[1] != 0xFFFFFFFF).mean() > 0.9
    # the render of the instanced scene sees the lifted surface in both cases
    opts = S.PathTracerOptions(samples=1, bounces=0, seed=2)
    gpu_ctx.dem_update(b.mesh, b.heights + 3000.0, refit=True)
    p1 = S.RenderPasses(64, 64)
    gpu_ctx.render(b.camera, opts, p1)
    gpu_ctx.blas_build(b.mesh)
    p2 = S.RenderPasses(64, 64)
    gpu_ctx.render(b.camera, opts, p2)
    torch.cuda.synchronize()
    assert torch.equal(p1.total_radiance, p2.total_radiance)
    assert torch.equal(p1.depth, p2.depth)


def test_refit_requires_built_blas(gpu_ctx):
    import torch
    sc = W.scaled(W.DEM257, image=64)
    h = gpu_ctx.synth_fbm(sc.dem, sc.dem, sc.seed, sc.amplitude, sc.octaves, sc.base_cells)
    mesh = gpu_ctx.dem_mesh(h, W.make_projection(sc), None)
    with pytest.raises(Exception) as e:
        gpu_ctx.blas_refit(mesh)
    assert "built" in str(e.value)
    # a DEM update on a mesh without a BLAS only moves the vertices
    gpu_ctx.dem_update(mesh, (h + 1.0).contiguous(), refit=True)
    torch.cuda.synchronize()


@pytest.mark.parametrize("side", [2, 3, 17])
def test_refit_tiny_dems(gpu_ctx, side):
    """A 2x2 DEM (two triangles, one leaf under the root), 3x3 and 17x17: the
    refit hits equal a rebuild's after the heights change."""
    import torch
    proj = S.projection(0, 30.0)
    h = torch.linspace(0.0, 50.0, side * side, device="cuda").reshape(side, side).contiguous()
    mesh = gpu_ctx.dem_mesh(h, proj, None)
    gpu_ctx.blas_build(mesh)
    ext = 0.5 * (side - 1) * 30.0
    rng = np.random.default_rng(side)
    n = 4000
    o = np.stack([rng.uniform(-ext, ext, n), rng.uniform(-ext, ext, n), np.full(n, 500.0)], 1)
    d = np.stack([rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n), -np.ones(n)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.zeros((n, 1)), d, np.full((n, 1), np.inf)], 1).astype(np.float32)
    h2 = (h * -1.5 + 120.0).contiguous()
    gpu_ctx.dem_update(mesh, h2, refit=True)
    refit = _trace(gpu_ctx, mesh, rays)
    gpu_ctx.blas_build(mesh)
    _same_hits(gpu_ctx, mesh, rays, refit)
    assert (refit[1] != 0xFFFFFFFF).any()

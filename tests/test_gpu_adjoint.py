"""GPU differentiable direct illumination (the 1025^2 DEM scene and scaled copies) against
the FP64 oracle model (oracle/oracle_diff.hpp), through the C-ABI.

Bar:
* forward image bit-exact with the path tracer at bounces = 0 and with the
  oracle's FP32 twin (same hit records);
* dL/dheights and dL/dtexels against the FP64 oracle adjoint on the same
  records (FP32 arithmetic and float atomics in run-to-run order on the GPU),
  relative L2 error: small scenes <= 1e-5 (heights) and 2e-5 (texels), measured
  5e-7 and 2.6e-6; the 1025^2 scene <= 5e-5 and 2.5e-4, measured 1.0e-5 and
  5.8e-5 (a texel sums the contributions of ~16 pixels' samples);
* dL/dheights against central finite differences of the FP64 L2 loss
  (h = 1e-3 m) on random nodes: |g - fd| <= 2e-4 max|fd| + 1e-3 |fd| (measured
  worst node 5e-5 max|fd| on the 1025^2 scene, 2e-6 on the small ones);
* hit records kept by the forward pass and re-traced ones give the same
  gradients (up to the order of the float atomics).
"""
from dataclasses import replace

import numpy as np
import pytest

import pyoracle as O
from vira_amd import scene as S
from vira_amd import workloads as W

pytestmark = pytest.mark.gpu


def _oracle(built):
    sc = built.scene
    heights = built.heights.cpu().numpy()
    proj = W.make_projection(sc)
    oproj = O.make_projection(proj.kind, proj.spacing, proj.radius, tuple(proj.center), tuple(proj.offset))
    om = O.dem_mesh(heights, oproj, None)
    cam = O.camera_from(built.camera.to_c())
    lights = [O.light_from(L.to_c()) for L in built.lights]
    tex = built.texture.cpu().numpy()
    mat = O.material(1.0, tex, sc.smooth_shading)
    dd = O.DemDiff(heights, oproj, None)
    return om, cam, lights, mat, dd, heights, tex


def _small(smooth=False, spp=1):
    sc = W.scaled(W.DEM1025_DI, dem=65, image=96, texture=64)
    return replace(sc, smooth_shading=smooth, spp=spp)


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_di_forward_equals_render_and_oracle(gpu_ctx):
    import torch
    sc = _small(spp=3)
    b = W.build(gpu_ctx, sc)
    opts = S.DIOptions(samples=3, seed=11)
    img = gpu_ctx.di_render(b.camera, opts).cpu().numpy()
    passes = S.RenderPasses(sc.image, sc.image)
    gpu_ctx.render(b.camera, S.PathTracerOptions(samples=3, bounces=0, seed=11, integrator="fused"), passes)
    torch.cuda.synchronize()
    ref = passes.total_radiance.cpu().numpy()
    assert np.array_equal(img.view(np.uint32), ref.reshape(img.shape).view(np.uint32))
    om, cam, lights, mat, dd, heights, tex = _oracle(b)
    out = om.render(cam, lights, mat, O.options(3, 0, False, 1, 11), precision=32, records=True)
    assert np.array_equal(img.reshape(-1, 8).view(np.uint32), out["radiance"].view(np.uint32))
    # FP64 model on the FP32 twin's records: only the shadow-ray offset and FP32 rounding differ
    img64 = dd.forward(heights, tex, out["records"], lights, mat)
    assert np.abs(img64 - img.reshape(-1, 8)).max() <= 1e-5 * np.abs(img64).max()


@pytest.mark.parametrize("smooth", [False, True])
def test_adjoint_matches_fp64_oracle(gpu_ctx, smooth):
    import torch
    sc = _small(smooth=smooth, spp=2)
    b = W.build(gpu_ctx, sc)
    opts = S.DIOptions(samples=2, seed=5)
    img = gpu_ctx.di_render(b.camera, opts)
    rng = np.random.default_rng(7)
    g = rng.standard_normal(img.shape).astype(np.float32)
    gh = torch.empty_like(b.heights)
    gt = torch.empty_like(b.texture)
    st = gpu_ctx.adjoint(b.camera, opts, torch.from_numpy(g).cuda(), gh, gt)
    assert st["used_records"] == 1
    om, cam, lights, mat, dd, heights, tex = _oracle(b)
    out = om.render(cam, lights, mat, O.options(2, 0, False, 1, 5), precision=32, records=True)
    assert np.array_equal(img.cpu().numpy().reshape(-1, 8).view(np.uint32), out["radiance"].view(np.uint32))
    gh64, gt64 = dd.adjoint(heights, tex, out["records"], lights, mat, g.reshape(-1, 8).astype(np.float64))
    eh = _rel_l2(gh.cpu().numpy().reshape(-1).astype(np.float64), gh64)
    et = _rel_l2(gt.cpu().numpy().reshape(-1).astype(np.float64), gt64)
    print(f"smooth={smooth}: rel L2 error heights {eh:.2e}, texels {et:.2e}")
    assert np.count_nonzero(gh64) > 100 and np.count_nonzero(gt64) > 100
    assert eh <= 1e-5 and et <= 2e-5


def _fd_check(ctx, b, opts, nodes_n, seed):
    """GPU dL/dh of the L2 loss against the target vs FP64 central differences."""
    import torch
    th, tt = W.perturbed_inputs(ctx, b.scene, b.heights, b.texture)
    # target image: the perturbed DEM rendered with its own visibility
    ctx.dem_update(b.mesh, th)
    ctx.material_set(b.mesh, texture=tt, smooth_shading=b.scene.smooth_shading)
    target = ctx.di_render(b.camera, opts).clone()
    ctx.dem_update(b.mesh, b.heights)
    ctx.material_set(b.mesh, texture=b.texture, smooth_shading=b.scene.smooth_shading)
    img = ctx.di_render(b.camera, opts)
    grad = 2.0 * (img - target)
    gh = torch.empty_like(b.heights)
    ctx.adjoint(b.camera, opts, grad, gh, None)
    om, cam, lights, mat, dd, heights, tex = _oracle(b)
    out = om.render(cam, lights, mat, O.options(opts.samples, 0, False, 1, opts.seed), precision=32, records=True)
    rng = np.random.default_rng(seed)
    nodes = rng.choice(heights.size, nodes_n, replace=False)
    fd = dd.fd_heights(heights, tex, out["records"], lights, mat,
                       target.cpu().numpy().reshape(-1, 8).astype(np.float64), nodes, 1e-3)
    g = gh.cpu().numpy().reshape(-1)[nodes].astype(np.float64)
    err = np.abs(g - fd)
    tol = 2e-4 * np.abs(fd).max() + 1e-3 * np.abs(fd)
    print(f"FD check on {nodes_n} nodes: max |g - fd| = {err.max():.3e}, max |fd| = {np.abs(fd).max():.3e}, "
          f"rel L2 {_rel_l2(g, fd):.2e}, nonzero {np.count_nonzero(fd)}")
    assert np.count_nonzero(fd) > nodes_n // 2
    assert np.all(err <= tol), (err / np.maximum(tol, 1e-300)).max()


@pytest.mark.parametrize("smooth", [False, True])
def test_adjoint_central_differences_small(gpu_ctx, smooth):
    b = W.build(gpu_ctx, _small(smooth=smooth))
    _fd_check(gpu_ctx, b, S.DIOptions(samples=1, seed=0), 400, 3)


def test_adjoint_records_kept_or_retraced(gpu_ctx):
    import torch
    b = W.build(gpu_ctx, _small(spp=4))
    rng = np.random.default_rng(2)
    g = torch.from_numpy(rng.standard_normal((96, 96, 8)).astype(np.float32)).cuda()
    res = []
    for budget in (1 << 30, 0):
        opts = S.DIOptions(samples=4, seed=9, record_budget=budget)
        gpu_ctx.di_render(b.camera, opts)
        gh, gt = torch.empty_like(b.heights), torch.empty_like(b.texture)
        st = gpu_ctx.adjoint(b.camera, opts, g, gh, gt)
        assert st["used_records"] == (1 if budget else 0)
        res.append((gh.cpu().numpy(), gt.cpu().numpy()))
    for a, c in zip(res[0], res[1]):
        assert _rel_l2(a.astype(np.float64), c.astype(np.float64)) <= 1e-5
    # a changed scene invalidates kept records: the adjoint re-traces
    opts = S.DIOptions(samples=4, seed=9)
    gpu_ctx.di_render(b.camera, opts)
    gpu_ctx.lights_set(b.lights)
    gh = torch.empty_like(b.heights)
    assert gpu_ctx.adjoint(b.camera, opts, g, gh, None)["used_records"] == 0


def test_dem_update_bit_exact_and_mask_rule(gpu_ctx):
    import torch
    sc = _small()
    b = W.build(gpu_ctx, sc)
    h2 = (b.heights + 3.5 * torch.sin(torch.arange(b.heights.numel(), device="cuda").reshape(b.heights.shape)
                                       .float())).contiguous()
    gpu_ctx.dem_update(b.mesh, h2)
    pos, idx, nrm = gpu_ctx.mesh_download(b.mesh)
    proj = W.make_projection(sc)
    oproj = O.make_projection(proj.kind, proj.spacing, proj.radius, tuple(proj.center), tuple(proj.offset))
    hn = h2.cpu().numpy()
    opos, opos64 = O.dem_positions(hn, oproj)
    onrm = O.vertex_normals(opos, O.dem_indices(hn), opos64)
    assert np.array_equal(pos.view(np.uint32), opos.view(np.uint32))
    assert np.array_equal(nrm.view(np.uint32), onrm.view(np.uint32))
    # the BLAS was rebuilt: closest hits equal the oracle's on the new geometry
    om = O.dem_mesh(hn, oproj, None)
    rng = np.random.default_rng(4)
    n = 4096
    ext = 0.45 * (sc.dem - 1) * sc.spacing
    o = np.stack([rng.uniform(-ext, ext, n), rng.uniform(-ext, ext, n), np.full(n, 2000.0)], 1)
    d = np.stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-0.3, 0.3, n), -np.ones(n)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.zeros((n, 1)), d, np.full((n, 1), np.inf)], 1).astype(np.float32)
    t, ids, _, _, _ = gpu_ctx.trace_rays(b.mesh, torch.from_numpy(rays).cuda())
    ot, oids, _, _, _ = om.trace(rays[:, [0, 1, 2, 4, 5, 6]])
    assert np.array_equal(ids.cpu().numpy().view(np.uint32), oids)
    # holes may not open or close
    h3 = h2.clone()
    h3[3, 3] = float("nan")
    with pytest.raises(Exception) as e:
        gpu_ctx.dem_update(b.mesh, h3)
    assert "finite" in str(e.value)


def test_adjoint_rejects_sphere_lights(gpu_ctx):
    import torch
    b = W.build(gpu_ctx, _small())
    gpu_ctx.lights_set([S.new_sun(b.lights[0].position)])
    opts = S.DIOptions()
    gpu_ctx.di_render(b.camera, opts)
    with pytest.raises(Exception) as e:
        gpu_ctx.adjoint(b.camera, opts, torch.zeros((96, 96, 8), device="cuda"), torch.empty_like(b.heights))
    assert getattr(e.value, "status", None) == 5 or "point lights" in str(e.value)


def test_autograd_function_and_descent(gpu_ctx):
    import torch
    d = W.build_diff(gpu_ctx, _small())
    h = d.built.heights.clone().requires_grad_(True)
    t = d.built.texture.clone().requires_grad_(True)
    img = d.dem(h, t)
    loss = ((img - d.target) ** 2).sum()
    loss.backward()
    # same numbers as a direct adjoint call with dL/dimage = 2 (img - target)
    gh, gt = torch.empty_like(h), torch.empty_like(t)
    d.dem._apply(h, t)
    gpu_ctx.di_render(d.dem.camera, d.dem.options)
    gpu_ctx.adjoint(d.dem.camera, d.dem.options, 2.0 * (img.detach() - d.target), gh, gt)
    assert _rel_l2(h.grad.cpu().numpy().astype(np.float64), gh.cpu().numpy().astype(np.float64)) <= 1e-5
    assert _rel_l2(t.grad.cpu().numpy().astype(np.float64), gt.cpu().numpy().astype(np.float64)) <= 1e-5
    # a few gradient steps on the texels lower the loss
    losses = []
    tt = d.built.texture.clone().requires_grad_(True)
    opt = torch.optim.Adam([tt], lr=0.005)
    for _ in range(5):
        opt.zero_grad()
        L = ((d.dem(d.built.heights, tt) - d.target) ** 2).sum()
        L.backward()
        opt.step()
        losses.append(float(L.detach()))
    assert losses[-1] < 0.8 * losses[0], losses


def test_dem1025_full_size_adjoint(gpu_ctx):
    """Full size: 1025^2 DEM (2,097,152 triangles), 1024^2 image."""
    import torch
    sc = W.DEM1025_DI
    b = W.build(gpu_ctx, sc)
    opts = S.DIOptions(samples=1, seed=0)
    img = gpu_ctx.di_render(b.camera, opts)
    fwd = gpu_ctx.last_stats
    rng = np.random.default_rng(8)
    g = rng.standard_normal(img.shape).astype(np.float32)
    gh, gt = torch.empty_like(b.heights), torch.empty_like(b.texture)
    st = gpu_ctx.adjoint(b.camera, opts, torch.from_numpy(g).cuda(), gh, gt)
    print(f"1025^2 DI: forward {fwd['ms']:.2f} ms, adjoint {st['ms']:.2f} ms (records kept: {st['used_records']})")
    om, cam, lights, mat, dd, heights, tex = _oracle(b)
    out = om.render(cam, lights, mat, O.options(1, 0, False, 1, 0), precision=32, records=True)
    assert np.array_equal(img.cpu().numpy().reshape(-1, 8).view(np.uint32), out["radiance"].view(np.uint32))
    gh64, gt64 = dd.adjoint(heights, tex, out["records"], lights, mat, g.reshape(-1, 8).astype(np.float64))
    eh = _rel_l2(gh.cpu().numpy().reshape(-1).astype(np.float64), gh64)
    et = _rel_l2(gt.cpu().numpy().reshape(-1).astype(np.float64), gt64)
    print(f"1025^2 DI: rel L2 error heights {eh:.2e}, texels {et:.2e}")
    assert eh <= 5e-5 and et <= 2.5e-4
    _fd_check(gpu_ctx, b, opts, 1000, 12)

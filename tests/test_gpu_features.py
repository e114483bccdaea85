"""GPU parity of the integrator options, lights and material features against
the oracle's FP32 twin, through the C-ABI, bit for bit: several lights (the
per-pixel renderer's MULTI variant, the wavefront schedule and a batched
request), ambient light, background on misses, normal maps, spectral albedo
maps, adaptive sampling, and the normal_camera / mesh_id / material_id passes.
The oracle's restatements are pinned in test_oracle_features.py."""
import numpy as np
import pytest

import pyoracle as O
from vira_amd import ViraGPUError
from vira_amd import scene as S
from vira_amd import workloads as W

pytestmark = pytest.mark.gpu

SPECTRAL = ("total_radiance", "radiance"), ("direct_radiance", "direct"), ("indirect_radiance", "indirect"), \
           ("albedo", "albedo")
OTHER = ("normal_global", "normal"), ("depth", "depth"), ("alpha", "alpha"), ("triangle_id", "tri_id"), \
        ("normal_camera", "normal_camera"), ("mesh_id", "mesh_id"), ("material_id", "material_id")


def _oracle_mesh(built, mesh_id):
    sc = built.scene
    p = W.make_projection(sc)
    proj = O.make_projection(p.kind, p.spacing, p.radius, tuple(p.center), tuple(p.offset))
    valb = None if built.vertex_albedo is None else built.vertex_albedo.cpu().numpy()
    h = built.heights.cpu().numpy()
    pos, pos64 = O.dem_positions(h, proj)
    idx = O.dem_indices(h)
    nrm = O.vertex_normals(pos, idx, pos64)
    a = np.ones(h.size, np.float32) if valb is None else valb
    return O.Mesh(pos, idx, nrm, np.repeat(a[:, None], 8, 1), O.dem_uv(h.shape[1], h.shape[0]), pos64,
                  mesh_id=mesh_id)


def _check(rp, o, names=SPECTRAL + OTHER):
    for attr, key in names:
        t = getattr(rp, attr)
        if t is None:
            continue
        g = t.cpu().numpy().reshape(o[key].shape)
        assert np.array_equal(g.view(np.uint32), o[key].view(np.uint32)), attr


def _opts(sc, **kw):
    return O.options(sc.spp, sc.bounces, sc.russian_roulette, 1, sc.render_seed, **kw)


@pytest.mark.parametrize("integrator", ["fused", "wavefront"])
def test_several_lights_bit_exact(gpu_ctx, integrator):
    """Two point lights and a sphere sun, 2 bounces: one shadow ray per light per
    vertex, terms summed in light order (cpu_path_tracer.ipp:399-428)."""
    sc = W.scaled(W.DEM4097, dem=97, image=64, spp=3, texture=64)
    built = W.build(gpu_ctx, sc)
    lights = built.lights + [S.directional_sun(55.0, 140.0, [0.5, 1, 1.5, 2, 2.5, 3, 3.5, 4], distance=1e6),
                             S.PointLight((200.0, -300.0, 2500.0), np.full(8, 3e6, np.float32))]
    gpu_ctx.lights_set(lights)
    try:
        rp = S.RenderPasses(sc.image, sc.image)
        built.options.integrator = integrator
        st = gpu_ctx.render(built.camera, built.options, rp)
        assert st["integrator"] == (1 if integrator == "fused" else 2)
        om = _oracle_mesh(built, built.mesh)
        tex = built.texture.cpu().numpy()
        o = om.render(O.camera_from(built.camera.to_c()), [O.light_from(L.to_c()) for L in lights],
                      O.material(1.0, tex), _opts(sc))
        assert (o["radiance"] > 0).mean() > 0.3
        names = SPECTRAL + OTHER if integrator == "fused" else SPECTRAL + OTHER[:4]
        _check(rp, o, names)
        assert st["shadow_rays"] == o["counts"]["shadow"] and st["bounce_rays"] == o["counts"]["bounce"]
    finally:
        gpu_ctx.lights_set(built.lights)


def test_several_lights_batched_request(gpu_ctx):
    """A batch mixing a two-light request and a one-light request: each request's
    passes equal the oracle's (the batch then runs the several-light variant for both)."""
    sc = W.scaled(W.DEM257, dem=65, image=48, spp=2, bounces=1)
    built = W.build(gpu_ctx, sc)
    two = built.lights + [S.directional_sun(50.0, 200.0, 1.5, distance=1e6)]
    reqs = [S.RenderRequest(built.mesh, built.camera, built.options, two, S.RenderPasses(48, 48)),
            S.RenderRequest(built.mesh, built.camera, built.options, built.lights[0], S.RenderPasses(48, 48))]
    gpu_ctx.batch_submit(reqs)
    om = _oracle_mesh(built, built.mesh)
    for req, ls in zip(reqs, (two, built.lights)):
        o = om.render(O.camera_from(built.camera.to_c()), [O.light_from(L.to_c()) for L in ls], O.material(),
                      _opts(sc))
        _check(req.passes, o)


def test_ambient_background_bit_exact(gpu_ctx):
    """Ambient light at the first hit and a spectral background on misses (a wide
    field of view, so part of the image is sky), sphere sun, 2 bounces + RR."""
    sc = W.scaled(W.DEM4097, dem=97, image=64, spp=2, texture=64)
    built = W.build(gpu_ctx, sc)
    import torch
    H, Wd = 16, 32
    jj, ii = np.mgrid[0:H, 0:Wd]
    bg = np.stack([np.sin(ii / Wd * 6.283 + k) * 0.3 + 1.0 + 0.2 * np.cos(jj / H * 3.14 * (k + 1) / 3)
                   for k in range(8)], -1).astype(np.float32) * 1e-3
    amb = np.linspace(1e-3, 3e-3, 8).astype(np.float32)
    gpu_ctx.environment_set(ambient=amb, background=torch.from_numpy(bg).cuda())
    try:
        cam = built.camera
        # a 75 deg field of view sees past the tile's edges: about half the pixels are sky
        tilted = S.Camera(cam.width, cam.height, 75.0, cam.position, cam.rotation)
        opt = built.options
        opt.show_background = True
        rp = S.RenderPasses(64, 64)
        gpu_ctx.render(tilted, opt, rp)
        om = _oracle_mesh(built, built.mesh)
        o = om.render(O.camera_from(tilted.to_c()), [O.light_from(L.to_c()) for L in built.lights],
                      O.material(1.0, built.texture.cpu().numpy()),
                      _opts(sc, show_background=True, ambient=amb, background=bg))
        miss = o["tri_id"] == 0xFFFFFFFF
        assert 0.1 < miss.mean() < 0.9
        _check(rp, o)
        with pytest.raises(ViraGPUError) as e:
            opt.integrator = "wavefront"
            gpu_ctx.render(tilted, opt, S.RenderPasses(64, 64))
        assert e.value.status == 5
    finally:
        built.options.show_background = False
        built.options.integrator = "auto"
        gpu_ctx.environment_set(None, None)


def test_normal_map_spectral_albedo_bit_exact(gpu_ctx):
    """A spectral albedo map (8 bands per texel) and a tangent-space normal map on
    the DEM, sphere sun, bounces: passes bit-exact, the normal pass holding the
    mapped normal."""
    import torch
    sc = W.scaled(W.DEM4097, dem=97, image=64, spp=2)
    built = W.build(gpu_ctx, sc)
    rng = np.random.default_rng(5)
    tex8 = rng.uniform(0.05, 0.4, (40, 36, 8)).astype(np.float32)
    nm = rng.normal(0, 0.3, (24, 28, 3)).astype(np.float32)
    nm[..., 2] = 1.0
    t_tex, t_nm = torch.from_numpy(tex8).cuda(), torch.from_numpy(nm).cuda()
    gpu_ctx.material_set(built.mesh, albedo=1.0, texture=t_tex, normal_map=t_nm, material_id=41)
    rp = S.RenderPasses(64, 64)
    gpu_ctx.render(built.camera, built.options, rp)
    om = _oracle_mesh(built, built.mesh)
    o = om.render(O.camera_from(built.camera.to_c()), [O.light_from(L.to_c()) for L in built.lights],
                  O.material(1.0, tex8, normal_map=nm, material_id=41), _opts(sc))
    _check(rp, o)
    hit = o["tri_id"] != 0xFFFFFFFF
    assert hit.mean() > 0.5 and np.all(o["material_id"][hit] == 41)


def test_adaptive_sampling_bit_exact(gpu_ctx):
    """Adaptive sampling per pixel (cpu_path_tracer.ipp:198-227) with its stop
    divisor; pixels stop at different samples, including never."""
    sc = W.scaled(W.DEM257, dem=65, image=64, spp=70, bounces=1)
    built = W.build(gpu_ctx, sc)
    opt = built.options
    opt.adaptive_sampling = True
    opt.samples_per_batch = 10
    opt.samples_to_detect_miss = 10
    opt.sampling_tolerance = 0.02
    rp = S.RenderPasses(64, 64)
    st = gpu_ctx.render(built.camera, opt, rp)
    om = _oracle_mesh(built, built.mesh)
    o = om.render(O.camera_from(built.camera.to_c()), [O.light_from(L.to_c()) for L in built.lights], O.material(),
                  _opts(sc, adaptive=True, samples_per_batch=10, samples_to_detect_miss=10, sampling_tolerance=0.02))
    _check(rp, o)
    a = o["alpha"][o["tri_id"] != 0xFFFFFFFF]
    assert len(np.unique(np.round(a, 5))) > 2  # several stopping samples
    assert st["primary_rays"] == o["counts"]["primary"] < 70 * 64 * 64
    with pytest.raises(ViraGPUError) as e:
        opt.integrator = "wavefront"
        gpu_ctx.render(built.camera, opt, S.RenderPasses(64, 64))
    assert e.value.status == 5


def test_passes_view_normal_given(gpu_ctx):
    """normal_camera with a caller-given view normal matrix."""
    sc = W.scaled(W.DEM257, dem=65, image=48)
    built = W.build(gpu_ctx, sc)
    cam = built.camera
    vn = np.diag([1.0, -1.0, -1.0]) @ np.asarray(cam.rotation).T
    cam2 = S.Camera(cam.width, cam.height, cam.fov_deg, cam.position, cam.rotation, view_normal=vn)
    rp = S.RenderPasses(48, 48)
    gpu_ctx.render(cam2, built.options, rp)
    om = _oracle_mesh(built, built.mesh)
    o = om.render(O.camera_from(cam2.to_c()), [O.light_from(L.to_c()) for L in built.lights], O.material(),
                  _opts(sc))
    _check(rp, o)

"""quipu pyramids read into GPU meshes (vira_gpu_qld_read): the decoded
heights equal the reference's fixedToFloat of the stored samples
(image_utils.ipp:508-527) bit for bit, compressed or not (the uncompressed
layout is stored i-outer, quipu_io.ipp:760-768), for 8-bit, 16-bit and float
heights; albedo of every stored kind reaches the vertices; a planar level's
mesh equals vira_gpu_dem_mesh of the decoded heights with the same grid
spacing (positions, triangles, normals bit-exact); an OGR equirectangular
level's vertices are the eqc inverse + ellipsoidToCartesian within one FP32
ulp (host and device libm); the mesh feeds the BVH builder and renders."""
import math

import numpy as np
import pytest

from vira_amd import _capi
from vira_amd import quipu as Q
from vira_amd import scene as S

pytestmark = pytest.mark.gpu


def _pyramid(ctx, base=129, levels=3, seed=5, amplitude=300.0, spacing=20.0, holes=True):
    out = []
    h0 = ctx.synth_fbm(base, base, seed, amplitude).cpu().numpy()
    if holes:
        h0[10:14, 40:47] = np.inf
    for k in range(levels):
        h = np.ascontiguousarray(h0[::1 << k, ::1 << k])
        W = h.shape[1]
        p = Q.DEMProjection(resolution=(W, h.shape[0]), pixel_scale=(spacing, spacing), dx=float(1 << k),
                            dy=float(1 << k))
        out.append(Q.DEMLevel(h, p))
    return out


def _fixed_to_float(q, mm):
    """fixedToFloat in FP32, the reference's operation order."""
    m = np.float32(np.iinfo(q.dtype).max)
    f = q.astype(np.float32) / (m - np.float32(1))
    v = (np.float32(mm[1]) - np.float32(mm[0])) * f + np.float32(mm[0])
    return np.where(q == 0, np.float32(np.inf), v).astype(np.float32)


def _expected_heights(level, base_heights, bits):
    mm = Q._valid_minmax(base_heights)
    if bits == 32:
        return np.asarray(level.heights, np.float32)
    q = Q.float_to_fixed(level.heights, mm, np.uint8 if bits == 8 else np.uint16)
    return _fixed_to_float(q, mm)


@pytest.mark.parametrize("compress", [True, False])
@pytest.mark.parametrize("spacing,bits", [(20.0, 16), (4000.0, 8), (0.001, 32)])
def test_heights_decode_bit_exact(gpu_ctx, tmp_path, compress, spacing, bits):
    lv = _pyramid(gpu_ctx, spacing=spacing)
    path = tmp_path / "dem.qld"
    Q.write_pyramid(str(path), lv, compress=compress)
    d = Q.QuipuDEM(str(path))
    mm = Q._valid_minmax(lv[0].heights)
    for k in range(len(lv)):
        _, h, _, info = d.read_level(gpu_ctx, k, mesh=False, samples=True)
        # writeDEMHeights' precision rule (quipu_io.ipp:641-658) for this level's GSD
        scale = np.float32(np.float32(min(lv[k].projection.gsd())) / np.float32(50))
        delta = np.float32(np.float32(mm[1]) - np.float32(mm[0]))
        rule = 8 if delta / np.float32(255) < scale else (16 if delta / np.float32(65535) < scale else 32)
        assert info["height_bits"] == rule and info["compressed"] == int(compress)
        if k == 0:
            assert rule == bits
        want = _expected_heights(lv[k], lv[0].heights, rule)
        got = h.cpu().numpy()
        assert got.shape == want.shape
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), k
    d.close()


def test_planar_mesh_equals_dem_mesh(gpu_ctx, tmp_path):
    import torch
    lv = _pyramid(gpu_ctx, spacing=20.0)
    path = tmp_path / "dem.qld"
    Q.write_pyramid(str(path), lv, compress=True)
    d = Q.QuipuDEM(str(path))
    mid, h, _, info = d.read_level(gpu_ctx, 0, mesh=True, samples=True)
    pos, idx, nrm = gpu_ctx.mesh_download(mid)
    proj = _capi.Projection()
    proj.kind, proj.spacing = 0, 20.0
    ref = gpu_ctx.dem_mesh(h, proj)
    pos2, idx2, nrm2 = gpu_ctx.mesh_download(ref)
    assert np.array_equal(idx, idx2)
    fin = np.isfinite(pos2).all(1)
    assert np.array_equal(pos[fin], pos2[fin]) and np.array_equal(nrm, nrm2)
    assert len(idx) < 2 * 128 * 128  # the hole removed triangles
    # it feeds the BVH builder and renders
    gpu_ctx.blas_build(mid)
    gpu_ctx.tlas_build([mid])
    cam = S.Camera(128, 128, 40.0, (0.0, 0.0, 3000.0))
    cam.look_in_direction((0.0, 0.0, -1.0))
    gpu_ctx.lights_set([S.PointLight((1e7, 0.0, 1e7), np.full(8, 1e14, np.float32))])
    passes = S.RenderPasses(128, 128, which=("radiance", "tri_id"))
    gpu_ctx.render(cam, S.PathTracerOptions(samples=1), passes)
    torch.cuda.synchronize()
    assert (passes.triangle_id >= 0).float().mean().item() > 0.9
    print(f"level 0 {info['width']}x{info['height']}: read {info['read_ms']:.2f} ms "
          f"(LZ4 {info['decode_ms']:.2f} ms), device {info['build_ms']:.2f} ms")
    d.close()


@pytest.mark.parametrize("kind", ["none", "constant", "grey", "color", "skip"])
def test_albedo_kinds(gpu_ctx, tmp_path, kind):
    lv = _pyramid(gpu_ctx, base=33, levels=2, holes=False)
    H, W = lv[0].heights.shape
    rng = np.random.default_rng(9)
    if kind == "constant":
        alb = [np.linspace(0.1, 0.3, 8).astype(np.float32)] * 2
    elif kind == "grey":
        alb = [(0.05 + 0.2 * rng.random(l.heights.shape)).astype(np.float32) for l in lv]
    elif kind == "color":
        alb = [(0.05 + 0.2 * rng.random(l.heights.shape + (8,))).astype(np.float32) for l in lv]
    else:
        alb = [None, None]
    for l, a in zip(lv, alb):
        l.albedo = a
    path = tmp_path / "alb.qld"
    Q.write_pyramid(str(path), lv, compress=True, write_albedos=(kind != "none"))
    opts = Q.QuipuReaderOptions(read_albedos=(kind != "skip"), albedo_profile=np.linspace(0.5, 1.5, 8))
    d = Q.QuipuDEM(str(path), opts)
    _, _, a, info = d.read_level(gpu_ctx, 0, mesh=False, samples=True)
    got = a.cpu().numpy()
    prof = np.linspace(0.5, 1.5, 8).astype(np.float32)
    if kind in ("none", "skip"):
        want = np.full((H, W, 8), np.float32(0.06))
        assert info["albedo_type"] == -1
    elif kind == "constant":
        want = np.broadcast_to(alb[0], (H, W, 8))
        assert info["albedo_type"] == 0
    elif kind == "grey":
        mm = Q._valid_minmax(alb[0])
        g = _fixed_to_float(Q.float_to_fixed(alb[0], mm, np.uint8), mm)
        want = prof[None, None, :] * g[..., None]
        assert info["albedo_type"] == 1
    else:
        chans = []
        for k in range(8):
            mm = Q._valid_minmax(alb[0][:, :, k])
            chans.append(_fixed_to_float(Q.float_to_fixed(alb[0][:, :, k], mm, np.uint8), mm))
        want = np.stack(chans, 2)
        assert info["albedo_type"] == 2
    assert np.array_equal(got, want.astype(np.float32))
    d.close()


def test_equirectangular_level(gpu_ctx, tmp_path):
    R = 1737400.0
    h = gpu_ctx.synth_fbm(65, 65, 3, 2000.0).cpu().numpy()
    deg = 0.01  # pixel size in degrees of longitude at the equator
    ps = R * math.radians(deg)
    p = Q.DEMProjection(resolution=(65, 65), pixel_scale=(ps, -ps), is_ogr=True,
                        proj_ref=Q.equirectangular_wkt(R, central_meridian=30.0),
                        model_tie_point=(-32.5 * ps, 32.5 * ps))
    lv = [Q.DEMLevel(h, p)]
    path = tmp_path / "eqc.qld"
    Q.write_pyramid(str(path), lv, compress=True)
    d = Q.QuipuDEM(str(path))
    mid, hd, _, info = d.read_level(gpu_ctx, 0, samples=True)
    assert info["is_ogr"] == 1
    pos, _, _ = gpu_ctx.mesh_download(mid)
    want = Q.level_positions(Q.DEMLevel(hd.cpu().numpy(), p)).reshape(-1, 3).astype(np.float32)
    ulp = np.spacing(np.abs(want))
    assert np.all(np.abs(pos - want) <= ulp), np.abs(pos - want).max()
    # the centre sample sits on the central meridian, on the equator
    c = pos.reshape(65, 65, 3)[32, 32].astype(np.float64)
    lon = math.degrees(math.atan2(c[1], c[0]))
    assert abs(lon - 30.0) < 1e-4 and abs(c[2]) < 1.0
    # other map projections are refused with their name
    p2 = Q.DEMProjection(resolution=(65, 65), pixel_scale=(ps, -ps), is_ogr=True,
                         proj_ref=Q.equirectangular_wkt(R).replace("Equirectangular\"]", "Lambert_Conformal_Conic_2SP\"]"))
    Q.write_pyramid(str(tmp_path / "lcc.qld"), [Q.DEMLevel(h, p2)], compress=False)
    d2 = Q.QuipuDEM(str(tmp_path / "lcc.qld"))
    with pytest.raises(_capi.ViraGPUError, match="lambert_conformal_conic_2sp"):
        d2.read_level(gpu_ctx, 0)
    d.close()
    d2.close()


@pytest.mark.parametrize("south,lat,lon0", [(False, None, 0.0), (True, None, 60.0), (False, 75.0, -20.0)])
def test_polar_stereographic_level(gpu_ctx, tmp_path, south, lat, lon0):
    """A 65^2 level around the pole (20 km pixels), read through the C-ABI:
    positions within one FP32 ulp of the FP64 inverse projection."""
    R = 1737400.0
    h = gpu_ctx.synth_fbm(65, 65, 5, 2000.0).cpu().numpy()
    ps = 20000.0
    p = Q.DEMProjection(resolution=(65, 65), pixel_scale=(ps, -ps), is_ogr=True,
                        proj_ref=Q.polar_stereographic_wkt(R, south=south, central_meridian=lon0,
                                                           latitude_of_origin=lat),
                        model_tie_point=(-32.5 * ps, 32.5 * ps))
    path = tmp_path / "ps.qld"
    Q.write_pyramid(str(path), [Q.DEMLevel(h, p)], compress=True)
    d = Q.QuipuDEM(str(path))
    mid, hd, _, info = d.read_level(gpu_ctx, 0, samples=True)
    pos, _, _ = gpu_ctx.mesh_download(mid)
    want = Q.level_positions(Q.DEMLevel(hd.cpu().numpy(), p)).reshape(-1, 3).astype(np.float32)
    ulp = np.spacing(np.abs(want))
    assert np.all(np.abs(pos - want) <= ulp), np.abs(pos - want).max()
    # the centre sample is the pole
    c = pos.reshape(65, 65, 3)[32, 32].astype(np.float64)
    assert np.sign(c[2]) == (-1 if south else 1) and np.hypot(c[0], c[1]) < 2 * ps
    d.close()
    # polar stereographic on an ellipsoid is refused
    p.proj_ref = p.proj_ref.replace(f"{R!r},0.0]", f"{R!r},298.257]")
    Q.write_pyramid(str(tmp_path / "pse.qld"), [Q.DEMLevel(h, p)], compress=False)
    d2 = Q.QuipuDEM(str(tmp_path / "pse.qld"))
    with pytest.raises(_capi.ViraGPUError, match="sphere only"):
        d2.read_level(gpu_ctx, 0)
    d2.close()


def test_read_buffers_keeps_the_loaded_level(gpu_ctx, tmp_path):
    lv = _pyramid(gpu_ctx, base=65, levels=3, holes=False)
    path = tmp_path / "lod.qld"
    Q.write_pyramid(str(path), lv, compress=True)
    d = Q.QuipuDEM(str(path))
    m0 = d.read_buffers(gpu_ctx)                 # default GSD +inf: the coarsest level
    assert d.current_level == 2 and d.default_gsd == d.gsds[2]
    assert d.read_buffers(gpu_ctx, 100.0) == m0  # still the coarsest: nothing reloaded
    m1 = d.read_buffers(gpu_ctx, 20.0)
    assert d.current_level == 0 and m1 != m0
    nv, nt = gpu_ctx.mesh_info(m1)
    assert nv == 65 * 65 and nt == 2 * 64 * 64
    d.close()


def test_highly_compressible_level_reads(gpu_ctx, tmp_path):
    """A nearly flat 513^2 DEM and a near-constant grey albedo image compress to a file
    far smaller than the decoded samples; the reader bounds a compressed image by its
    LZ4 block (at most ~255-fold expansion), not by the file size, so it reads them."""
    n = 513
    h = np.zeros((n, n), np.float32)
    h[0, 0] = 1.0                      # min != max, so the fixed-point range is defined
    alb = np.full((n, n), 0.12, np.float32)
    alb[0, 0] = 0.13
    p = Q.DEMProjection(resolution=(n, n), pixel_scale=(20.0, 20.0))
    lv = [Q.DEMLevel(h, p, albedo=alb)]
    path = tmp_path / "flat.qld"
    Q.write_pyramid(str(path), lv, compress=True)
    size = path.stat().st_size
    assert size < n * n  # smaller than even the 8-bit samples
    d = Q.QuipuDEM(str(path))
    mid, got_h, got_a, info = d.read_level(gpu_ctx, 0, mesh=True, samples=True)
    assert info["compressed"] == 1 and info["albedo_type"] == 1
    mm = Q._valid_minmax(h)
    bits = info["height_bits"]
    want = _fixed_to_float(Q.float_to_fixed(h, mm, np.uint8 if bits == 8 else np.uint16), mm) if bits < 32 else h
    assert np.array_equal(got_h.cpu().numpy().view(np.uint32), want.view(np.uint32))
    am = Q._valid_minmax(alb)
    wa = _fixed_to_float(Q.float_to_fixed(alb, am, np.uint8), am)
    assert np.array_equal(got_a.cpu().numpy()[..., 0], wa)
    assert gpu_ctx.mesh_info(mid)[1] == 2 * (n - 1) ** 2
    d.close()

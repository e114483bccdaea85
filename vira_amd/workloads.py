"""The benchmark scenes: seeded fBm DEMs, camera, sun and render options.

Each scene is plain data plus a builder that creates it on the GPU through
the C-ABI.  Synthetic inputs are generated on the device (value-noise fBm,
hashed per-vertex albedo, a seeded log-normal albedo texture).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np

from . import scene as S

MOON_RADIUS = 1737.4e3


@dataclass(frozen=True)
class DemScene:
    name: str
    dem: int                 # DEM samples per side
    seed: int
    amplitude: float = 500.0
    spacing: float = 30.0
    octaves: int = 6
    base_cells: float = 64.0
    sphere_radius: float = 0.0   # > 0: gnomonic drape on a sphere + camera-relative FP64 rebasing
    altitude: float = 20000.0
    image: int = 512
    fov_deg: float = 40.0
    spp: int = 1
    bounces: int = 0
    russian_roulette: bool = False
    sun_elevation_deg: float = 30.0
    sun_kind: str = "point"      # "point": distant PointLight, flat 1 W/m^2/band; "sphere": newSun at 1 AU
    albedo: str = "vertex"       # "vertex": 0.12 +- 0.02 per vertex; "texture": log-normal texture
    texture: int = 0             # texture side
    smooth_shading: bool = False
    render_seed: int = 0

    @property
    def triangles(self) -> int:
        return 2 * (self.dem - 1) ** 2


# 257^2 DEM, 512^2, 1 spp, primary + shadow, deterministic.
DEM257 = DemScene("dem257_512px_1spp_direct", dem=257, seed=0)
# 4097^2 DEM (33.5M triangles) on a 1737 km sphere, 2048^2, 16 spp, 3 bounces + RR.
DEM4097 = DemScene("dem4097_2048px_16spp_3b", dem=4097, seed=1, sphere_radius=MOON_RADIUS, altitude=100e3,
                   image=2048, spp=16, bounces=3, russian_roulette=True, sun_kind="sphere",
                   albedo="texture", texture=4096)

# 1025^2 DEM (2.1M triangles), 1024^2, direct illumination with fixed
# visibility; per-texel albedo (1024^2 log-normal texture) and the heights are
# the differentiable parameters.  40 km altitude: the DEM fills the image.
DEM1025_DI = DemScene("dem1025_1024px_di_adjoint", dem=1025, seed=2, altitude=40000.0, image=1024,
                   albedo="texture", texture=1024)


def scaled(sc: DemScene, dem: int | None = None, image: int | None = None, spp: int | None = None,
           bounces: int | None = None, texture: int | None = None) -> DemScene:
    """A smaller copy of a scene (same generator and geometry rules); the camera
    altitude shrinks with the DEM so the DEM covers the same part of the image."""
    kw = {}
    if dem is not None:
        kw["dem"] = dem
        kw["altitude"] = sc.altitude * (dem - 1) / (sc.dem - 1)
    if image is not None:
        kw["image"] = image
    if spp is not None:
        kw["spp"] = spp
    if bounces is not None:
        kw["bounces"] = bounces
    if texture is not None:
        kw["texture"] = texture
    return replace(sc, **kw)


def camera_position(sc: DemScene) -> tuple:
    if sc.sphere_radius > 0:
        return (sc.sphere_radius + sc.altitude, 0.0, 0.0)
    return (0.0, 0.0, sc.altitude)


def rebase_offset(sc: DemScene) -> tuple:
    """FP64 origin subtracted from every position before the FP32 store."""
    return camera_position(sc) if sc.sphere_radius > 0 else (0.0, 0.0, 0.0)


def make_projection(sc: DemScene):
    if sc.sphere_radius > 0:
        return S.projection(1, sc.spacing, sc.sphere_radius, (0.0, 0.0, 0.0), rebase_offset(sc))
    return S.projection(0, sc.spacing)


def make_camera(sc: DemScene) -> S.Camera:
    pos = np.asarray(camera_position(sc)) - np.asarray(rebase_offset(sc))
    cam = S.Camera(sc.image, sc.image, sc.fov_deg, tuple(pos))
    if sc.sphere_radius > 0:
        cam.look_in_direction((-1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0))
    else:
        cam.look_in_direction((0.0, 0.0, -1.0))
    return cam


def make_lights(sc: DemScene):
    off = np.asarray(rebase_offset(sc))
    if sc.sphere_radius > 0:
        up, east = (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)
        ground = np.array([sc.sphere_radius, 0.0, 0.0])
    else:
        up, east = (0.0, 0.0, 1.0), (1.0, 0.0, 0.0)
        ground = np.zeros(3)
    origin = ground - off
    if sc.sun_kind == "point":
        return [S.directional_sun(sc.sun_elevation_deg, 0.0, 1.0, origin=origin, up=up, east=east)]
    pl = S.directional_sun(sc.sun_elevation_deg, 0.0, 1.0, origin=origin, up=up, east=east)
    return [S.new_sun(pl.position)]


def make_options(sc: DemScene) -> S.PathTracerOptions:
    return S.PathTracerOptions(samples=sc.spp, bounces=sc.bounces, russian_roulette=sc.russian_roulette,
                               rr_start=1, seed=sc.render_seed)


def make_texture(ctx: S.Context, sc: DemScene):
    """Seeded log-normal grey albedo texture (median 0.11, sigma 0.25)."""
    import torch
    g = torch.Generator(device=f"cuda:{ctx.device}")
    g.manual_seed(1000 + sc.seed)
    z = torch.randn((sc.texture, sc.texture), generator=g, device=f"cuda:{ctx.device}", dtype=torch.float32)
    return (0.11 * torch.exp(0.25 * z)).contiguous()


@dataclass
class BuiltScene:
    scene: DemScene
    mesh: int
    heights: object
    vertex_albedo: object
    texture: object
    camera: S.Camera
    lights: list
    options: S.PathTracerOptions
    bvh: dict


def build(ctx: S.Context, sc: DemScene, bvh_max_leaf: int = 4, bvh_bins: int = 16, bvh_width: int = 4,
          bvh_traversal_cost: float = 0.0, bvh_method: str = "auto") -> BuiltScene:
    heights = ctx.synth_fbm(sc.dem, sc.dem, sc.seed, sc.amplitude, sc.octaves, sc.base_cells)
    valb = ctx.synth_vertex_albedo(sc.dem * sc.dem, sc.seed) if sc.albedo == "vertex" else None
    mesh = ctx.dem_mesh(heights, make_projection(sc), valb)
    bvh = ctx.blas_build(mesh, bvh_max_leaf, bvh_bins, bvh_width, bvh_traversal_cost, bvh_method)
    ctx.tlas_build([mesh])
    lights = make_lights(sc)
    ctx.lights_set(lights)
    tex = make_texture(ctx, sc) if sc.albedo == "texture" else None
    ctx.material_set(mesh, albedo=1.0, texture=tex, smooth_shading=sc.smooth_shading)
    return BuiltScene(sc, mesh, heights, valb, tex, make_camera(sc), lights, make_options(sc), bvh)


@dataclass
class DiffScene:
    built: BuiltScene
    dem: object               # autograd.DifferentiableDem
    target: object            # (H, W, 8) image of the perturbed DEM
    target_heights: object
    target_texture: object


def perturbed_inputs(ctx: S.Context, sc: DemScene, heights, texture, height_sigma: float = 2.0,
                     albedo_sigma: float = 0.05, seed: int | None = None):
    """Target parameters: heights + N(0, height_sigma) and texture x (1 + N(0, albedo_sigma)),
    from a torch generator seeded with 2000 + scene seed."""
    import torch
    g = torch.Generator(device=f"cuda:{ctx.device}")
    g.manual_seed(2000 + sc.seed if seed is None else seed)
    dh = torch.randn(heights.shape, generator=g, device=heights.device, dtype=torch.float32) * height_sigma
    th = heights + torch.where(torch.isfinite(heights), dh, torch.zeros_like(dh))
    tt = None
    if texture is not None:
        dt = torch.randn(texture.shape, generator=g, device=texture.device, dtype=torch.float32) * albedo_sigma
        tt = texture * (1.0 + dt)
    return th.contiguous(), None if tt is None else tt.contiguous()


def build_diff(ctx: S.Context, sc: DemScene, options: S.DIOptions | None = None, bvh_width: int = 4) -> DiffScene:
    """Differentiable scene of the 1025^2 kind: the base DEM on the GPU plus the
    target image rendered from the perturbed heights and texture."""
    from .autograd import DifferentiableDem
    b = build(ctx, sc, bvh_width=bvh_width)
    dem = DifferentiableDem(ctx, b.mesh, b.camera, options or S.DIOptions(), smooth_shading=sc.smooth_shading)
    th, tt = perturbed_inputs(ctx, sc, b.heights, b.texture)
    import torch
    with torch.no_grad():
        target = dem(th, tt).clone()
    dem._apply(b.heights, b.texture)
    return DiffScene(b, dem, target, th, tt)


# ---------------------------------------------------------------- batched mixed workload (64 requests)
def _lowbias32(x):
    x = np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def _lattice3(ix, iy, iz, octave, seed):
    h = _lowbias32(np.uint64(seed) ^ np.uint64(0x9E3779B9))
    h = _lowbias32(h ^ np.uint64(octave))
    for c in (ix, iy, iz):
        h = _lowbias32(h ^ (np.asarray(c, np.int64).astype(np.uint64) & 0xFFFFFFFF))
    return h.astype(np.float64) / 4294967295.0 * 2.0 - 1.0


def fbm3(p: np.ndarray, seed: int, octaves: int = 6) -> np.ndarray:
    """3-D value-noise fBm in [-1, 1] at points p (n, 3): lacunarity 2, gain 0.5,
    quintic fade, lattice values hashed with lowbias32 (seed, octave, ix, iy, iz)."""
    total = np.zeros(p.shape[0])
    amp, norm, freq = 1.0, 0.0, 1.0
    for o in range(octaves):
        q = p * freq
        i0 = np.floor(q).astype(np.int64)
        f = q - i0
        w = f * f * f * (f * (f * 6.0 - 15.0) + 10.0)
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    v = _lattice3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, o, seed)
                    wt = ((w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) *
                          (w[:, 2] if dz else 1 - w[:, 2]))
                    acc += wt * v
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def small_body_mesh(seed: int = 3, axes=(500.0, 350.0, 300.0), nlat: int = 251, nlon: int = 400,
                    relief: float = 0.08):
    """Seeded small-body shape model: a latitude/longitude tessellated ellipsoid
    (semi-axes `axes`, metres) with radial fBm relief; 2 * nlon * (nlat - 1)
    triangles (200,000 by default), outward winding.  Returns (pos, idx)."""
    th = np.pi * np.arange(1, nlat) / nlat
    ph = 2.0 * np.pi * np.arange(nlon) / nlon
    T, P = np.meshgrid(th, ph, indexing="ij")
    ring = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dirs = np.concatenate([[[0.0, 0.0, 1.0]], ring, [[0.0, 0.0, -1.0]]])
    a, b, c = axes
    re = 1.0 / np.sqrt((dirs[:, 0] / a) ** 2 + (dirs[:, 1] / b) ** 2 + (dirs[:, 2] / c) ** 2)
    r = re * (1.0 + relief * fbm3(dirs * 2.0 + 7.0, seed))
    pos = dirs * r[:, None]
    N, S = 0, 1 + (nlat - 1) * nlon
    vid = lambda i, j: 1 + i * nlon + (j % nlon)
    tris = []
    for j in range(nlon):
        tris.append((N, vid(0, j), vid(0, j + 1)))
    for i in range(nlat - 2):
        for j in range(nlon):
            tris.append((vid(i, j), vid(i + 1, j), vid(i + 1, j + 1)))
            tris.append((vid(i, j), vid(i + 1, j + 1), vid(i, j + 1)))
    for j in range(nlon):
        tris.append((S, vid(nlat - 2, j + 1), vid(nlat - 2, j)))
    idx = np.asarray(tris, np.uint32)
    p0, p1, p2 = pos[idx[:, 0]], pos[idx[:, 1]], pos[idx[:, 2]]
    fn = np.cross(p1 - p0, p2 - p0)
    if np.mean(np.sum(fn * (p0 + p1 + p2), 1)) < 0:
        idx = idx[:, [0, 2, 1]].copy()
    return pos.astype(np.float32), idx


BATCH_SIDES = (128, 256, 512, 1024, 2048)
BATCH_SPP = (1, 4, 16, 64)


@dataclass
class BatchMeshes:
    ids: list          # mesh ids: 257^2 DEM, 4097^2 DEM, 1025^2 DEM, small body
    scenes: list       # DemScene (or None for the small body)
    keep: list         # tensors the meshes' materials point at


def build_batch_meshes(ctx: S.Context, dem_built: "BuiltScene | None" = None, bvh_width: int = 4) -> BatchMeshes:
    """The four meshes of the batched workload: the 257^2, 4097^2 and 1025^2
    DEMs (with their albedo) and the 200k-triangle small body (albedo 0.12)."""
    ids, scenes, keep = [], [], []
    for sc in (DEM257, DEM4097, DEM1025_DI):
        b = dem_built if (sc is DEM4097 and dem_built is not None) else build(ctx, sc, bvh_width=bvh_width)
        ids.append(b.mesh)
        scenes.append(sc)
        keep.append(b.texture)
        keep.append(b.heights)
    pos, idx = small_body_mesh()
    mid = ctx.mesh_upload(pos, idx)
    ctx.blas_build(mid, width=bvh_width)
    ctx.material_set(mid, albedo=0.12)
    ids.append(mid)
    scenes.append(None)
    # the 4097^2 DEM's material must point at its texture again (build() of other scenes does not touch it)
    return BatchMeshes(ids, scenes, keep)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def make_batch(ctx: S.Context, meshes: BatchMeshes, n: int = 64, seed: int = 3, bounces: int = 3,
               passes=("radiance", "depth", "tri_id"), sides=BATCH_SIDES, spps=BATCH_SPP):
    """Mixed batch: n requests, image side uniform over `sides`, spp uniform over
    `spps`, mesh uniform over the four meshes, random camera and sun per request
    (numpy Generator seeded with `seed`)."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        side = int(rng.choice(sides))
        spp = int(rng.choice(spps))
        k = int(rng.integers(0, len(meshes.ids)))
        sc = meshes.scenes[k]
        tilt = np.radians(rng.uniform(0.0, 20.0))
        az = rng.uniform(0.0, 2.0 * np.pi)
        elev = rng.uniform(20.0, 60.0)
        sun_az = rng.uniform(0.0, 360.0)
        if sc is None:  # small body: camera 1.5-3 km away looking at its centre
            d = _unit(rng.standard_normal(3))
            pos = d * rng.uniform(1500.0, 3000.0)
            cam = S.Camera(side, side, 30.0, tuple(pos))
            up = (0.0, 0.0, 1.0) if abs(d[2]) < 0.9 else (1.0, 0.0, 0.0)
            cam.look_in_direction(tuple(-d), up=up)
            sd = _unit(rng.standard_normal(3))
            e = _unit(np.cross(sd, [0.0, 0.0, 1.0] if abs(sd[2]) < 0.9 else [1.0, 0.0, 0.0]))
            light = S.directional_sun(90.0, 0.0, 1.0, up=tuple(sd), east=tuple(e))
        elif sc.sphere_radius > 0:  # 4097^2 DEM: rebased frame, ground 100 km below along -x
            ext = (sc.dem - 1) * sc.spacing
            pos = (rng.uniform(-0.3, 0.3) * sc.altitude, rng.uniform(-0.15, 0.15) * ext,
                   rng.uniform(-0.15, 0.15) * ext)
            cam = S.Camera(side, side, sc.fov_deg, pos)
            cam.look_in_direction((-np.cos(tilt), np.sin(tilt) * np.cos(az), np.sin(tilt) * np.sin(az)),
                                  up=(0.0, 0.0, 1.0))
            off = np.asarray(rebase_offset(sc))
            ground = np.array([sc.sphere_radius, 0.0, 0.0]) - off
            light = S.directional_sun(elev, sun_az, 1.0, origin=ground, up=(1.0, 0.0, 0.0), east=(0.0, 1.0, 0.0))
        else:  # the flat 257^2 and 1025^2 DEMs
            ext = (sc.dem - 1) * sc.spacing
            pos = (rng.uniform(-0.15, 0.15) * ext, rng.uniform(-0.15, 0.15) * ext,
                   sc.altitude * rng.uniform(0.7, 1.3))
            cam = S.Camera(side, side, sc.fov_deg, pos)
            cam.look_in_direction((np.sin(tilt) * np.cos(az), np.sin(tilt) * np.sin(az), -np.cos(tilt)))
            light = S.directional_sun(elev, sun_az, 1.0)
        opts = S.PathTracerOptions(samples=spp, bounces=bounces, russian_roulette=True, rr_start=1, seed=1000 + i)
        reqs.append(S.RenderRequest(meshes.ids[k], cam, opts, light, S.RenderPasses(side, side, which=passes)))
    return reqs


def render_plain(ctx: S.Context, reqs) -> None:
    """The per-request form: one synchronous render call per request."""
    for r in reqs:
        ctx.tlas_build([r.mesh])
        ctx.lights_set([] if r.light is None else [r.light])
        ctx.render(r.camera, r.options, r.passes)


# ---------------------------------------------------------------- instanced multi-body scene at 1 AU
AU = S.AU


def box_mesh(size, center=(0.0, 0.0, 0.0), n: int = 8):
    """Axis-aligned box (size = full extents) with each face split into an n x n
    grid of quads (12 n^2 triangles), outward winding, per-face vertices."""
    sx, sy, sz = (0.5 * float(v) for v in size)
    c = np.asarray(center, np.float64)
    g = np.linspace(-1.0, 1.0, n + 1)
    U, V = np.meshgrid(g, g, indexing="ij")
    pos, idx = [], []
    faces = [  # (axis, sign)
        (0, 1.0), (0, -1.0), (1, 1.0), (1, -1.0), (2, 1.0), (2, -1.0)]
    half = np.array([sx, sy, sz])
    for axis, sign in faces:
        a1, a2 = [k for k in range(3) if k != axis]
        P = np.zeros(((n + 1) ** 2, 3))
        P[:, axis] = sign * half[axis]
        P[:, a1] = U.reshape(-1) * half[a1]
        P[:, a2] = V.reshape(-1) * half[a2]
        base = sum(len(p) for p in pos)
        q = np.arange(n)
        i, j = np.meshgrid(q, q, indexing="ij")
        v00 = (i * (n + 1) + j).reshape(-1)
        v10, v01, v11 = v00 + (n + 1), v00 + 1, v00 + (n + 1) + 1
        t = np.concatenate([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)]) + base
        # outward: cross(e1, e2) along +axis*sign
        p0, p1, p2 = P[t[0, 0] - base], P[t[0, 1] - base], P[t[0, 2] - base]
        if np.dot(np.cross(p1 - p0, p2 - p0), np.eye(3)[axis] * sign) < 0:
            t = t[:, [0, 2, 1]]
        pos.append(P + c)
        idx.append(t)
    return np.concatenate(pos).astype(np.float32), np.concatenate(idx).astype(np.uint32)


def spacecraft_mesh():
    """Box-and-panel spacecraft (50,496 triangles): a 3 x 2 x 2 m bus, two
    6 x 2 m solar panels and a 3 m boom along -x (local frame: bus centre)."""
    parts = [box_mesh((3.0, 2.0, 2.0), (0.0, 0.0, 0.0), 40),
             box_mesh((0.05, 6.0, 2.0), (0.0, 4.0, 0.0), 36),
             box_mesh((0.05, 6.0, 2.0), (0.0, -4.0, 0.0), 36),
             box_mesh((3.0, 0.1, 0.1), (-3.0, 0.6, 0.5), 4)]
    pos, idx, base = [], [], 0
    for p, t in parts:
        pos.append(p)
        idx.append(t + base)
        base += len(p)
    return np.concatenate(pos), np.concatenate(idx)


def _rot_z(deg):
    a = np.radians(deg)
    R = np.eye(4)
    R[0, 0], R[0, 1], R[1, 0], R[1, 1] = np.cos(a), -np.sin(a), np.sin(a), np.cos(a)
    return R


def _rot_axis(axis, deg):
    k = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    a = np.radians(deg)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    R = np.eye(4)
    R[:3, :3] = np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * (K @ K)
    return R


def _trans(t):
    T = np.eye(4)
    T[:3, 3] = t
    return T


@dataclass
class InstancedScene:
    meshes: list        # mesh ids: DEM tile, sphere, small body, spacecraft
    transforms: list    # FP64 4x4 model matrices (row-major)
    camera: S.Camera
    lights: list
    options: S.PathTracerOptions
    keep: list
    local: dict         # host copies of the generic meshes (for the oracle)


def build_instanced(ctx: S.Context, dem_built: "BuiltScene | None" = None, dem: int = 4097, image: int = 4096,
                    spp: int = 8, bounces: int = 2, sphere=(161, 200), body=(251, 400), sun: str = "sphere",
                    bvh_width: int = 4) -> InstancedScene:
    """Instanced scene: the Moon (a 4097^2 DEM tile + a 64k-triangle sphere for the
    rest) at 1 AU from the Sun, a 200k-triangle small body 30 km from the
    camera and a 50,496-triangle spacecraft carrying the camera, every body
    placed by an FP64 model matrix at ~1.5e8 km from the origin; camera 100 km
    above the tile (scaled with the DEM), 4096^2, 8 spp, 2 bounces."""
    sc = DEM4097 if dem == DEM4097.dem else scaled(DEM4097, dem=dem, texture=min(DEM4097.texture, 4 * (dem - 1)))
    b = dem_built if dem_built is not None else build(ctx, sc, bvh_width=bvh_width)
    R = sc.sphere_radius
    alt = sc.altitude
    P_moon = np.array([AU, 0.0, 0.0])
    Rz = _rot_z(120.0)                                   # tile centre direction u = Rz (1, 0, 0)
    u = Rz[:3, :3] @ np.array([1.0, 0.0, 0.0])
    v = Rz[:3, :3] @ np.array([0.0, 1.0, 0.0])
    P_cam = P_moon + (R + alt) * u                       # the DEM mesh's local origin
    M_tile = _trans(P_cam) @ Rz
    sp_pos, sp_idx = small_body_mesh(seed=0, axes=(R - 1000.0,) * 3, nlat=sphere[0], nlon=sphere[1], relief=0.0)
    sphere_id = ctx.mesh_upload(sp_pos, sp_idx)
    ctx.blas_build(sphere_id, width=bvh_width)
    ctx.material_set(sphere_id, albedo=0.11)
    M_sphere = _trans(P_moon)
    sb_pos, sb_idx = small_body_mesh(nlat=body[0], nlon=body[1])
    sb_id = ctx.mesh_upload(sb_pos, sb_idx)
    ctx.blas_build(sb_id, width=bvh_width)
    ctx.material_set(sb_id, albedo=0.12)
    d_sb = 0.3 * alt
    M_sb = _trans(P_cam - d_sb * u + 0.04 * d_sb / 0.3 * v) @ _rot_axis((1.0, 1.0, 1.0), 40.0)
    sc_pos, sc_idx = spacecraft_mesh()
    sc_id = ctx.mesh_upload(sc_pos, sc_idx)
    ctx.blas_build(sc_id, width=bvh_width)
    ctx.material_set(sc_id, albedo=0.5)
    M_sc = _trans(P_cam + 2.0 * u) @ Rz
    meshes = [b.mesh, sphere_id, sb_id, sc_id]
    transforms = [M_tile, M_sphere, M_sb, M_sc]
    ctx.tlas_build(meshes, transforms)
    cam = S.Camera(image, image, 40.0, tuple(P_cam))
    cam.look_in_direction(tuple(-u), up=(0.0, 0.0, 1.0))
    if sun == "sphere":
        lights = [S.new_sun((0.0, 0.0, 0.0))]
    else:  # a point sun at the origin giving 1 W/m^2 per band at 1 AU
        lights = [S.PointLight((0.0, 0.0, 0.0), np.full(8, AU * AU, np.float32))]
    ctx.lights_set(lights)
    opts = S.PathTracerOptions(samples=spp, bounces=bounces, seed=4)
    local = {"sphere": (sp_pos, sp_idx), "body": (sb_pos, sb_idx), "craft": (sc_pos, sc_idx)}
    return InstancedScene(meshes, transforms, cam, lights, opts, [b.heights, b.texture], local)


# ---------------------------------------------------------------- the 100,000-star field
def johnson_v_band(n: int = 1000):
    """johnsonVBandApproximation (magnitudes.ipp): skew-normal V-band response on 450-725 nm (FP64)."""
    from scipy.special import erfc
    lam = np.linspace(450e-9, 725e-9, n)
    sigma, mu, alpha, scale = 63e-9, 505e-9, 6e7, 5.4729e6
    x = lam - mu
    s2 = sigma * sigma
    a = 1.0 / np.sqrt(2 * np.pi * s2)
    b = np.exp(-(x * x) / (2 * s2))
    cdf = 0.5 * erfc(-(alpha * x) * (1.0 / np.sqrt(2.0)))
    return lam, a * b * cdf / scale


def star_temperature_omega(vmag, bv: float = 0.3):
    """Temperature (K) and solid angle (sr) of stars of V magnitude vmag and colour
    index bv (Tycho2Interface::initIrradiance, tycho2_interface.ipp:112-150): T
    from B-V, solid angle from the V-band photon flux over the black body's."""
    c, h, kb = 299792458.0, 6.62607015e-34, 1.380649e-23
    T = 4600.0 * (1.0 / (0.92 * bv + 1.7) + 1.0 / (0.92 * bv + 0.62))
    # Band(550 nm, fwhm 0.16, 3640 Jy).fluxFromMagnitude (magnitudes.hpp)
    freq = c / 550.0
    jy2photon = (1e-23 * 10000 * freq) / ((h * freq) / 1e-7)
    flux = np.power(10.0, np.asarray(vmag, np.float64) / -2.5) * 3640.0 * jy2photon * 0.16
    lam, v = johnson_v_band()
    planck = 2 * h * c * c / lam ** 5 / (np.exp(h * c / (lam * kb * T)) - 1.0)
    counts = v * planck / (h * c / lam)
    flux2 = np.sum(0.5 * (counts[1:] + counts[:-1]) * np.diff(lam))
    return T, flux / flux2


def star_irradiance(vmag, bv: float = 0.3):
    """Spectral irradiance (n, 8) of stars of V magnitude vmag and colour index bv
    (Star::getIrradiance, star.ipp:13-16): solid angle times the black-body band
    radiance."""
    T, omega = star_temperature_omega(vmag, bv)
    band = S.planck_band_radiance(T, steps=100)
    return (omega[:, None] * band[None, :]).astype(np.float32)


def make_star_field(n: int = 100_000, seed: int = 4, mag_max: float = 10.0):
    """n stars: V magnitudes from an exponential law on [0, mag_max] (density
    proportional to e^m, faint stars most numerous), directions uniform on the
    sky; numpy Generator(seed).  Returns (directions (n, 3), irradiance (n, 8),
    magnitudes)."""
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    mag = np.log1p(u * np.expm1(mag_max))
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    r = np.sqrt(1.0 - z * z)
    dirs = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)
    return dirs, star_irradiance(mag), mag


def make_star_catalogue(n: int = 100_000, seed: int = 4, mag_max: float = 10.0, pm_sigma: float = 1e-8):
    """The star field of make_star_field(n, seed, mag_max) as a star catalogue:
    64-byte star quipu records (RA = atan2(y, x), DE = asin(z) of the same
    directions, temperature and solid angle from the V magnitude at B-V 0.3) with
    proper motions N(0, pm_sigma) rad/yr from numpy Generator(seed + 1)."""
    from . import quipu as Q
    dirs, _, mag = make_star_field(n, seed, mag_max)
    d = dirs.astype(np.float64)
    T, omega = star_temperature_omega(mag)
    pm = np.random.default_rng(seed + 1).normal(0.0, pm_sigma, (2, n))
    return Q.make_stars(np.arctan2(d[:, 1], d[:, 0]), np.arcsin(np.clip(d[:, 2], -1, 1)), mag, T, omega,
                        pm_ra=pm[0], pm_de=pm[1])


def star_field_camera(image: int = 4096, direction=(0.6, -0.48, 0.64)):
    """The instanced scene's camera (4096^2, 40 degree field, f = 50 mm, f/2.8) for its star
    field, with the Airy-disk PSF of those optics (camera.ipp:90-97: focal
    length, aperture diameter f/N, pixel size = sensor size / resolution)."""
    from . import sensor as SN
    cam = S.Camera(image, image, 40.0)
    cam.look_in_direction(direction, up=(0.0, 0.0, 1.0))
    sw, sh = cam.sensor_size()
    f32 = np.float32
    psf = SN.AiryDiskPSF(focal_length=cam.focal_length,
                         aperture_diameter=float(f32(cam.focal_length) / f32(cam.f_stop)),
                         pixel_size=(float(f32(sw) / f32(image)), float(f32(sh) / f32(image))))
    return cam, psf

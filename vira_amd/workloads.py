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


# Config 0: 257^2 DEM, 512^2, 1 spp, primary + shadow, deterministic.
CONFIG0 = DemScene("dem257_512px_1spp_direct", dem=257, seed=0)
# Config 1: 4097^2 DEM (33.5M triangles) on a 1737 km sphere, 2048^2, 16 spp, 3 bounces + RR.
CONFIG1 = DemScene("dem4097_2048px_16spp_3b", dem=4097, seed=1, sphere_radius=MOON_RADIUS, altitude=100e3,
                   image=2048, spp=16, bounces=3, russian_roulette=True, sun_kind="sphere",
                   albedo="texture", texture=4096)

# Config 2: 1025^2 DEM (2.1M triangles), 1024^2, direct illumination with fixed
# visibility; per-texel albedo (1024^2 log-normal texture) and the heights are
# the differentiable parameters.  40 km altitude: the DEM fills the image.
CONFIG2 = DemScene("dem1025_1024px_di_adjoint", dem=1025, seed=2, altitude=40000.0, image=1024,
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


def build(ctx: S.Context, sc: DemScene, bvh_max_leaf: int = 4, bvh_bins: int = 16, bvh_width: int = 4) -> BuiltScene:
    heights = ctx.synth_fbm(sc.dem, sc.dem, sc.seed, sc.amplitude, sc.octaves, sc.base_cells)
    valb = ctx.synth_vertex_albedo(sc.dem * sc.dem, sc.seed) if sc.albedo == "vertex" else None
    mesh = ctx.dem_mesh(heights, make_projection(sc), valb)
    bvh = ctx.blas_build(mesh, bvh_max_leaf, bvh_bins, bvh_width)
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
    """Config-2 style differentiable scene: the base DEM on the GPU plus the
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

"""Host-side scene description and the GPU renderer, mirroring vira's API.

The reference's user-facing objects are ``Camera`` (include/vira/cameras/
camera.hpp), the lights (include/vira/lights), ``Scene`` and
``CPUPathTracer`` with ``options`` / ``renderPasses``
(include/vira/rendering/cpu_path_tracer.hpp:30-71).  This module provides the
same objects over the C-ABI of include/vira_gpu.h; all compute happens in the
HIP library.  Device memory is held in torch tensors (plumbing only).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi

NUM_BINS = capi.NUM_BINS
SOLAR_RADIUS = 696000000.0  # include/vira/math.hpp:55
AU = 1.495978707e11

# SpectralData<8, UNIFORM_RANGE, 400, 1000>: 8 uniform bands (spectral_data.ipp:64-132)
BAND_EDGES_NM = np.linspace(400.0, 1000.0, NUM_BINS + 1)


# ---------------------------------------------------------------- camera
def intrinsics_inverse(width: int, height: int, focal_length: float, sensor_w: float, sensor_h: float,
                       z_dir: float = 1.0) -> np.ndarray:
    """K^-1 as in Camera::initializeIntrinsicMatrix (camera.ipp:1526-1568), FP32.

    Returns the 6 entries [col][row] = (k00, k01, k10, k11, k20, k21).
    """
    f32 = np.float32
    kx = f32(width) / f32(sensor_w)
    ky_ = f32(height) / f32(sensor_h)
    px_ = f32(width) / f32(2)
    py_ = f32(height) / f32(2)
    zd = f32(z_dir)
    ky = zd * ky_
    px = zd * px_
    py = zd * py_
    kxy = zd * f32(0)
    kyx = zd * f32(0)
    fx = f32(focal_length) * kx
    fy = f32(focal_length) * ky
    det = (fx * fy) - (kxy * kyx)
    k00 = fy / det
    k10 = -kxy / det
    k20 = ((kxy * py) - (fy * px)) / det
    k01 = -kyx / det
    k11 = fx / det
    k21 = ((kyx * px) - (fx * py)) / det
    return np.array([k00, k01, k10, k11, k20, k21], dtype=np.float32)


def look_in_direction(direction, up=(0.0, 0.0, 1.0), blender_frame: bool = False) -> np.ndarray:
    """Camera::lookInDirection (camera.ipp:1455-1500); returns R (3x3, columns = axes)."""
    def nrm(v):
        v = np.asarray(v, dtype=np.float64)
        return v / np.sqrt(np.dot(v, v))
    up = nrm(up)
    d = nrm(direction)
    if abs(np.dot(d, up)) > 0.999:
        for cand in ((0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0)):
            if abs(np.dot(d, cand)) < 0.999:
                up = np.asarray(cand)
                break
    if blender_frame:
        y_axis, z_axis = up, -d
    else:
        y_axis, z_axis = -up, d
    x_axis = nrm(np.cross(y_axis, z_axis))
    y_axis = nrm(np.cross(z_axis, x_axis))
    return np.stack([x_axis, y_axis, z_axis], axis=1)


@dataclass
class Camera:
    """Pinhole camera: resolution, square pixels, horizontal field of view."""
    width: int
    height: int
    fov_deg: float = 40.0
    position: tuple = (0.0, 0.0, 0.0)
    rotation: np.ndarray = field(default_factory=lambda: np.eye(3))
    focal_length: float = 0.05
    z_dir: float = 1.0
    f_stop: float = 2.8            # camera.hpp:235
    exposure_time: float = 0.01    # camera.hpp:233
    optical_efficiency: float = 1.0
    view_normal: object = None     # Camera::getViewNormalMatrix (3x3); None = transpose of rotation

    def aperture_area(self) -> float:
        """CircularAperture of diameter f/N (camera.ipp:82, aperture.ipp:14-16), FP32 as the reference."""
        f32 = np.float32
        d = f32(self.focal_length) / f32(self.f_stop)
        r = d / f32(2)
        return float(f32(np.pi) * r * r)

    def sensor(self) -> capi.Sensor:
        s = capi.Sensor()
        eta = np.broadcast_to(np.asarray(self.optical_efficiency, np.float32), (NUM_BINS,))
        for k in range(NUM_BINS):
            s.optical_efficiency[k] = float(eta[k])
        s.aperture_area = self.aperture_area()
        return s

    def look_in_direction(self, direction, up=(0.0, 0.0, 1.0)):
        self.rotation = look_in_direction(direction, up)
        return self

    def sensor_size(self):
        sw = 2.0 * self.focal_length * math.tan(math.radians(self.fov_deg) / 2.0)
        return sw, sw * self.height / self.width

    def to_c(self) -> capi.Camera:
        sw, sh = self.sensor_size()
        kinv = intrinsics_inverse(self.width, self.height, self.focal_length, sw, sh, self.z_dir)
        c = capi.Camera()
        c.width, c.height = self.width, self.height
        for k in range(6):
            c.kinv[k] = float(kinv[k])
        R = np.asarray(self.rotation, dtype=np.float64).astype(np.float32)
        for r in range(3):
            for k in range(3):
                c.rot[3 * r + k] = float(R[r, k])
        for k in range(3):
            c.position[k] = float(self.position[k])
        c.z_dir = float(self.z_dir)
        if self.view_normal is not None:
            vn = np.asarray(self.view_normal, np.float32).reshape(9)
            for k in range(9):
                c.view_normal[k] = float(vn[k])
        return c


# ---------------------------------------------------------------- lights
def planck_band_radiance(temperature: float, steps: int = 1000) -> np.ndarray:
    """blackBodyRadiance (spectral_data.ipp:613-629): trapezoid of Planck's law per band."""
    h, c, kb = 6.62607015e-34, 299792458.0, 1.380649e-23
    out = np.zeros(NUM_BINS)
    for i in range(NUM_BINS):
        lo, hi = BAND_EDGES_NM[i] * 1e-9, BAND_EDGES_NM[i + 1] * 1e-9
        lam = np.linspace(lo, hi, steps)
        rad = (2 * h * c * c / lam**5) / (np.exp(h * c / (lam * kb * temperature)) - 1.0)
        out[i] = np.sum(0.5 * (rad[1:] + rad[:-1]) * np.diff(lam))
    return out.astype(np.float32)


@dataclass
class PointLight:
    """PointLight (point_light.ipp:10-38): spectral power P, intensity P / 4 pi."""
    position: tuple
    intensity: np.ndarray  # P / (4 pi) per band

    def to_c(self) -> capi.Light:
        L = capi.Light()
        L.kind = 0
        for k in range(3):
            L.position[k] = float(self.position[k])
        L.radius = 0.0
        for k in range(NUM_BINS):
            L.value[k] = float(self.intensity[k])
        return L


@dataclass
class SphereLight:
    """SphereLight (sphere_light.ipp:16-125) with spectral radiance."""
    position: tuple
    radius: float
    radiance: np.ndarray

    def to_c(self) -> capi.Light:
        L = capi.Light()
        L.kind = 1
        for k in range(3):
            L.position[k] = float(self.position[k])
        L.radius = float(self.radius)
        for k in range(NUM_BINS):
            L.value[k] = float(self.radiance[k])
        return L


def directional_sun(elevation_deg: float, azimuth_deg: float, irradiance, distance: float = AU,
                    origin=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), east=(1.0, 0.0, 0.0)) -> PointLight:
    """A distant PointLight giving the stated irradiance (W/m^2 per band) at `origin`."""
    up = np.asarray(up, dtype=np.float64)
    east = np.asarray(east, dtype=np.float64)
    north = np.cross(up, east)
    el, az = math.radians(elevation_deg), math.radians(azimuth_deg)
    d = math.cos(el) * (math.cos(az) * east + math.sin(az) * north) + math.sin(el) * up
    pos = np.asarray(origin, dtype=np.float64) + distance * d
    e = np.broadcast_to(np.asarray(irradiance, dtype=np.float64), (NUM_BINS,))
    return PointLight(tuple(pos), (e * distance * distance).astype(np.float32))


def new_sun(position) -> SphereLight:
    """Group::newSun (group.ipp:585-591): SphereLight, R_sun, 5800 K blackbody radiance."""
    return SphereLight(tuple(position), SOLAR_RADIUS, planck_band_radiance(5800.0, 1000))


# ---------------------------------------------------------------- options / passes
@dataclass
class DenoiseOptions:
    """EATWTOptions (cpu_denoise.hpp:9-18)."""
    max_levels_direct: int = 5
    max_levels_indirect: int = 5
    depth_threshold: float = 0.2
    normal_threshold: float = 0.5
    min_normal_weight: float = 0.3
    epsilon: float = 1e-6

    def to_c(self) -> capi.DenoiseOpts:
        return capi.DenoiseOpts(int(self.max_levels_direct), int(self.max_levels_indirect),
                                float(self.depth_threshold), float(self.normal_threshold),
                                float(self.min_normal_weight), float(self.epsilon))


@dataclass
class PathTracerOptions:
    """CPUPathTracerOptions (cpu_path_tracer.hpp:30-43) + the GPU additions.
    denoise / denoiser: CPUPathTracer::options.denoise and denoiserOptions
    (cpu_path_tracer.hpp:35,57), applied after the render."""
    samples: int = 1
    bounces: int = 0
    russian_roulette: bool = False
    rr_start: int = 1
    seed: int = 0
    integrator: str = "auto"  # "auto" | "fused" | "wavefront"
    denoise: bool = False
    denoiser: DenoiseOptions = field(default_factory=DenoiseOptions)
    show_background: bool = False
    adaptive_sampling: bool = False
    samples_per_batch: int = 30
    sampling_tolerance: float = 0.05
    samples_to_detect_miss: int = 30
    count_visits: bool = False  # BVH visit counters into the stats (per-pixel integrator)

    def to_c(self) -> capi.RenderOpts:
        o = capi.RenderOpts()
        o.samples, o.bounces = int(self.samples), int(self.bounces)
        o.russian_roulette, o.rr_start, o.seed = int(self.russian_roulette), int(self.rr_start), int(self.seed)
        o.integrator = {"auto": 0, "fused": 1, "wavefront": 2}[self.integrator]
        o.show_background, o.adaptive_sampling = int(self.show_background), int(self.adaptive_sampling)
        o.samples_per_batch, o.sampling_tolerance = int(self.samples_per_batch), float(self.sampling_tolerance)
        o.samples_to_detect_miss = int(self.samples_to_detect_miss)
        o.count_visits = int(self.count_visits)
        return o


@dataclass
class RenderRequest:
    """One request of a batch: mesh, camera, options, lights, passes."""
    mesh: int
    camera: "Camera"
    options: PathTracerOptions
    light: object  # PointLight | SphereLight | a list of them (<= 8) | None
    passes: "RenderPasses"

    def to_c(self) -> capi.Request:
        r = capi.Request()
        r.mesh_id = self.mesh
        r.camera = self.camera.to_c()
        r.opts = self.options.to_c()
        ls = [] if self.light is None else (list(self.light) if isinstance(self.light, (list, tuple)) else [self.light])
        if len(ls) > capi.MAX_LIGHTS:
            raise ValueError(f"at most {capi.MAX_LIGHTS} lights per request")
        r.num_lights = len(ls)
        for i, L in enumerate(ls):
            r.lights[i] = L.to_c()
        r.out = self.passes.to_c()
        return r


@dataclass
class BatchConfig:
    bvh_width: int = 4
    waves_per_block: int = 1
    waves_per_cu: int = 16
    tiles_per_grab: int = 1
    order: int = 0

    def to_c(self) -> capi.BatchConfig:
        return capi.BatchConfig(self.bvh_width, self.waves_per_block, self.waves_per_cu, self.tiles_per_grab,
                                self.order)


@dataclass
class DIOptions:
    """Fixed-visibility direct-illumination render (forward + adjoint)."""
    samples: int = 1
    seed: int = 0
    record_budget: int = 1 << 30   # bytes of hit records kept for the adjoint

    def to_c(self) -> capi.DIOpts:
        return capi.DIOpts(self.samples, self.seed, self.record_budget)


class RenderPasses:
    """RenderPasses (passes.hpp:51-95) as device tensors, pixel index i + j*W.
    `which` limits the passes allocated (the others are not written)."""

    ALL = ("radiance", "direct", "indirect", "albedo", "normal", "depth", "alpha", "tri_id", "instance_id",
           "normal_camera", "mesh_id", "material_id")

    def __init__(self, width: int, height: int, device="cuda", which=ALL):
        import torch
        self.width, self.height = width, height
        w = set(which)
        spec = lambda: torch.zeros((height, width, NUM_BINS), dtype=torch.float32, device=device)
        self.total_radiance = spec() if "radiance" in w else None
        self.direct_radiance = spec() if "direct" in w else None
        self.indirect_radiance = spec() if "indirect" in w else None
        self.albedo = spec() if "albedo" in w else None
        self.normal_global = (torch.zeros((height, width, 3), dtype=torch.float32, device=device)
                              if "normal" in w else None)
        self.depth = torch.zeros((height, width), dtype=torch.float32, device=device) if "depth" in w else None
        self.alpha = torch.zeros((height, width), dtype=torch.float32, device=device) if "alpha" in w else None
        self.triangle_id = (torch.zeros((height, width), dtype=torch.int32, device=device)
                            if "tri_id" in w else None)
        self.instance_id = (torch.zeros((height, width), dtype=torch.int32, device=device)
                            if "instance_id" in w else None)
        self.normal_camera = (torch.zeros((height, width, 3), dtype=torch.float32, device=device)
                              if "normal_camera" in w else None)
        self.mesh_id = torch.zeros((height, width), dtype=torch.int32, device=device) if "mesh_id" in w else None
        self.material_id = (torch.zeros((height, width), dtype=torch.int32, device=device)
                            if "material_id" in w else None)

    def to_c(self) -> capi.Outputs:
        p = lambda t: None if t is None else t.data_ptr()
        o = capi.Outputs()
        o.radiance_dev = p(self.total_radiance)
        o.direct_dev = p(self.direct_radiance)
        o.indirect_dev = p(self.indirect_radiance)
        o.albedo_dev = p(self.albedo)
        o.normal_dev = p(self.normal_global)
        o.depth_dev = p(self.depth)
        o.alpha_dev = p(self.alpha)
        o.tri_id_dev = p(self.triangle_id)
        o.instance_id_dev = p(self.instance_id)
        o.normal_camera_dev = p(self.normal_camera)
        o.mesh_id_dev = p(self.mesh_id)
        o.material_id_dev = p(self.material_id)
        return o


class Context:
    """One GPU, one HIP stream (torch's current stream by default)."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        import torch
        self.lib = capi.load()
        self.torch = torch
        stream = torch.cuda.current_stream(device).cuda_stream if use_torch_stream else None
        opts = capi.Opts(device, stream)
        h = C.c_void_p()
        capi.check(self.lib.vira_gpu_create(C.byref(h), C.byref(opts)))
        self.h = h
        self.device = device
        self.last_stats = None

    def _call(self, fn, *args):
        """Calls into the library after the work queued on torch's current stream
        has finished (the library runs on its own stream and returns with its
        work complete, so outputs are ready for any stream afterwards)."""
        self.torch.cuda.current_stream(self.device).synchronize()
        capi.check(fn(*args))

    def close(self):
        if self.h:
            capi.check(self.lib.vira_gpu_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- synthetic inputs
    def synth_fbm(self, width: int, height: int, seed: int, amplitude: float = 500.0, octaves: int = 6,
                  base_cells: float = 64.0):
        t = self.torch.empty((height, width), dtype=self.torch.float32, device=f"cuda:{self.device}")
        self._call(self.lib.vira_gpu_synth_fbm, self.h, width, height, seed, amplitude, octaves, base_cells,
                                               t.data_ptr())
        return t

    def synth_vertex_albedo(self, nv: int, seed: int):
        t = self.torch.empty((nv,), dtype=self.torch.float32, device=f"cuda:{self.device}")
        self._call(self.lib.vira_gpu_synth_vertex_albedo, self.h, nv, seed, t.data_ptr())
        return t

    # ---- meshes
    def dem_mesh(self, heights, proj: capi.Projection, grey_albedo=None) -> int:
        H, W = heights.shape
        mid = C.c_int32()
        self._call(self.lib.vira_gpu_dem_mesh, self.h, heights.data_ptr(), W, H, C.byref(proj),
                                              None if grey_albedo is None else grey_albedo.data_ptr(),
                                              C.byref(mid))
        return mid.value

    def mesh_upload(self, pos, idx, albedo=None, uv=None, normals=None) -> int:
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        arr = [None if a is None else np.ascontiguousarray(a, dtype=np.float32) for a in (albedo, uv, normals)]
        p = lambda a: None if a is None else a.ctypes.data
        mid = C.c_int32()
        self._call(self.lib.vira_gpu_mesh_upload, self.h, p(pos), pos.size // 3, p(idx), idx.size // 3,
                                                 p(arr[0]), p(arr[1]), p(arr[2]), C.byref(mid))
        return mid.value

    def mesh_info(self, mid: int):
        nv, nt = C.c_size_t(), C.c_size_t()
        self._call(self.lib.vira_gpu_mesh_info, self.h, mid, C.byref(nv), C.byref(nt))
        return nv.value, nt.value

    def mesh_download(self, mid: int):
        nv, nt = self.mesh_info(mid)
        pos = np.empty((nv, 3), np.float32)
        idx = np.empty((nt, 3), np.uint32)
        nrm = np.empty((nv, 3), np.float32)
        self._call(self.lib.vira_gpu_mesh_download, self.h, mid, pos.ctypes.data, idx.ctypes.data, nrm.ctypes.data)
        return pos, idx, nrm

    def mesh_release(self, mid: int):
        self._call(self.lib.vira_gpu_mesh_release, self.h, mid)

    def blas_build(self, mid: int, max_leaf: int = 4, bins: int = 16, width: int = 4, traversal_cost: float = 0.0,
                   method: str = "auto", compress: bool = False):
        """method: "auto", "sah" (binned SAH) or "grid" (meshes made from a DEM grid, width 4).
        compress: also keep the 64 B quantised node copy (width 4)."""
        o = capi.BVHOpts(max_leaf, bins, width, traversal_cost, {"auto": 0, "sah": 1, "grid": 2}[method],
                         int(compress))
        self._call(self.lib.vira_gpu_blas_build, self.h, mid, C.byref(o))
        s = capi.BVHStats()
        self._call(self.lib.vira_gpu_blas_stats, self.h, mid, C.byref(s))
        return {f: getattr(s, f) for f, _ in capi.BVHStats._fields_}

    def tlas_build(self, mesh_ids, transforms=None):
        """Instances of built BLASes; transforms: FP64 4x4 model matrices (row-major numpy
        arrays, identity if None)."""
        arr = (capi.Instance * len(mesh_ids))()
        for i, m in enumerate(mesh_ids):
            arr[i].mesh_id = m
            M = np.eye(4) if transforms is None or transforms[i] is None else np.asarray(transforms[i], np.float64)
            for c in range(4):
                for r in range(4):
                    arr[i].transform[c * 4 + r] = float(M[r, c])  # column-major
        self._call(self.lib.vira_gpu_tlas_build, self.h, arr, len(mesh_ids))

    def lights_set(self, lights):
        arr = (capi.Light * max(1, len(lights)))()
        for i, L in enumerate(lights):
            arr[i] = L.to_c()
        self._call(self.lib.vira_gpu_lights_set, self.h, arr, len(lights))

    def material_set(self, mid: int, albedo=None, texture=None, smooth_shading: bool = False, normal_map=None,
                     material_id: int = 0):
        """Material of mesh mid: constant albedo (8 bands), texture (H, W) grey x albedo or
        (H, W, 8) spectral albedo map, normal_map (H, W, 3) tangent-space normals (device
        tensors, kept alive here), material_id for the material_id pass."""
        m = capi.Material()
        a = np.ones(NUM_BINS, np.float32) if albedo is None else np.broadcast_to(np.asarray(albedo, np.float32), (NUM_BINS,))
        for k in range(NUM_BINS):
            m.albedo[k] = float(a[k])
        m.tex_channels = 1
        if texture is not None:
            if not texture.is_contiguous() or texture.dtype != self.torch.float32:
                raise ValueError("texture: contiguous float32 device tensor")
            m.texture_dev = texture.data_ptr()
            m.tex_height, m.tex_width = texture.shape[:2]
            m.tex_channels = 1 if texture.dim() == 2 else int(texture.shape[2])
        if normal_map is not None:
            if not normal_map.is_contiguous() or normal_map.dim() != 3 or normal_map.shape[2] != 3:
                raise ValueError("normal_map: contiguous (H, W, 3) float32 device tensor")
            m.normal_map_dev = normal_map.data_ptr()
            m.nmap_height, m.nmap_width = normal_map.shape[:2]
        m.smooth_shading = int(smooth_shading)
        m.material_id = int(material_id)
        if not hasattr(self, "_materials"):
            self._materials = {}
        self._materials[mid] = (texture, normal_map)  # keep alive
        self._texture = texture
        self._call(self.lib.vira_gpu_material_set, self.h, mid, C.byref(m))

    def environment_set(self, ambient=None, background=None):
        """Scene::setAmbient (8 bands or a scalar; None = none) and setBackgroundEmission
        ((H, W, 8) float32 device tensor, kept alive here; None = zero)."""
        amb = None
        if ambient is not None:
            a = np.broadcast_to(np.asarray(ambient, np.float32), (NUM_BINS,))
            amb = (C.c_float * NUM_BINS)(*[float(v) for v in a])
        bw = bh = 0
        if background is not None:
            if not background.is_contiguous() or background.dim() != 3 or background.shape[2] != NUM_BINS:
                raise ValueError("background: contiguous (H, W, 8) float32 device tensor")
            bh, bw = background.shape[:2]
        self._background = background
        self._call(self.lib.vira_gpu_environment_set, self.h, amb,
                   None if background is None else background.data_ptr(), bw, bh)

    def render(self, camera: Camera, options: PathTracerOptions, passes: RenderPasses):
        st = capi.RenderStats()
        cam = camera.to_c()
        opt = options.to_c()
        out = passes.to_c()
        if options.denoise and any(t is None for t in (passes.direct_radiance, passes.indirect_radiance,
                                                        passes.albedo, passes.depth, passes.normal_global)):
            raise ValueError("denoise needs the direct, indirect, albedo, depth and normal passes")
        self._call(self.lib.vira_gpu_render, self.h, C.byref(cam), C.byref(opt), C.byref(out), C.byref(st))
        self.last_stats = capi.struct_dict(st)
        if options.denoise:  # cpu_path_tracer.ipp:124-134
            self.denoise(passes, options.denoiser)
        return self.last_stats

    def render_noalloc(self, camera: Camera, options: PathTracerOptions, passes: RenderPasses):
        """vira_gpu_render_noalloc: the per-pixel render without any device
        allocation (scratch in the library's fixed per-GPU array; at most 4096
        8x8 tiles, 64 instances).  Same passes as render(); no denoise step."""
        st = capi.RenderStats()
        cam = camera.to_c()
        opt = options.to_c()
        out = passes.to_c()
        self._call(self.lib.vira_gpu_render_noalloc, self.h, C.byref(cam), C.byref(opt), C.byref(out), C.byref(st))
        self.last_stats = capi.struct_dict(st)
        return self.last_stats

    def denoise(self, passes: RenderPasses, options: "DenoiseOptions" = None):
        """denoiseSpectralRadianceEATWT on the passes' direct and indirect radiance (in
        place), then total = direct + indirect when the passes hold a total."""
        o = (options or DenoiseOptions()).to_c()
        need = (passes.direct_radiance, passes.indirect_radiance, passes.albedo, passes.depth, passes.normal_global)
        if any(t is None or not t.is_contiguous() for t in need):
            raise ValueError("denoise needs contiguous direct, indirect, albedo, depth and normal passes")
        tot = passes.total_radiance
        self._call(self.lib.vira_gpu_denoise, self.h, C.byref(o), passes.width, passes.height,
                   passes.direct_radiance.data_ptr(), passes.indirect_radiance.data_ptr(), passes.albedo.data_ptr(),
                   passes.depth.data_ptr(), passes.normal_global.data_ptr(),
                   None if tot is None else tot.data_ptr())

    def trace_rays(self, mid: int, rays, any_hit: bool = False):
        """rays: (n, 8) float32 CUDA tensor (ox oy oz 0 dx dy dz tmax)."""
        n = rays.shape[0]
        dev = rays.device
        t = self.torch.empty(n, dtype=self.torch.float32, device=dev)
        ids = self.torch.empty(n, dtype=self.torch.int32, device=dev)
        b1 = self.torch.empty(n, dtype=self.torch.float32, device=dev)
        b2 = self.torch.empty(n, dtype=self.torch.float32, device=dev)
        ms = C.c_double()
        self._call(self.lib.vira_gpu_trace_rays, self.h, mid, rays.data_ptr(), n, int(any_hit), t.data_ptr(),
                                                ids.data_ptr(), b1.data_ptr(), b2.data_ptr(), C.byref(ms))
        return t, ids, b1, b2, ms.value

    def synchronize(self):
        self._call(self.lib.vira_gpu_synchronize, self.h)

    def received_power(self, camera: Camera, radiance=None, solid_angle: bool = False):
        """Received power (H, W, 8) from a radiance image (Camera::calculateReceivedPower) and,
        if asked, the per-pixel solid angles (H, W)."""
        power = None if radiance is None else self.torch.empty_like(radiance)
        omega = (self.torch.empty((camera.height, camera.width), dtype=self.torch.float32,
                                  device=f"cuda:{self.device}") if solid_angle else None)
        cam = camera.to_c()
        sen = camera.sensor()
        self._call(self.lib.vira_gpu_received_power, self.h, C.byref(cam), C.byref(sen),
                   None if radiance is None else radiance.contiguous().data_ptr(),
                   None if power is None else power.data_ptr(), None if omega is None else omega.data_ptr())
        return power, omega

    # ---- sensor tail
    def psf_kernel(self, psf, size: int):
        """(size, size, 8) kernel of a GaussianPSF / AiryDiskPSF (PointSpreadFunction::makeKernel)."""
        k = self.torch.empty((size, size, NUM_BINS), dtype=self.torch.float32, device=f"cuda:{self.device}")
        c = psf.to_c()
        self._call(self.lib.vira_gpu_psf_kernel, self.h, C.byref(c), size, k.data_ptr())
        return k

    def convolve(self, image, kernel):
        """Image::convolve of an (H, W, 8) image with a (K, K, 8) kernel, zero-padded."""
        img = image.contiguous()
        out = self.torch.empty_like(img)
        H, W = img.shape[:2]
        self._call(self.lib.vira_gpu_convolve, self.h, img.data_ptr(), W, H, kernel.contiguous().data_ptr(),
                   kernel.shape[0], out.data_ptr())
        return out

    def simulate_sensor(self, config, power, photons: bool = False, filter_mosaic=None):
        """Camera::simulateSensor: (H, W, 8) received power -> (H, W) normalised image
        (and the (H, W, 8) photon counts if asked); filter_mosaic: an optional (H, W, 8)
        colour filter array on the device (e.g. sensor.bayer_mosaic)."""
        p = power.contiguous()
        H, W = p.shape[:2]
        img = self.torch.empty((H, W), dtype=self.torch.float32, device=p.device)
        ph = self.torch.empty_like(p) if photons else None
        c = config.to_c()
        if filter_mosaic is not None:
            if tuple(filter_mosaic.shape) != (H, W, NUM_BINS) or not filter_mosaic.is_contiguous():
                raise ValueError("filter_mosaic: contiguous (H, W, 8) float32 device tensor")
            c.filter_mosaic_dev = filter_mosaic.data_ptr()
        self._call(self.lib.vira_gpu_simulate_sensor, self.h, C.byref(c), p.data_ptr(), W, H,
                   None if ph is None else ph.data_ptr(), img.data_ptr())
        return img, ph

    def star_psf(self, psf, camera: "Camera", sizes=(3, 9, 27, 81), cull_behind: bool = False):
        """PSF kernels of ascending sizes (PointSpreadFunction::initKernels default,
        psf.ipp) with their edge maxima, the minimum detectable power and K, for
        render_stars.  cull_behind=False projects sources behind the camera as
        the reference does.  psf=None: a camera without PSF, each source's power
        assigned to one pixel."""
        from . import sensor as SN
        sp = capi.StarPSF()
        if psf is None:
            sizes = ()
        sp.count = len(sizes)
        kernels = []
        for i, K in enumerate(sizes):
            k = self.psf_kernel(psf, K)
            kernels.append(k)
            sp.size[i] = K
            sp.kernel_dev[i] = k.data_ptr()
            em = SN.edge_max(k.cpu().numpy())
            for b in range(NUM_BINS):
                sp.edge_max[i][b] = float(em[b])
        sp.minimum_power = SN.minimum_detectable_power(camera.aperture_area(), camera.exposure_time)
        sp.cull_behind = int(bool(cull_behind))
        K = SN.intrinsic_matrix(camera)
        for i in range(6):
            sp.intrinsic[i] = float(K[i])
        sp._keep = kernels
        return sp

    def render_stars(self, camera: "Camera", star_psf, directions, irradiance, out=None):
        """Unresolved point sources (unit directions (n, 3), irradiance (n, 8)) added into
        an (H, W, 8) received-power image (zeros if out is None)."""
        if out is None:
            out = self.torch.zeros((camera.height, camera.width, NUM_BINS), dtype=self.torch.float32,
                                   device=f"cuda:{self.device}")
        d = directions.contiguous()
        e = irradiance.contiguous()
        n = d.shape[0]
        for name, t, shape in (("directions", d, (n, 3)), ("irradiance", e, (n, NUM_BINS)),
                               ("out", out, (camera.height, camera.width, NUM_BINS))):
            if tuple(t.shape) != shape or t.dtype != self.torch.float32 or not t.is_cuda:
                raise ValueError(f"render_stars: {name} must be a float32 device tensor of shape {shape}")
        if not out.is_contiguous():
            raise ValueError("render_stars: out must be contiguous")
        cam = camera.to_c()
        sen = camera.sensor()
        self._call(self.lib.vira_gpu_render_stars, self.h, C.byref(cam), C.byref(sen), C.byref(star_psf),
                   d.data_ptr(), e.data_ptr(), n, out.data_ptr())
        return out

    def star_light(self, stars, et: float = 0.0):
        """StarCatalogue::makeStarLight (star_catalogue.ipp:55-68) on the device:
        stars a uint8 device tensor holding n 64-byte star quipu records (n x 64),
        et seconds past J2000.  Returns (directions (n, 3), irradiance (n, 8)),
        float32 device tensors for render_stars."""
        if stars.dtype != self.torch.uint8 or not stars.is_cuda or stars.dim() != 2 or stars.shape[1] != 64 \
                or not stars.is_contiguous():
            raise ValueError("star_light: stars must be a contiguous (n, 64) uint8 device tensor of star records")
        n = stars.shape[0]
        dev = f"cuda:{self.device}"
        d = self.torch.empty((n, 3), dtype=self.torch.float32, device=dev)
        e = self.torch.empty((n, NUM_BINS), dtype=self.torch.float32, device=dev)
        self._call(self.lib.vira_gpu_star_light, self.h, stars.data_ptr(), n, float(et), d.data_ptr(), e.data_ptr())
        return d, e

    # ---- batched rendering
    @staticmethod
    def _requests(requests):
        arr = (capi.Request * len(requests))()
        for i, r in enumerate(requests):
            arr[i] = r.to_c()
        return arr

    def batch_submit(self, requests, config: "BatchConfig | None" = None):
        """Renders every RenderRequest in one persistent launch; returns the launch stats.
        Requests whose options ask for denoising are denoised afterwards, one by one."""
        arr = self._requests(requests)
        st = capi.BatchStats()
        cfg = None if config is None else C.byref(config.to_c())
        self._call(self.lib.vira_gpu_batch_submit, self.h, arr, len(requests), cfg, C.byref(st))
        self.last_stats = {f: getattr(st, f) for f, _ in capi.BatchStats._fields_}
        for r in requests:
            if r.options.denoise:
                self.denoise(r.passes, r.options.denoiser)
        return self.last_stats

    def batch_autotune(self, requests, sample_stride: int = 16):
        """Picks and keeps the fastest batch configuration for these requests."""
        arr = self._requests(requests)
        best = capi.BatchConfig()
        ms = C.c_double()
        self._call(self.lib.vira_gpu_batch_autotune, self.h, arr, len(requests), sample_stride, C.byref(best),
                   C.byref(ms))
        return BatchConfig(best.bvh_width, best.waves_per_block, best.waves_per_cu, best.tiles_per_grab,
                           best.order), ms.value

    # ---- differentiable direct illumination
    def dem_update(self, mid: int, heights, refit: bool = False):
        """New heights (H, W) for a DEM mesh; rebuilds its BLAS if one was built,
        or refits it (same tree, new boxes) with refit=True."""
        fn = self.lib.vira_gpu_dem_update_refit if refit else self.lib.vira_gpu_dem_update
        self._call(fn, self.h, mid, heights.contiguous().data_ptr())

    def blas_refit(self, mid: int):
        """Refits a built BLAS to the mesh's current vertices (same tree, new
        boxes); returns the BLAS statistics (build_ms = the refit's time)."""
        self._call(self.lib.vira_gpu_blas_refit, self.h, mid)
        s = capi.BVHStats()
        self._call(self.lib.vira_gpu_blas_stats, self.h, mid, C.byref(s))
        return {f: getattr(s, f) for f, _ in capi.BVHStats._fields_}

    def di_render(self, camera: Camera, options: DIOptions, radiance=None):
        """Radiance (H, W, 8) of the scene mesh under point lights, bounce 0 only."""
        if radiance is None:
            radiance = self.torch.empty((camera.height, camera.width, NUM_BINS), dtype=self.torch.float32,
                                        device=f"cuda:{self.device}")
        st = capi.DIStats()
        cam = camera.to_c()
        opt = options.to_c()
        self._call(self.lib.vira_gpu_di_render, self.h, C.byref(cam), C.byref(opt), radiance.data_ptr(),
                                               C.byref(st))
        self.last_stats = {f: getattr(st, f) for f, _ in capi.DIStats._fields_}
        return radiance

    def adjoint(self, camera: Camera, options: DIOptions, grad_image, grad_heights=None, grad_texture=None):
        """dL/dheights and dL/dtexels from dL/dimage (H, W, 8); pass output tensors to fill (or None)."""
        st = capi.DIStats()
        cam = camera.to_c()
        opt = options.to_c()
        g = grad_image.contiguous()
        self._call(self.lib.vira_gpu_adjoint, self.h, C.byref(cam), C.byref(opt), g.data_ptr(),
                                             None if grad_heights is None else grad_heights.data_ptr(),
                                             None if grad_texture is None else grad_texture.data_ptr(),
                                             C.byref(st))
        self.last_stats = {f: getattr(st, f) for f, _ in capi.DIStats._fields_}
        return self.last_stats


def projection(kind: int = 0, spacing: float = 30.0, radius: float = 0.0, center=(0.0, 0.0, 0.0),
               offset=(0.0, 0.0, 0.0)) -> capi.Projection:
    p = capi.Projection()
    p.kind, p.spacing, p.radius = kind, spacing, radius
    for k in range(3):
        p.center[k] = float(center[k])
        p.offset[k] = float(offset[k])
    return p

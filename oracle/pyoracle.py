"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the measured or
shipped path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "liboracle.so"
NB = 8


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


class Projection(C.Structure):
    _fields_ = [("kind", C.c_int), ("spacing", C.c_double), ("radius", C.c_double),
                ("center", C.c_double * 3), ("offset", C.c_double * 3)]


class Camera(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("kinv", C.c_float * 6), ("rot", C.c_float * 9),
                ("pos", C.c_double * 3), ("zdir", C.c_float), ("view_normal", C.c_float * 9)]


class Light(C.Structure):
    _fields_ = [("kind", C.c_int), ("pos", C.c_double * 3), ("radius", C.c_float), ("value", C.c_float * NB)]


class Material(C.Structure):
    _fields_ = [("albedo", C.c_float * NB), ("tex_w", C.c_int), ("tex_h", C.c_int),
                ("tex", C.c_void_p), ("smooth", C.c_int), ("tex_channels", C.c_int), ("nmap", C.c_void_p),
                ("nmap_w", C.c_int), ("nmap_h", C.c_int), ("material_id", C.c_uint32)]


class PSF(C.Structure):
    _fields_ = [("kind", C.c_int), ("sigma_x", C.c_float * NB), ("sigma_y", C.c_float * NB),
                ("angle_deg", C.c_float), ("focal_length", C.c_float), ("aperture_diameter", C.c_float),
                ("pixel_size", C.c_float * 2), ("wavelength", C.c_float * NB), ("supersample", C.c_int)]


class DenoiseOpts(C.Structure):
    _fields_ = [("levels_direct", C.c_int), ("levels_indirect", C.c_int), ("depth_threshold", C.c_float),
                ("normal_threshold", C.c_float), ("min_normal_weight", C.c_float), ("eps", C.c_float)]


class Options(C.Structure):
    _fields_ = [("spp", C.c_int), ("bounces", C.c_int), ("rr", C.c_int), ("rr_start", C.c_int),
                ("seed", C.c_uint32), ("show_background", C.c_int), ("adaptive", C.c_int),
                ("samples_per_batch", C.c_int), ("sampling_tolerance", C.c_float),
                ("samples_to_detect_miss", C.c_int), ("ambient", C.c_float * NB), ("has_ambient", C.c_int),
                ("background", C.c_void_p), ("bg_w", C.c_int), ("bg_h", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    L = C.CDLL(str(LIB))
    P = C.c_void_p
    sig = {
        "orc_fbm_heights": (None, [C.c_int, C.c_int, C.c_uint32, C.c_double, C.c_int, C.c_double, P, C.c_int]),
        "orc_vertex_albedo": (None, [C.c_size_t, C.c_uint32, P]),
        "orc_dem_indices": (C.c_int64, [P, C.c_int, C.c_int, P]),
        "orc_dem_positions": (None, [P, C.c_int, C.c_int, C.POINTER(Projection), P, P]),
        "orc_vertex_normals": (None, [P, P, C.c_size_t, P, C.c_size_t, P]),
        "orc_mesh_create": (P, [P, P, C.c_size_t, P, C.c_size_t, P, P, P, C.c_int]),
        "orc_mesh_destroy": (None, [P]),
        "orc_mesh_build_ms": (C.c_double, [P]),
        "orc_bvh_num_nodes": (C.c_size_t, [P]),
        "orc_bvh_check": (C.c_int, [P, P, P]),
        "orc_trace": (None, [P, P, C.c_size_t, C.c_int, P, C.c_int, P, P, P, P, P, C.c_int]),
        "orc_trace64": (None, [P, P, C.c_size_t, C.c_int, P, P, P, P, P, C.c_int]),
        "orc_intersect_tri": (C.c_int, [P, P, P, P, P]),
        "orc_lowbias32": (C.c_uint32, [C.c_uint32]),
        "orc_bvh_nodes": (None, [P, P, P, P]),
        "orc_bvh_tri_index": (None, [P, P]),
        "orc_offset_point": (None, [P, P, P]),
        "orc_rng_u01": (C.c_float, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
        "orc_render": (None, [P, C.c_int, C.POINTER(Camera), C.POINTER(Light), C.c_int, C.POINTER(Material),
                              C.POINTER(Options), P, P, P, P, P, P, P, P, P, P, C.c_int, P, C.c_size_t, P, P, P]),
        "orc_render_scene": (None, [P, P, P, C.c_int, C.c_int, C.POINTER(Camera), C.POINTER(Light), C.c_int,
                                    C.POINTER(Options), P, P, P, P, P, P, P, P, P, P, C.c_int, P, C.c_size_t,
                                    P, P, P]),
        "orc_mesh_set_id": (None, [P, C.c_uint32]),
        "orc_atan2_f32": (None, [P, P, C.c_size_t, P]),
        "orc_plant": (None, [C.c_int]),
        "orc_pixel_solid_angles": (None, [P, C.c_float, C.c_int, C.c_int, P]),
        "orc_psf_kernel": (None, [C.POINTER(PSF), C.c_int, P]),
        "orc_convolve": (None, [P, C.c_int, C.c_int, P, C.c_int, P]),
        "orc_denoise": (None, [C.POINTER(DenoiseOpts), C.c_int, C.c_int, P, P, P, P, P, C.c_int]),
        "orc_render_stars": (None, [P, C.c_int, C.c_int, P, C.c_float, C.c_int, P, P, C.c_float, P, C.c_int,
                                    P, P, C.c_size_t, P]),
        "orc_dem_diff_create": (P, [C.c_int, C.c_int, C.POINTER(Projection), P, P]),
        "orc_dem_diff_destroy": (None, [P]),
        "orc_di_forward64": (None, [P, P, P, P, C.c_size_t, C.c_int, C.POINTER(Light), C.c_int,
                                    C.POINTER(Material), P, C.c_int]),
        "orc_di_adjoint64": (None, [P, P, P, P, C.c_size_t, C.c_int, C.POINTER(Light), C.c_int,
                                    C.POINTER(Material), P, P, P]),
        "orc_sincos_f32": (None, [P, C.c_size_t, P, P]),
        "orc_light_sample": (None, [C.c_int, C.POINTER(Light), P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_size_t,
                                    P, P, P, P]),
        "orc_light_pdf": (None, [C.c_int, C.POINTER(Light), P, P, C.c_size_t, P]),
        "orc_bounce_dirs": (None, [C.c_int, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_size_t, P, P]),
        "orc_tex_bilinear": (None, [C.c_int, C.POINTER(Material), P, C.c_size_t, P]),
        "orc_kahan_sum_f32": (C.c_float, [P, C.c_size_t]),
        "orc_di_fd64": (None, [P, P, P, P, C.c_size_t, C.c_int, C.POINTER(Light), C.c_int, C.POINTER(Material),
                               P, P, C.c_size_t, C.c_double, P, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def nthreads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def fbm_heights(W, H, seed, amplitude=500.0, octaves=6, base_cells=64.0) -> np.ndarray:
    out = np.empty((H, W), np.float32)
    lib().orc_fbm_heights(W, H, seed, amplitude, octaves, base_cells, out.ctypes.data, nthreads())
    return out


def vertex_albedo(nv, seed) -> np.ndarray:
    out = np.empty(nv, np.float32)
    lib().orc_vertex_albedo(nv, seed, out.ctypes.data)
    return out


def dem_indices(heights: np.ndarray) -> np.ndarray:
    h = np.ascontiguousarray(heights, np.float32)
    H, W = h.shape
    out = np.empty((2 * (W - 1) * (H - 1), 3), np.uint32)
    nt = lib().orc_dem_indices(h.ctypes.data, W, H, out.ctypes.data)
    return out[:nt].copy()


def make_projection(kind=0, spacing=30.0, radius=0.0, center=(0, 0, 0), offset=(0, 0, 0)) -> Projection:
    p = Projection()
    p.kind, p.spacing, p.radius = kind, spacing, radius
    for k in range(3):
        p.center[k], p.offset[k] = center[k], offset[k]
    return p


def dem_positions(heights: np.ndarray, proj: Projection):
    h = np.ascontiguousarray(heights, np.float32)
    H, W = h.shape
    pos = np.empty((H * W, 3), np.float32)
    pos64 = np.empty((H * W, 3), np.float64)
    lib().orc_dem_positions(h.ctypes.data, W, H, C.byref(proj), pos.ctypes.data, pos64.ctypes.data)
    return pos, pos64


def dem_uv(W, H) -> np.ndarray:
    i = np.arange(W, dtype=np.float64)[None, :].repeat(H, 0)
    j = np.arange(H, dtype=np.float64)[:, None].repeat(W, 1)
    return np.stack([(i / (W - 1)).astype(np.float32), (j / (H - 1)).astype(np.float32)], -1).reshape(-1, 2)


def vertex_normals(pos: np.ndarray, idx: np.ndarray, pos64: np.ndarray | None = None) -> np.ndarray:
    """Mesh::calculateNormals over face normals evaluated in FP64 from pos64 (the
    FP64 positions the mesh was built from) or, without it, from pos widened."""
    pos = np.ascontiguousarray(pos, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    p64 = None if pos64 is None else np.ascontiguousarray(pos64, np.float64)
    out = np.empty_like(pos)
    lib().orc_vertex_normals(pos.ctypes.data, _p(p64), pos.shape[0], idx.ctypes.data, idx.size // 3, out.ctypes.data)
    return out


class Mesh:
    """Mesh + reference binned-SAH BVH2 on the CPU."""

    def __init__(self, pos, idx, nrm=None, alb=None, uv=None, pos64=None, mesh_id=0):
        self.pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        self.idx = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
        self.nrm = None if nrm is None else np.ascontiguousarray(nrm, np.float32)
        self.alb = None if alb is None else np.ascontiguousarray(alb, np.float32)
        self.uv = None if uv is None else np.ascontiguousarray(uv, np.float32)
        self.pos64 = None if pos64 is None else np.ascontiguousarray(pos64, np.float64)
        self.h = lib().orc_mesh_create(_p(self.pos), _p(self.pos64), self.pos.shape[0], _p(self.idx),
                                       self.idx.shape[0], _p(self.nrm), _p(self.alb), _p(self.uv), nthreads())
        lib().orc_mesh_set_id(self.h, int(mesh_id))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_destroy(self.h)
            self.h = None

    @property
    def build_ms(self):
        return lib().orc_mesh_build_ms(self.h)

    def check(self):
        md, ml = C.c_uint32(), C.c_uint32()
        code = lib().orc_bvh_check(self.h, C.byref(md), C.byref(ml))
        return code, md.value, ml.value

    def nodes(self):
        n = lib().orc_bvh_num_nodes(self.h)
        lohi = np.empty((n, 6), np.float64)
        left = np.empty(n, np.uint32)
        cnt = np.empty(n, np.uint32)
        lib().orc_bvh_nodes(self.h, lohi.ctypes.data, left.ctypes.data, cnt.ctypes.data)
        tri = np.empty(self.idx.shape[0], np.uint32)
        lib().orc_bvh_tri_index(self.h, tri.ctypes.data)
        return lohi, left, cnt, tri

    def trace(self, rays: np.ndarray, any_hit=False, tmax=None, brute=False):
        """rays: (n, 6) float32.  Returns t, id, b1, b2, (internal visits, tri tests)."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        t = np.empty(n, np.float32)
        ids = np.empty(n, np.uint32)
        b1 = np.empty(n, np.float32)
        b2 = np.empty(n, np.float32)
        cnt = np.zeros(2, np.uint64)
        tm = None if tmax is None else np.ascontiguousarray(tmax, np.float32)
        lib().orc_trace(self.h, rays.ctypes.data, n, int(any_hit), _p(tm), int(brute), t.ctypes.data,
                        ids.ctypes.data, b1.ctypes.data, b2.ctypes.data, cnt.ctypes.data, nthreads())
        return t, ids, b1, b2, (int(cnt[0]), int(cnt[1]))

    def trace64(self, rays: np.ndarray, any_hit=False, tmax=None):
        rays = np.ascontiguousarray(rays, np.float64)
        n = rays.shape[0]
        t = np.empty(n, np.float64)
        ids = np.empty(n, np.uint32)
        b1 = np.empty(n, np.float64)
        b2 = np.empty(n, np.float64)
        tm = None if tmax is None else np.ascontiguousarray(tmax, np.float64)
        lib().orc_trace64(self.h, rays.ctypes.data, n, int(any_hit), _p(tm), t.ctypes.data, ids.ctypes.data,
                          b1.ctypes.data, b2.ctypes.data, nthreads())
        return t, ids, b1, b2

    def render(self, cam: Camera, lights, mat: Material, opt: Options, precision=32, pixels=None,
               threads=None, records=False):
        W, H = cam.W, cam.H
        n = W * H
        out = {
            "radiance": np.zeros((n, NB), np.float32), "direct": np.zeros((n, NB), np.float32),
            "indirect": np.zeros((n, NB), np.float32), "albedo": np.zeros((n, NB), np.float32),
            "normal": np.zeros((n, 3), np.float32), "depth": np.zeros(n, np.float32),
            "alpha": np.zeros(n, np.float32), "tri_id": np.full(n, 0xFFFFFFFF, np.uint32),
            "normal_camera": np.zeros((n, 3), np.float32), "mesh_id": np.zeros(n, np.uint32),
            "material_id": np.zeros(n, np.uint32),
        }
        cnt = np.zeros(7, np.uint64)
        rec = np.zeros((opt.spp, n, 4), np.uint32) if records else None
        larr = (Light * max(1, len(lights)))(*lights) if lights else (Light * 1)()
        pl = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
        lib().orc_render(self.h, precision, C.byref(cam), larr, len(lights), C.byref(mat), C.byref(opt),
                         *(out[k].ctypes.data for k in ("radiance", "direct", "indirect", "albedo", "normal",
                                                        "depth", "alpha", "tri_id")),
                         cnt.ctypes.data, _p(rec), threads or nthreads(), _p(pl), 0 if pl is None else pl.size,
                         *(out[k].ctypes.data for k in ("normal_camera", "mesh_id", "material_id")))
        if rec is not None:
            out["records"] = rec
        out["counts"] = {"primary": int(cnt[0]), "shadow": int(cnt[1]), "bounce": int(cnt[2]),
                         "closest_internal": int(cnt[3]), "closest_tris": int(cnt[4]),
                         "any_internal": int(cnt[5]), "any_tris": int(cnt[6])}
        return out


# Planted mistakes (the oracle's PLANT_* codes): test instrument for showing that each pin catches one.
PLANTS = {"mis_balance": 1, "drop_cosine": 2, "cone_omega": 3, "cone_square": 4, "bounce_pdf": 5,
          "rr_no_reweight": 6, "tex_off_by_one": 7, "tex_transposed": 8, "kahan_off": 9, "normal_matrix": 10,
          "background_swap": 11, "adaptive_divisor": 12, "nmap_unnormalised": 13, "ambient_no_albedo": 14}


def plant(name: str | None):
    """Plants the named mistake in the oracle (None removes it)."""
    lib().orc_plant(C.c_int(0 if name is None else PLANTS[name]))


def atan2_f32(y, x):
    """The kernels' FP32 atan2 (vira_amd/csrc/fp_trig.hpp)."""
    y = np.ascontiguousarray(y, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(y)
    lib().orc_atan2_f32(y.ctypes.data, x.ctypes.data, y.size, out.ctypes.data)
    return out


def sincos_f32(x):
    """The kernels' FP32 sin/cos routine (vira_amd/csrc/fp_trig.hpp) on an array."""
    x = np.ascontiguousarray(x, np.float32)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().orc_sincos_f32(x.ctypes.data, x.size, s.ctypes.data, c.ctypes.data)
    return s, c


def light_sample(L: Light, p, n: int, seed=0, pix=0, dim0=16, precision=64):
    """Light::sample at point p for samples 0..n-1: (dirs (n,3), distance (n), pdf (n), radiance (n,8))."""
    pp = np.ascontiguousarray(p, np.float64)
    dirs = np.empty((n, 3)); dist = np.empty(n); pdf = np.empty(n); rad = np.empty((n, NB))
    lib().orc_light_sample(precision, C.byref(L), pp.ctypes.data, seed, pix, dim0, n, dirs.ctypes.data,
                           dist.ctypes.data, pdf.ctypes.data, rad.ctypes.data)
    return dirs, dist, pdf, rad


def light_pdf(L: Light, p, dirs, precision=64):
    pp = np.ascontiguousarray(p, np.float64)
    d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    out = np.empty(d.shape[0])
    lib().orc_light_pdf(precision, C.byref(L), pp.ctypes.data, d.ctypes.data, d.shape[0], out.ctypes.data)
    return out


def bounce_dirs(N, n: int, seed=0, pix=0, bounce=0, precision=64):
    """Cosine-weighted bounce directions (n,3) and pdfs (n) around normal N, samples 0..n-1."""
    NN = np.ascontiguousarray(N, np.float64)
    dirs = np.empty((n, 3)); pdf = np.empty(n)
    lib().orc_bounce_dirs(precision, NN.ctypes.data, seed, pix, bounce, n, dirs.ctypes.data, pdf.ctypes.data)
    return dirs, pdf


def tex_bilinear(texture: np.ndarray, uv, precision=32):
    """Image::sampleUVs of a grey texture (H, W) at uv (n, 2)."""
    m = material(1.0, texture)
    u = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
    out = np.empty(u.shape[0])
    lib().orc_tex_bilinear(precision, C.byref(m), u.ctypes.data, u.shape[0], out.ctypes.data)
    return out


def kahan_sum_f32(x) -> float:
    x = np.ascontiguousarray(x, np.float32)
    return float(lib().orc_kahan_sum_f32(x.ctypes.data, x.size))


def offset_point(p, n):
    p = np.ascontiguousarray(p, np.float32)
    n = np.ascontiguousarray(n, np.float32)
    out = np.empty(3, np.float32)
    lib().orc_offset_point(p.ctypes.data, n.ctypes.data, out.ctypes.data)
    return out


def camera_from(c) -> Camera:
    """Copies a vira_amd camera record (same numbers) into the oracle's struct."""
    o = Camera()
    o.W, o.H = c.width, c.height
    for k in range(6):
        o.kinv[k] = c.kinv[k]
    for k in range(9):
        o.rot[k] = c.rot[k]
    for k in range(3):
        o.pos[k] = c.position[k]
    o.zdir = c.z_dir
    vn = getattr(c, "view_normal", None)
    if vn is not None:
        for k in range(9):
            o.view_normal[k] = vn[k]
    return o


def light_from(L) -> Light:
    o = Light()
    o.kind = L.kind
    for k in range(3):
        o.pos[k] = L.position[k]
    o.radius = L.radius
    for k in range(NB):
        o.value[k] = L.value[k]
    return o


def material(albedo=1.0, texture: np.ndarray | None = None, smooth=False, normal_map: np.ndarray | None = None,
             material_id: int = 0) -> Material:
    """texture (H, W) grey x albedo, or (H, W, 8) spectral albedo map; normal_map (H, W, 3) tangent space."""
    m = Material()
    a = np.broadcast_to(np.asarray(albedo, np.float32), (NB,))
    for k in range(NB):
        m.albedo[k] = float(a[k])
    keep = []
    m.tex_channels = 1
    if texture is not None:
        texture = np.ascontiguousarray(texture, np.float32)
        m.tex_h, m.tex_w = texture.shape[:2]
        m.tex_channels = 1 if texture.ndim == 2 else texture.shape[2]
        m.tex = texture.ctypes.data
        keep.append(texture)
    if normal_map is not None:
        normal_map = np.ascontiguousarray(normal_map, np.float32)
        m.nmap_h, m.nmap_w = normal_map.shape[:2]
        m.nmap = normal_map.ctypes.data
        keep.append(normal_map)
    m._keep = keep
    m.smooth = int(smooth)
    m.material_id = int(material_id)
    return m


def options(spp=1, bounces=0, rr=False, rr_start=1, seed=0, show_background=False, adaptive=False,
            samples_per_batch=30, sampling_tolerance=0.05, samples_to_detect_miss=30, ambient=None,
            background: np.ndarray | None = None) -> Options:
    """PathTracer options plus the scene's ambient light and background image ((H, W, 8))."""
    o = Options()
    o.spp, o.bounces, o.rr, o.rr_start, o.seed = spp, bounces, int(rr), rr_start, seed
    o.show_background, o.adaptive = int(show_background), int(adaptive)
    o.samples_per_batch, o.sampling_tolerance, o.samples_to_detect_miss = samples_per_batch, sampling_tolerance, \
        samples_to_detect_miss
    if ambient is not None:
        a = np.broadcast_to(np.asarray(ambient, np.float32), (NB,))
        for k in range(NB):
            o.ambient[k] = float(a[k])
        o.has_ambient = int(np.any(a != 0))
    if background is not None:
        bg = np.ascontiguousarray(background, np.float32)
        o.bg_h, o.bg_w = bg.shape[:2]
        o.background = bg.ctypes.data
        o._keep = bg
    return o


def dem_mesh(heights: np.ndarray, proj: Projection, grey_albedo: np.ndarray | None = None,
             with_pos64=True) -> Mesh:
    """DEM -> mesh exactly as the oracle restates dem.ipp / image_utils.ipp / mesh.ipp.
    The mesh keeps its FP64 positions (face normals are evaluated from them, and
    FP64 renders trace them); with_pos64 is accepted for older callers."""
    H, W = heights.shape
    pos, pos64 = dem_positions(heights, proj)
    idx = dem_indices(heights)
    nrm = vertex_normals(pos, idx, pos64)
    a = np.ones(H * W, np.float32) if grey_albedo is None else np.asarray(grey_albedo, np.float32)
    alb = np.repeat(a[:, None], NB, axis=1)
    return Mesh(pos, idx, nrm, alb, dem_uv(W, H), pos64)


def pixel_solid_angles(cam: Camera) -> np.ndarray:
    """(H, W) FP32 solid angle of every pixel (camera.ipp:1594-1666)."""
    k = np.array([cam.kinv[i] for i in range(6)], np.float32)
    out = np.empty((cam.H, cam.W), np.float32)
    lib().orc_pixel_solid_angles(k.ctypes.data, cam.zdir, cam.W, cam.H, out.ctypes.data)
    return out


def psf_from(c) -> PSF:
    """Copy of a library PSF struct (same field layout)."""
    p = PSF()
    C.memmove(C.byref(p), C.byref(c), C.sizeof(PSF))
    return p


def psf_kernel(p: PSF, K: int) -> np.ndarray:
    out = np.empty((K, K, NB), np.float32)
    lib().orc_psf_kernel(C.byref(p), K, out.ctypes.data)
    return out


def convolve(img: np.ndarray, kern: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img, np.float32)
    kern = np.ascontiguousarray(kern, np.float32)
    H, W = img.shape[:2]
    out = np.empty_like(img)
    lib().orc_convolve(img.ctypes.data, W, H, kern.ctypes.data, kern.shape[0], out.ctypes.data)
    return out


def denoise(direct, indirect, albedo, depth, normal, levels=(5, 5), depth_threshold=0.2, normal_threshold=0.5,
            min_normal_weight=0.3, eps=1e-6, literal=False):
    """EATWT denoiser on (H, W, NB) direct/indirect/albedo, (H, W) depth, (H, W, 3) normal.
    Returns (direct, indirect) denoised (new arrays)."""
    d = np.ascontiguousarray(direct, np.float32).copy()
    i = np.ascontiguousarray(indirect, np.float32).copy()
    a = np.ascontiguousarray(albedo, np.float32)
    z = np.ascontiguousarray(depth, np.float32)
    n = np.ascontiguousarray(normal, np.float32)
    H, W = d.shape[:2]
    o = DenoiseOpts(int(levels[0]), int(levels[1]), depth_threshold, normal_threshold, min_normal_weight, eps)
    if literal and W != H:
        raise ValueError("the reference's indexing is defined for square images only")
    lib().orc_denoise(C.byref(o), W, H, d.ctypes.data, i.ctypes.data, a.ctypes.data, z.ctypes.data, n.ctypes.data,
                      int(bool(literal)))
    return d, i


SECONDS_PER_YEAR = 365.25 * 86400.0   # math.hpp:53


def band_edges_m() -> np.ndarray:
    """The 8 bands' [min, max] wavelengths in metres as the spectrum holds them:
    FP32 nm bounds 400 + 75 i scaled by 1e-9f (spectral_data.ipp:64-125), read
    back as FP64 by blackBodyRadiance (spectral_data.ipp:617-619)."""
    nm = np.float32(400.0) + np.arange(NB + 1, dtype=np.float32) * np.float32(75.0)
    return (nm * np.float32(1e-9)).astype(np.float64)


def planck(temperature, lam):
    """plancksLaw (math.ipp:45-59), FP64, spectral radiance per metre."""
    c, h, kB = 299792458.0, 6.62607015e-34, 1.380649e-23
    coeff = 2 * h * c * c / lam ** 5.0
    return coeff * 1.0 / (np.exp(h * c / (lam * kB * temperature)) - 1.0)


def black_body_bands(temperature, steps: int = 100) -> np.ndarray:
    """blackBodyRadiance (spectral_data.ipp:613-629) for an array of temperatures:
    per band, Planck at linspace(min, max, steps) (math.ipp:33-42) integrated by
    trapezoids summed in order (math.ipp:131-172), rounded to FP32.  (n, 8)."""
    T = np.atleast_1d(np.asarray(temperature, np.float64))[:, None]
    e = band_edges_m()
    out = np.zeros((T.shape[0], NB), np.float32)
    for k in range(NB):
        lo, hi = e[k], e[k + 1]
        step = (hi - lo) / float(steps - 1)
        x = lo + np.arange(steps, dtype=np.float64) * step
        y = planck(T, x[None, :])
        area = np.zeros(T.shape[0])
        for i in range(steps - 1):
            area = area + (x[i + 1] - x[i]) * (y[:, i] + y[:, i + 1]) / 2.0
        out[:, k] = area.astype(np.float32)
    return out


def star_light(stars, et: float):
    """StarCatalogue::makeStarLight (star_catalogue.ipp:55-68) over STAR_DTYPE
    records: unit vectors per Star::getUnitVector (star.ipp:19-32; FP32, the
    right-ascension rate added to the declination and the declination rate to
    the right ascension, as the reference does) and irradiance per
    Star::getIrradiance (star.ipp:13-16; (float)omega times the FP32 band
    radiance).  Returns (directions (n, 3), irradiance (n, 8)), float32."""
    dt = float(et) / SECONDS_PER_YEAR
    delta = (stars["de"] + stars["pm_ra"] * dt).astype(np.float32)
    alpha = (stars["ra"] + stars["pm_de"] * dt).astype(np.float32)
    d = np.stack([np.cos(delta) * np.cos(alpha), np.cos(delta) * np.sin(alpha), np.sin(delta)], axis=1)
    irr = stars["omega"].astype(np.float32)[:, None] * black_body_bands(stars["temperature"], 100)
    return d.astype(np.float32), irr.astype(np.float32)


def render_stars(rot, width: int, height: int, eta, area: float, kernels, minimum_power: float, intrinsic,
                 directions, irradiance, cull_behind: bool = False, img=None) -> np.ndarray:
    """Unresolved point sources: rot (3x3, columns = camera axes), kernels a list of
    (K, K, NB) arrays indexed [y][x] in ascending size, intrinsic the 6 entries of K
    ([col][row]), directions (n, 3), irradiance (n, NB).  Returns (H, W, NB)."""
    rot = np.ascontiguousarray(np.asarray(rot, np.float32).reshape(3, 3))
    eta = np.ascontiguousarray(np.broadcast_to(np.asarray(eta, np.float32), (NB,)))
    ks = [np.ascontiguousarray(k, np.float32) for k in kernels]
    sizes = np.array([k.shape[0] for k in ks], np.int32)
    ptrs = (C.c_void_p * len(ks))(*[k.ctypes.data for k in ks])
    intr = np.ascontiguousarray(intrinsic, np.float32)
    d = np.ascontiguousarray(directions, np.float32)
    e = np.ascontiguousarray(irradiance, np.float32)
    out = np.zeros((height, width, NB), np.float32) if img is None else np.ascontiguousarray(img, np.float32).copy()
    lib().orc_render_stars(rot.ctypes.data, width, height, eta.ctypes.data, float(area), len(ks),
                           sizes.ctypes.data, C.cast(ptrs, C.c_void_p), float(minimum_power), intr.ctypes.data,
                           int(bool(cull_behind)), d.ctypes.data, e.ctypes.data, d.shape[0], out.ctypes.data)
    return out


def render_scene(meshes, transforms, mats, cam: Camera, lights, opt: Options, precision=32, pixels=None,
                 threads=None):
    """Two-level scene: meshes (Mesh), transforms (4x4 FP64, row-major numpy), mats (Material)."""
    n = len(meshes)
    W, H = cam.W, cam.H
    npx = W * H
    out = {
        "radiance": np.zeros((npx, NB), np.float32), "direct": np.zeros((npx, NB), np.float32),
        "indirect": np.zeros((npx, NB), np.float32), "albedo": np.zeros((npx, NB), np.float32),
        "normal": np.zeros((npx, 3), np.float32), "depth": np.zeros(npx, np.float32),
        "alpha": np.zeros(npx, np.float32), "tri_id": np.full(npx, 0xFFFFFFFF, np.uint32),
        "instance_id": np.full(npx, 0xFFFFFFFF, np.uint32),
        "normal_camera": np.zeros((npx, 3), np.float32), "mesh_id": np.zeros(npx, np.uint32),
        "material_id": np.zeros(npx, np.uint32),
    }
    cnt = np.zeros(7, np.uint64)
    handles = (C.c_void_p * n)(*[m.h for m in meshes])
    tr = np.ascontiguousarray(np.stack([np.asarray(t, np.float64).T.reshape(-1) for t in transforms]))  # column-major
    marr = (Material * n)(*mats)
    larr = (Light * max(1, len(lights)))(*lights) if lights else (Light * 1)()
    pl = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
    lib().orc_render_scene(handles, tr.ctypes.data, marr, n, precision, C.byref(cam), larr, len(lights),
                           C.byref(opt), *(out[k].ctypes.data for k in ("radiance", "direct", "indirect", "albedo",
                                                                        "normal", "depth", "alpha", "tri_id",
                                                                        "instance_id")),
                           cnt.ctypes.data, threads or nthreads(), _p(pl), 0 if pl is None else pl.size,
                           *(out[k].ctypes.data for k in ("normal_camera", "mesh_id", "material_id")))
    out["counts"] = {"primary": int(cnt[0]), "shadow": int(cnt[1]), "bounce": int(cnt[2])}
    return out


class DemDiff:
    """FP64 fixed-visibility direct-illumination model of a DEM (oracle_diff.hpp):
    forward image, adjoint and central finite differences over hit records
    (Mesh.render(..., records=True)["records"], shape (spp, W*H, 4))."""

    def __init__(self, heights: np.ndarray, proj: Projection, grey_albedo: np.ndarray | None = None):
        self.shape = heights.shape
        h = np.ascontiguousarray(heights, np.float32)
        g = None if grey_albedo is None else np.ascontiguousarray(grey_albedo, np.float32)
        self.h = lib().orc_dem_diff_create(h.shape[1], h.shape[0], C.byref(proj), h.ctypes.data, _p(g))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_dem_diff_destroy(self.h)
            self.h = None

    @staticmethod
    def _args(records, lights, mat):
        rec = np.ascontiguousarray(records, np.uint32)
        larr = (Light * max(1, len(lights)))(*lights)
        return rec, rec.shape[0], rec.shape[1], larr

    def forward(self, heights, texture, records, lights, mat: Material) -> np.ndarray:
        rec, spp, npix, larr = self._args(records, lights, mat)
        h = np.ascontiguousarray(heights, np.float64).reshape(-1)
        t = None if texture is None else np.ascontiguousarray(texture, np.float64).reshape(-1)
        img = np.zeros((npix, NB), np.float64)
        lib().orc_di_forward64(self.h, h.ctypes.data, _p(t), rec.ctypes.data, npix, spp, larr, len(lights),
                               C.byref(mat), img.ctypes.data, nthreads())
        return img

    def adjoint(self, heights, texture, records, lights, mat: Material, grad_image):
        rec, spp, npix, larr = self._args(records, lights, mat)
        h = np.ascontiguousarray(heights, np.float64).reshape(-1)
        t = None if texture is None else np.ascontiguousarray(texture, np.float64).reshape(-1)
        g = np.ascontiguousarray(grad_image, np.float64).reshape(npix, NB)
        dh = np.zeros(h.size, np.float64)
        dt = None if t is None else np.zeros(t.size, np.float64)
        lib().orc_di_adjoint64(self.h, h.ctypes.data, _p(t), rec.ctypes.data, npix, spp, larr, len(lights),
                               C.byref(mat), g.ctypes.data, dh.ctypes.data, _p(dt))
        return dh, dt

    def fd_heights(self, heights, texture, records, lights, mat: Material, target, nodes, step=1e-3):
        rec, spp, npix, larr = self._args(records, lights, mat)
        h = np.ascontiguousarray(heights, np.float64).reshape(-1)
        t = None if texture is None else np.ascontiguousarray(texture, np.float64).reshape(-1)
        tg = np.ascontiguousarray(target, np.float64).reshape(npix, NB)
        nd = np.ascontiguousarray(nodes, np.int64)
        out = np.zeros(nd.size, np.float64)
        lib().orc_di_fd64(self.h, h.ctypes.data, _p(t), rec.ctypes.data, npix, spp, larr, len(lights),
                          C.byref(mat), tg.ctypes.data, nd.ctypes.data, nd.size, step, out.ctypes.data, nthreads())
        return out

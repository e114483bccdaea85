"""quipu DEM pyramids (.qld): the reader over the library's vira_gpu_qld_* entry
points, a writer, and the level-of-detail rule that picks a level.

* `QuipuDEM` mirrors DEMQuipu (dem_quipu.hpp:34-81): open a file, read its
  header, and `read_buffers(ctx, required_gsd)` builds the chosen level's mesh
  on the GPU (dem_quipu.ipp:131-236), keeping the last level like the
  reference does.
* `write_pyramid` writes the format DEMQuipu::write produces
  (dem_quipu.ipp:238-333, field codecs in quipu_io.ipp), with LZ4 blocks from
  the system liblz4 when `compress` is set.
* `target_gsd` / `lod_update` are LevelOfDetailManager::updateByMapping's GSD
  rule (lod_manager.ipp:149-195) with Camera::calculateGSD (camera.ipp:1408-1414).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import struct
from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi

NUM_BINS = capi.NUM_BINS

# ViraClassID (class_ids.hpp:8-56)
ID_FLOAT, ID_DOUBLE, ID_UINT8, ID_UINT16, ID_DEM_PYRAMID = 1, 2, 3, 4, 2002
# AlbedoType (dem.hpp:27-31)
CONSTANT_ALBEDO, FLOAT_IMAGE_ALBEDO, COLOR_IMAGE_ALBEDO = 0, 1, 2


# ----------------------------------------------------------------- the level record
@dataclass
class DEMProjection:
    """The fields of vira::dems::DEMProjection (dem_projection.hpp:16-35) a
    quipu level stores."""
    resolution: tuple = (0, 0)
    pixel_scale: tuple = (1.0, 1.0)
    tile_definitions: tuple = (0, 0)
    tile_row: int = 0
    tile_col: int = 0
    xoff: float = 0.0
    yoff: float = 0.0
    dx: float = 1.0
    dy: float = 1.0
    is_ogr: bool = False
    proj_ref: str = ""
    row_col_rotation: tuple = (0.0, 0.0)
    model_tie_point: tuple = (0.0, 0.0)

    def gsd(self):
        """DEM::getGSD (dem.ipp:615-621)."""
        return (self.dx * self.pixel_scale[0], self.dy * self.pixel_scale[1])


@dataclass
class DEMLevel:
    """One level: heights (H, W) float32 (non-finite = no data), its projection,
    and an albedo: None, a constant spectrum (8,), a grey image (H, W) or eight
    channels (H, W, 8)."""
    heights: np.ndarray
    projection: DEMProjection
    albedo: object = None


def equirectangular_wkt(radius: float, central_meridian: float = 0.0, standard_parallel: float = 0.0,
                        name: str = "Moon") -> str:
    """A WKT1 equirectangular projection on a sphere, as lunar DEM products carry."""
    return (f'PROJCS["{name}_Equirectangular",GEOGCS["GCS_{name}",DATUM["D_{name}",SPHEROID["{name}",'
            f'{radius!r},0.0]],PRIMEM["Reference_Meridian",0.0],UNIT["Degree",0.0174532925199433]],'
            f'PROJECTION["Equirectangular"],PARAMETER["False_Easting",0.0],PARAMETER["False_Northing",0.0],'
            f'PARAMETER["Central_Meridian",{central_meridian!r}],PARAMETER["Standard_Parallel_1",'
            f'{standard_parallel!r}],UNIT["Meter",1.0]]')


def polar_stereographic_wkt(radius: float, south: bool = False, central_meridian: float = 0.0,
                            latitude_of_origin: float | None = None, scale_factor: float = 1.0,
                            name: str = "Moon") -> str:
    """A WKT1 polar stereographic projection on a sphere, as polar lunar DEM
    products carry: latitude_of_origin +-90 (the pole, with scale_factor) or
    the latitude of true scale."""
    lat = (-90.0 if south else 90.0) if latitude_of_origin is None else latitude_of_origin
    return (f'PROJCS["{name}_Polar_Stereographic",GEOGCS["GCS_{name}",DATUM["D_{name}",SPHEROID["{name}",'
            f'{radius!r},0.0]],PRIMEM["Reference_Meridian",0.0],UNIT["Degree",0.0174532925199433]],'
            f'PROJECTION["Polar_Stereographic"],PARAMETER["latitude_of_origin",{lat!r}],'
            f'PARAMETER["central_meridian",{central_meridian!r}],PARAMETER["scale_factor",{scale_factor!r}],'
            f'PARAMETER["false_easting",0.0],PARAMETER["false_northing",0.0],UNIT["Meter",1.0]]')


# ----------------------------------------------------------------- writer
def _lz4():
    try:
        lib = C.CDLL("liblz4.so.1")
    except OSError as e:
        raise RuntimeError("compressed quipu files need the system liblz4 (liblz4.so.1)") from e
    lib.LZ4_compressBound.restype = C.c_int
    lib.LZ4_compressBound.argtypes = [C.c_int]
    lib.LZ4_compress_default.restype = C.c_int
    lib.LZ4_compress_default.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int]
    return lib


def lz4_compress_block(data: bytes) -> bytes:
    """LZ4_compress_default of one block (compressData, quipu_io.ipp:176-217)."""
    lib = _lz4()
    bound = lib.LZ4_compressBound(len(data))
    out = C.create_string_buffer(bound)
    n = lib.LZ4_compress_default(data, out, len(data), bound)
    if n <= 0:
        raise RuntimeError("LZ4 compression failed")
    return out.raw[:n]


def float_to_fixed(img: np.ndarray, minmax, dtype) -> np.ndarray:
    """floatToFixed (image_utils.ipp:484-506) in FP32: inf -> 0, else
    (T)(((m - 1) * ((v - min) / (max - min))) + 1)."""
    f = np.float32
    v = np.asarray(img, np.float32)
    m = f(np.iinfo(dtype).max)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        val = (v - f(minmax[0])) / (f(minmax[1]) - f(minmax[0]))
        q = ((m - f(1)) * val) + f(1)
    out = np.where(np.isinf(v), 0, np.trunc(np.nan_to_num(q, nan=0.0))).astype(np.int64)
    return out.astype(dtype)


def _valid_minmax(img: np.ndarray):
    """Image::minmax (image.ipp:693-737): over finite values; (FLT_MAX, lowest) if none."""
    v = np.asarray(img, np.float32)
    fin = v[np.isfinite(v)]
    if fin.size == 0:
        return (np.finfo(np.float32).max, -np.finfo(np.float32).max)
    return (float(fin.min()), float(fin.max()))


class _W:
    def __init__(self):
        self.parts = []
        self.n = 0

    def raw(self, b: bytes):
        self.parts.append(b)
        self.n += len(b)
        return len(b)

    def fmt(self, f: str, *v):
        return self.raw(struct.pack("<" + f, *v))

    def vec(self, values, double=True):   # writeVec (quipu_io.ipp:322-341)
        return self.fmt("HB", ID_DOUBLE if double else ID_FLOAT, len(values)) + \
            self.fmt(("d" if double else "f") * len(values), *values)

    def mat3(self, R):                    # writeMat, m[column][row] (quipu_io.ipp:414-445)
        vals = [float(R[k][i]) for i in range(3) for k in range(3)]  # column i, row k
        return self.fmt("HBB", ID_DOUBLE, 3, 3) + self.fmt("9d", *vals)

    def image(self, img: np.ndarray, class_id: int, compress: bool):   # writeImage (quipu_io.ipp:744-772)
        H, W = img.shape
        n = self.fmt("HIIB", class_id, W, H, 1 if compress else 0)
        if compress:
            data = np.ascontiguousarray(img).tobytes()
            block = lz4_compress_block(data)
            n += self.fmt("QQ", len(data), len(block)) + self.raw(block)
        else:
            n += self.raw(np.ascontiguousarray(img.T).tobytes())  # i outer, j inner
        return n


def _transform_parts(M):
    """getPositionFromTransformation / getScaleFromTransformation /
    getRotationFromTransformation (reference_frame.ipp:503-560) of a
    column-major 4x4 given as a row-major numpy matrix."""
    M = np.asarray(M, np.float64)
    pos = M[:3, 3]
    scale = np.linalg.norm(M[:3, :3], axis=0)
    R = M[:3, :3] / np.abs(scale)
    return pos, R, scale


def _grid_normals(P: np.ndarray, valid: np.ndarray) -> np.ndarray:
    """Vertex normals of the DEM grid for the normal cone: imageToIndexBuffer's
    two triangles per quad (image_utils.ipp:371-400), A = ((i, j+1), (i+1, j),
    (i, j)) and B = ((i+1, j), (i, j+1), (i+1, j+1)), unit face normals
    (v1 - v0) x (v2 - v0) (triangle.ipp:36-44) summed per vertex and
    normalised (Mesh::calculateNormals, mesh.ipp:199-233)."""
    N = np.zeros(P.shape)
    P00, P01, P10, P11 = P[:-1, :-1], P[:-1, 1:], P[1:, :-1], P[1:, 1:]
    V00, V01, V10, V11 = valid[:-1, :-1], valid[:-1, 1:], valid[1:, :-1], valid[1:, 1:]

    def unit(v0, v1, v2, ok):
        n = np.cross(v1 - v0, v2 - v0)
        ln = np.linalg.norm(n, axis=2, keepdims=True)
        return np.where((ln > 0) & ok[..., None], n / np.where(ln > 0, ln, 1.0), 0.0)

    nA = unit(P10, P01, P00, V10 & V01 & V00)
    nB = unit(P01, P10, P11, V01 & V10 & V11)
    N[1:, :-1] += nA
    N[:-1, 1:] += nA
    N[:-1, :-1] += nA
    N[:-1, 1:] += nB
    N[1:, :-1] += nB
    N[1:, 1:] += nB
    ln = np.linalg.norm(N, axis=2, keepdims=True)
    return np.where(ln > 0, N / np.where(ln > 0, ln, 1.0), 0.0)


def level_positions(level: DEMLevel) -> np.ndarray:
    """DEMProjection::transformToWorld (dem_projection.ipp:102-144) in FP64 for
    every sample (planar, or OGR equirectangular or polar stereographic on a
    sphere)."""
    h = np.asarray(level.heights, np.float64)
    H, W = h.shape
    p = level.projection
    i = np.arange(W)[None, :].astype(np.float64)
    j = np.arange(H)[:, None].astype(np.float64)
    if not p.is_ogr:
        tx, ty = (W - 1.0) / 2.0, (H - 1.0) / 2.0
        sx, sy = (tx + 0.5) / W, (ty + 0.5) / H
        Ix, Iy = i + p.xoff / p.dx, j + p.yoff / p.dy
        x = p.dx * p.pixel_scale[0] * (-tx + 2.0 * (Ix * sx))
        y = p.dy * p.pixel_scale[1] * (ty + 2.0 * (-Iy * sy))
        return np.stack(np.broadcast_arrays(x, y, h), axis=2)
    a, lon0, lat_ts = _eqc_params(p.proj_ref)
    Ix = p.dx * i + p.xoff + 0.5
    Iy = p.dy * j + p.yoff + 0.5
    X = p.model_tie_point[0] + Ix * p.pixel_scale[0] + Iy * p.row_col_rotation[0]
    Y = p.model_tie_point[1] + Ix * p.row_col_rotation[1] + Iy * p.pixel_scale[1]
    ps = _polar_params(p.proj_ref)
    if ps is not None:
        # spherical polar stereographic inverse (PROJ stere): c = 2 atan(rh / 2k0)
        pole, akm1 = ps
        x, y = X / a, Y / a
        rh = np.hypot(x, y)
        lat = np.where(rh <= 1e-10, pole * np.pi / 2, np.arcsin(pole * np.cos(2.0 * np.arctan(rh / akm1))))
        lon = np.radians(lon0) + np.where(rh <= 1e-10, 0.0, np.arctan2(x, -pole * y))
        lat, lon = np.broadcast_arrays(lat, lon)
    else:
        lon = np.radians(lon0) + X / (a * math.cos(math.radians(lat_ts)))
        lat = Y / a
    r = a + h
    return np.stack([r * np.cos(lon) * np.cos(lat), r * np.sin(lon) * np.cos(lat), r * np.sin(lat)], axis=2)


def _eqc_params(wkt: str):
    import re
    lw = wkt.lower()
    a = float(re.search(r'spheroid\["[^"]*",\s*([-+0-9.eE]+)', lw).group(1))
    cm = re.search(r'parameter\["central_meridian",\s*([-+0-9.eE]+)', lw)
    sp = re.search(r'parameter\["standard_parallel_1",\s*([-+0-9.eE]+)', lw)
    return a, float(cm.group(1)) if cm else 0.0, float(sp.group(1)) if sp else 0.0


def _polar_params(wkt: str):
    """(pole +-1, 2 k0) of a polar stereographic WKT, None for other projections."""
    import re
    lw = wkt.lower()
    m = re.search(r'projection\["([^"]*)"', lw)
    name = m.group(1) if m else ""
    if name not in ("polar_stereographic", "stereographic_north_pole", "stereographic_south_pole"):
        return None
    lo = re.search(r'parameter\["(?:latitude_of_origin|standard_parallel_1)",\s*([-+0-9.eE]+)', lw)
    k = re.search(r'parameter\["scale_factor",\s*([-+0-9.eE]+)', lw)
    lat = float(lo.group(1)) if lo else (-90.0 if name.endswith("south_pole") else 90.0)
    pole = 1.0 if name.endswith("north_pole") else -1.0 if name.endswith("south_pole") else (-1.0 if lat < 0 else 1.0)
    phits = math.radians(abs(lat))
    if abs(phits - math.pi / 2) >= 1e-10:
        return pole, math.cos(phits) / math.tan(math.pi / 4 - 0.5 * phits)
    return pole, 2.0 * (float(k.group(1)) if k else 1.0)


def normal_cone(level: DEMLevel):
    """DEMQuipu::write's normal cone (dem_quipu.ipp:263-276): the normalised sum
    of the coarsest level's vertex normals and the largest angle to it."""
    P = level_positions(level)
    valid = np.isfinite(np.asarray(level.heights))
    N = _grid_normals(np.where(valid[..., None], P, 0.0), valid)
    s = N.reshape(-1, 3).sum(0)
    n = (s / np.linalg.norm(s)).astype(np.float32)
    dots = np.clip((N.reshape(-1, 3).astype(np.float32) @ n), -1.0, 1.0)
    cone = float(np.max(np.arccos(dots))) if dots.size else 0.0
    return n, np.float32(cone)


def write_pyramid(path: str, levels, transformation=np.eye(4), compress: bool = False, write_albedos: bool = True,
                  normal=None, cone_angle=None) -> None:
    """DEMQuipu::write (dem_quipu.ipp:238-333): levels from the finest to the
    coarsest; the heights' fixed-point range is the finest level's min/max."""
    base = levels[0]
    if base.albedo is None:
        albedo_type = CONSTANT_ALBEDO
    else:
        a = np.asarray(base.albedo)
        albedo_type = CONSTANT_ALBEDO if a.ndim == 1 else (FLOAT_IMAGE_ALBEDO if a.ndim == 2 else COLOR_IMAGE_ALBEDO)
    if normal is None or cone_angle is None:
        normal, cone_angle = normal_cone(levels[-1])
    w = _W()
    w.raw(b"QUIPU")
    w.fmt("H", ID_DEM_PYRAMID)
    pos, R, scale = _transform_parts(transformation)
    w.vec(list(pos))
    w.mat3(R)
    w.vec(list(scale))
    w.fmt("H?", len(levels), bool(write_albedos))
    w.vec([float(x) for x in normal], double=False)
    w.fmt("f", float(cone_angle))
    header = w.n
    toc_at = len(w.parts)
    offset = header + len(levels) * 16
    w.raw(b"\0" * (len(levels) * 16))  # the table, filled in below
    hmm = _valid_minmax(base.heights)
    sizes = []
    for lv in levels:
        start = w.n
        p = lv.projection
        H, W = np.asarray(lv.heights).shape
        w.fmt("II", W, H)
        w.fmt("dd", *p.pixel_scale)
        w.fmt("IIII", *p.tile_definitions, p.tile_row, p.tile_col)
        w.fmt("dddd", p.xoff, p.yoff, p.dx, p.dy)
        w.fmt("?", bool(p.is_ogr))
        if p.is_ogr:
            b = p.proj_ref.encode()
            w.fmt("I", len(b))
            w.raw(b)
            w.fmt("dd", *p.row_col_rotation)
            w.fmt("dd", *p.model_tie_point)
        # writeDEMHeights (quipu_io.ipp:635-660)
        g = np.float32(min(p.gsd()))
        gsd_scale = np.float32(g / np.float32(50.0))
        delta = np.float32(np.float32(hmm[1]) - np.float32(hmm[0]))
        w.fmt("ff", *hmm)
        h = np.asarray(lv.heights, np.float32)
        if np.float32(delta / np.float32(255)) < gsd_scale:
            w.fmt("H", ID_UINT8)
            w.image(float_to_fixed(h, hmm, np.uint8), ID_UINT8, compress)
        elif np.float32(delta / np.float32(65535)) < gsd_scale:
            w.fmt("H", ID_UINT16)
            w.image(float_to_fixed(h, hmm, np.uint16), ID_UINT16, compress)
        else:
            w.fmt("H", ID_FLOAT)
            w.image(h, ID_FLOAT, compress)
        if write_albedos:
            w.fmt("B", albedo_type)
            a = lv.albedo
            if albedo_type == CONSTANT_ALBEDO:
                spec = np.broadcast_to(np.asarray(1.0 if a is None else a, np.float32), (NUM_BINS,))
                w.fmt("8f", *[float(x) for x in spec])
            elif albedo_type == FLOAT_IMAGE_ALBEDO:
                mm = _valid_minmax(a)
                w.fmt("ff", *mm)
                w.image(float_to_fixed(a, mm, np.uint8), ID_UINT8, compress)
            else:
                a = np.asarray(a, np.float32)
                w.fmt("I?", a.shape[2], bool(compress))
                for k in range(a.shape[2]):
                    mm = _valid_minmax(a[:, :, k])
                    w.fmt("ff", *mm)
                    w.image(float_to_fixed(a[:, :, k], mm, np.uint8), ID_UINT8, compress)
        sizes.append(w.n - start)
    toc = b""
    for lv, sz in zip(levels, sizes):
        toc += struct.pack("<dQ", lv.projection.gsd()[0], offset)
        offset += sz
    w.parts[toc_at] = toc
    with open(path, "wb") as f:
        for b in w.parts:
            f.write(b)


# ----------------------------------------------------------------- reader
# The star quipu record (star.hpp:27-38): the Star object's bytes, 64 per star.
STAR_DTYPE = np.dtype([("initialized", "u1"), ("pad", "u1", 7), ("ra", "<f8"), ("de", "<f8"), ("pm_ra", "<f8"),
                       ("pm_de", "<f8"), ("vmag", "<f8"), ("temperature", "<f8"), ("omega", "<f8")])


def make_stars(ra, de, vmag, temperature, omega, pm_ra=0.0, pm_de=0.0) -> np.ndarray:
    """A star catalogue as STAR_DTYPE records (the Star constructors, star.ipp:34-46)."""
    ra = np.atleast_1d(np.asarray(ra, np.float64))
    out = np.zeros(ra.shape[0], STAR_DTYPE)
    out["initialized"] = 1
    for name, v in (("ra", ra), ("de", de), ("pm_ra", pm_ra), ("pm_de", pm_de), ("vmag", vmag),
                    ("temperature", temperature), ("omega", omega)):
        out[name] = v
    return out


def write_star_quipu(path, stars: np.ndarray) -> str:
    """StarQuipu::write (star_quipu.ipp:31-41): the extension becomes .qsc, then
    "QUIPU", uint32 count, uint32 byte size and the raw records (writeVector,
    quipu_io.ipp:106-116)."""
    stars = np.ascontiguousarray(stars, STAR_DTYPE)
    p = os.path.splitext(str(path))[0] + ".qsc"
    os.makedirs(os.path.dirname(os.path.abspath(p)), exist_ok=True)
    with open(p, "wb") as f:
        f.write(b"QUIPU" + struct.pack("<II", stars.shape[0], stars.nbytes) + stars.tobytes())
    return p


def read_star_quipu(path) -> np.ndarray:
    """StarQuipu::read (star_quipu.ipp:20-29) through vira_gpu_qsc_read: the
    catalogue as STAR_DTYPE records."""
    lib = capi.load()
    n = C.c_size_t(0)
    capi.check(lib.vira_gpu_qsc_read(str(path).encode(), None, 0, C.byref(n)))
    out = np.zeros(n.value, STAR_DTYPE)
    capi.check(lib.vira_gpu_qsc_read(str(path).encode(), out.ctypes.data if n.value else None, n.value,
                                     C.byref(n)))
    return out


@dataclass
class QuipuReaderOptions:
    """DEMQuipuReaderOptions (dem_quipu.hpp:24-30)."""
    read_albedos: bool = True
    default_gsd: float = math.inf
    default_albedo: float = 0.06
    albedo_profile: object = 1.0


def lz4_decode_block(block: bytes, size: int) -> bytes:
    """The library's LZ4 block decoder (host)."""
    lib = capi.load()
    out = C.create_string_buffer(max(1, size))
    n = lib.vira_gpu_lz4_decode_block(block, len(block), out, size)
    if n < 0:
        raise ValueError("malformed LZ4 block")
    return out.raw[:n]


class QuipuDEM:
    """DEMQuipu over vira_gpu_qld_*: the header at construction, levels on demand."""

    def __init__(self, path: str, options: QuipuReaderOptions | None = None):
        self.lib = capi.load()
        self.options = options or QuipuReaderOptions()
        self.h = C.c_void_p()
        info = capi.QldInfo()
        capi.check(self.lib.vira_gpu_qld_open(str(path).encode(), C.byref(self.h), C.byref(info)))
        self.path = str(path)
        self.levels = info.levels
        self.has_albedos = bool(info.has_albedos)
        self.transformation = np.array(info.transformation[:]).reshape(4, 4).T  # row-major view
        self.normal = np.array(info.normal[:], np.float32)
        self.cone_angle = float(info.cone_angle)
        self.gsds = []
        for i in range(self.levels):
            g = C.c_double()
            capi.check(self.lib.vira_gpu_qld_level_gsd(self.h, i, C.byref(g)))
            self.gsds.append(g.value)
        self.default_gsd = self.options.default_gsd
        self.current_gsd = math.inf
        self.current_level = -1
        self.mesh = None

    def close(self):
        if self.h:
            self.lib.vira_gpu_qld_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def select(self, required_gsd: float) -> int:
        lvl = C.c_int32()
        capi.check(self.lib.vira_gpu_qld_select(self.h, float(required_gsd), C.byref(lvl)))
        return lvl.value

    def _opts(self, offset):
        o = capi.QldReadOpts()
        o.read_albedos = int(bool(self.options.read_albedos))
        o.default_albedo = float(self.options.default_albedo)
        prof = np.broadcast_to(np.asarray(self.options.albedo_profile, np.float32), (NUM_BINS,))
        for k in range(NUM_BINS):
            o.albedo_profile[k] = float(prof[k])
        for k in range(3):
            o.offset[k] = float(offset[k])
        return o

    def read_level(self, ctx, level: int, mesh: bool = True, samples: bool = False, offset=(0.0, 0.0, 0.0)):
        """Reads one level: returns (mesh id or None, heights (H, W) or None,
        albedo (H, W, 8) or None, level info dict)."""
        torch = ctx.torch
        info = capi.QldLevel()
        hdev = adev = None
        if samples:
            # the resolution is known only after the read: decode once without outputs to size them
            capi.check(self.lib.vira_gpu_qld_read(ctx.h, self.h, level, C.byref(self._opts(offset)), None, None,
                                                  None, C.byref(info)))
            dev = f"cuda:{ctx.device}"
            hdev = torch.empty((info.height, info.width), dtype=torch.float32, device=dev)
            adev = torch.empty((info.height, info.width, NUM_BINS), dtype=torch.float32, device=dev)
        mid = C.c_int32(-1)
        ctx._call(self.lib.vira_gpu_qld_read, ctx.h, self.h, level, C.byref(self._opts(offset)),
                  C.byref(mid) if mesh else None, None if hdev is None else hdev.data_ptr(),
                  None if adev is None else adev.data_ptr(), C.byref(info))
        d = {f: getattr(info, f) for f, _ in capi.QldLevel._fields_}
        return (mid.value if mesh else None), hdev, adev, d

    def read_buffers(self, ctx, required_gsd: float | None = None, offset=(0.0, 0.0, 0.0)):
        """DEMQuipu::readBuffers (dem_quipu.ipp:131-236): the level for the
        requested GSD (default_gsd if None) as a mesh id; the mesh already
        loaded is kept when the choice does not change."""
        req = self.default_gsd if required_gsd is None else required_gsd
        level = self.select(req)
        if level == self.current_level and self.mesh is not None:
            return self.mesh
        mid, _, _, info = self.read_level(ctx, level, offset=offset)
        if math.isinf(self.default_gsd):
            self.default_gsd = self.gsds[level]
        if self.mesh is not None:
            ctx.mesh_release(self.mesh)
        self.mesh, self.current_level, self.current_gsd = mid, level, self.gsds[level]
        self.last_info = info
        return mid


# ----------------------------------------------------------------- level of detail
def calculate_gsd(camera, distance: float) -> float:
    """Camera::calculateGSD (camera.ipp:1408-1414)."""
    sw, sh = camera.sensor_size()
    gx = distance * sw / (camera.focal_length * camera.width)
    gy = distance * sh / (camera.focal_length * camera.height)
    return min(gx, gy)


def target_gsd(camera, distance: float, target_triangle_size: float = 1.0) -> float:
    """The GSD a mesh at this distance needs (lod_manager.ipp:174-175)."""
    return float(np.float32(target_triangle_size) * np.float32(calculate_gsd(camera, distance)))


def lod_update(current_gsd: float, target: float):
    """lod_manager.ipp:177-193: (the GSD to load, whether to reload).  A mesh
    coarser than needed is reloaded; a finer one is kept unless the target is
    at least twice its GSD."""
    if current_gsd > target:
        return target, True
    if current_gsd < target:
        if current_gsd * 2.0 > target:
            return current_gsd, False
        return target, True
    return target, False

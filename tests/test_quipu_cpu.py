"""quipu (.qld) reader on the host: the library's LZ4 block decoder against
the system liblz4 (the codec the reference links, quipu_io.ipp:176-235), the
header and level table written by vira_amd.quipu.write_pyramid
(DEMQuipu::write, dem_quipu.ipp:238-333) and read back by vira_gpu_qld_open
(DEMQuipu::DEMQuipu, dem_quipu.ipp:29-56), the field sizes the reference's own
quipu I/O test fixes (tests/quipu/test_quipu_io.cpp:12-70), readBuffers' level
choice and the level-of-detail rule.  No GPU is needed."""
import ctypes as C
import math
import struct

import numpy as np
import pytest

from vira_amd import quipu as Q
from vira_amd import scene as S


def _liblz4():
    try:
        return Q._lz4()
    except RuntimeError:
        pytest.skip("system liblz4 not present")


def _lz4_decompress_ref(block: bytes, size: int) -> bytes:
    lib = _liblz4()
    lib.LZ4_decompress_safe.restype = C.c_int
    lib.LZ4_decompress_safe.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int]
    out = C.create_string_buffer(max(1, size))
    n = lib.LZ4_decompress_safe(block, out, len(block), size)
    return out.raw[:n] if n >= 0 else None


def _payloads():
    rng = np.random.default_rng(0)
    yield b""
    yield b"a"
    yield b"abcd" * 3
    yield bytes(rng.integers(0, 256, 100_000, dtype=np.uint8))           # incompressible
    yield b"\x07" * 300_000                                               # offset-1 runs, long match lengths
    yield (b"0123456789" * 50_000)[:499_999]
    h = (np.cumsum(rng.integers(-3, 4, 1 << 18)) & 0xFFFF).astype(np.uint16)
    yield h.tobytes()                                                     # a smooth 16-bit height profile
    words = [b"crater ", b"regolith ", b"mare ", b"highland "]
    yield b"".join(words[k] for k in rng.integers(0, 4, 40_000))


@pytest.mark.parametrize("k", range(8))
def test_lz4_decoder_matches_liblz4(k):
    data = list(_payloads())[k]
    block = Q.lz4_compress_block(data) if data else Q.lz4_compress_block(b"")
    assert Q.lz4_decode_block(block, len(data)) == data
    if data:
        assert _lz4_decompress_ref(block, len(data)) == data


def test_lz4_decoder_rejects_malformed_blocks():
    from vira_amd import _capi
    lib = _capi.load()
    data = (b"lunar south pole " * 400) + bytes(range(256))
    block = Q.lz4_compress_block(data)

    def dec(b, cap):
        out = C.create_string_buffer(max(1, cap))
        return lib.vira_gpu_lz4_decode_block(b, len(b), out, cap)

    assert dec(block, len(data)) == len(data)
    # every truncation fails cleanly (or, if it ends on a sequence boundary, is short)
    for n in range(0, len(block)):
        r = dec(block[:n], len(data))
        assert r == -1 or r < len(data)
    assert dec(block, len(data) - 1) == -1           # output does not fit
    # a match offset of zero, or reaching before the start of the output
    assert dec(bytes([0x10, 0x41, 0x00, 0x00]), 64) == -1
    assert dec(bytes([0x10, 0x41, 0x05, 0x00]), 64) == -1
    # a literal run longer than the input
    assert dec(bytes([0xF0, 0xFF, 0x10]), 1 << 20) == -1


def test_field_sizes_as_the_reference_io_test():
    """writeVec: class id (2 B) + length (1 B) + N values; writeMat: class id +
    two sizes + N*M values (tests/quipu/test_quipu_io.cpp:15, 46)."""
    for n in (1, 2, 3, 4):
        w = Q._W()
        assert w.vec([2.3 * i + 4.5 for i in range(n)]) == 2 + 1 + 8 * n
        assert w.vec([2.3 * i + 4.5 for i in range(n)], double=False) == 2 + 1 + 4 * n
    w = Q._W()
    assert w.mat3(np.eye(3)) == 2 + 2 + 9 * 8


def _levels(n=3, base=33, scale=30.0, albedo=None, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        W = (base - 1) // (1 << k) + 1
        h = (100.0 * rng.random((W + 2, W))).astype(np.float32)
        p = Q.DEMProjection(resolution=(W, W + 2), pixel_scale=(scale, scale), dx=float(1 << k), dy=float(1 << k))
        out.append(Q.DEMLevel(h, p, albedo))
    return out


def _rotation(ax, ang):
    ax = np.asarray(ax, float) / np.linalg.norm(ax)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def test_header_and_level_table_round_trip(tmp_path):
    M = np.eye(4)
    M[:3, :3] = _rotation((1, 2, 3), 0.7) @ np.diag([2.0, 3.0, 0.5])
    M[:3, 3] = (1.5e11, -2.0e10, 3.25e9)
    lv = _levels()
    path = tmp_path / "p.qld"
    Q.write_pyramid(str(path), lv, transformation=M, compress=False, normal=(0.0, 0.6, 0.8), cone_angle=0.25)
    d = Q.QuipuDEM(str(path))
    assert d.levels == 3 and d.has_albedos
    assert d.gsds == [30.0, 60.0, 120.0]
    assert np.allclose(d.transformation, M, rtol=1e-12, atol=1e-3)
    assert np.allclose(d.normal, (0.0, 0.6, 0.8)) and abs(d.cone_angle - 0.25) < 1e-7
    # the table's offsets point at each level's projection record
    raw = path.read_bytes()
    toc_at = raw.index(struct.pack("<d", 30.0))
    for k in range(3):
        gsd, off = struct.unpack_from("<dQ", raw, toc_at + 16 * k)
        W, H = struct.unpack_from("<II", raw, off)
        assert gsd == d.gsds[k] and (W, H) == (lv[k].heights.shape[1], lv[k].heights.shape[0])
    d.close()


def test_open_rejects_bad_files(tmp_path):
    from vira_amd._capi import ViraGPUError
    bad = tmp_path / "bad.qld"
    bad.write_bytes(b"NOTQUIPU")
    with pytest.raises(ViraGPUError, match="not a valid Quipu"):
        Q.QuipuDEM(str(bad))
    good = tmp_path / "good.qld"
    Q.write_pyramid(str(good), _levels(1), compress=False)
    raw = good.read_bytes()
    for cut in (3, 7, 40, 90):
        (tmp_path / "cut.qld").write_bytes(raw[:cut])
        with pytest.raises(ViraGPUError):
            Q.QuipuDEM(str(tmp_path / "cut.qld"))
    other = bytearray(raw)
    other[5:7] = struct.pack("<H", 2001)  # a single DEM, not a pyramid
    (tmp_path / "cls.qld").write_bytes(bytes(other))
    with pytest.raises(ViraGPUError, match="2002"):
        Q.QuipuDEM(str(tmp_path / "cls.qld"))
    with pytest.raises(ViraGPUError, match="cannot open"):
        Q.QuipuDEM(str(tmp_path / "missing.qld"))


def test_level_choice_as_read_buffers(tmp_path):
    def pyramid(gsds):
        lv = []
        for g in gsds:
            p = Q.DEMProjection(resolution=(3, 3), pixel_scale=(g, g))
            lv.append(Q.DEMLevel(np.zeros((3, 3), np.float32), p))
        path = tmp_path / f"g{len(gsds)}_{gsds[-1]}.qld"
        Q.write_pyramid(str(path), lv, normal=(0, 0, 1), cone_angle=0.0)
        return Q.QuipuDEM(str(path))

    d = pyramid([1.0, 2.0, 4.0, 8.0])
    for req, want in ((0.5, 0), (1.0, 0), (1.99, 0), (2.0, 1), (3.0, 1), (8.0, 3), (math.inf, 3)):
        assert d.select(req) == want, (req, want)
    # the scan stops at the first level coarser than asked
    d = pyramid([1.0, 4.0, 2.0])
    assert d.select(3.0) == 0 and d.select(5.0) == 2


def test_lod_rule():
    cam = S.Camera(1000, 800, fov_deg=30.0, focal_length=0.1)
    sw, sh = cam.sensor_size()
    d = 2.0e5
    assert math.isclose(Q.calculate_gsd(cam, d), min(d * sw / (0.1 * 1000), d * sh / (0.1 * 800)))
    assert math.isclose(Q.target_gsd(cam, d, 2.0), 2.0 * Q.calculate_gsd(cam, d), rel_tol=1e-6)
    assert Q.lod_update(100.0, 50.0) == (50.0, True)      # too coarse: reload finer
    assert Q.lod_update(40.0, 50.0) == (40.0, False)      # finer but within 2x: keep
    assert Q.lod_update(20.0, 50.0) == (50.0, True)       # more than 2x finer: reload coarser
    assert Q.lod_update(50.0, 50.0) == (50.0, False)


def test_fixed_point_codec_is_the_reference_formula():
    """floatToFixed then fixedToFloat (image_utils.ipp:484-527) on the host:
    the reader's dequantisation of a value round-trips within one step (the
    reference's +1 offset on write is not undone on read, so values come back
    one quantum high), and no-data maps to 0 and back to +inf."""
    h = np.array([[-12.5, 0.0, 7.25, np.inf], [100.0, 55.5, -12.5, 3.0]], np.float32)
    mm = Q._valid_minmax(h)
    assert mm == (-12.5, 100.0)
    q = Q.float_to_fixed(h, mm, np.uint16)
    assert q[0, 3] == 0 and q[0, 0] == 1 and q[1, 0] == 65535
    back = (q.astype(np.float32) / np.float32(65534)) * np.float32(mm[1] - mm[0]) + np.float32(mm[0])
    step = (mm[1] - mm[0]) / 65534
    fin = np.isfinite(h)
    assert np.all(np.abs(back[fin] - h[fin] - step) <= step)


# ---- star quipu (.qsc): StarQuipu::write / StarQuipu::read

def test_star_quipu_round_trip_through_the_c_abi(tmp_path):
    from vira_amd import quipu as Q
    rng = np.random.default_rng(5)
    n = 1000
    st = Q.make_stars(rng.uniform(0, 2 * np.pi, n), rng.uniform(-1.5, 1.5, n), rng.uniform(-1, 12, n),
                      rng.uniform(2500, 40000, n), 10 ** rng.uniform(-18, -14, n),
                      pm_ra=rng.normal(0, 1e-8, n), pm_de=rng.normal(0, 1e-8, n))
    p = Q.write_star_quipu(tmp_path / "cat.bin", st)
    assert p.endswith(".qsc")
    raw = open(p, "rb").read()
    assert raw[:5] == b"QUIPU" and len(raw) == 5 + 8 + 64 * n
    assert np.frombuffer(raw[5:13], "<u4").tolist() == [n, 64 * n]
    back = Q.read_star_quipu(p)
    assert back.dtype == Q.STAR_DTYPE and back.tobytes() == st.tobytes()


def test_star_quipu_empty_and_malformed(tmp_path):
    from vira_amd import quipu as Q
    from vira_amd import _capi as capi
    p = Q.write_star_quipu(tmp_path / "empty", Q.make_stars([], [], [], [], []))
    assert Q.read_star_quipu(p).shape == (0,)
    bad = tmp_path / "bad.qsc"
    bad.write_bytes(b"QUIPX" + np.array([1, 64], "<u4").tobytes() + bytes(64))
    with pytest.raises(capi.ViraGPUError, match="not a valid Quipu"):
        Q.read_star_quipu(bad)
    bad.write_bytes(b"QUIPU" + np.array([2, 64], "<u4").tobytes() + bytes(128))
    with pytest.raises(capi.ViraGPUError, match="does not hold"):
        Q.read_star_quipu(bad)
    bad.write_bytes(b"QUIPU" + np.array([2, 128], "<u4").tobytes() + bytes(100))
    with pytest.raises(capi.ViraGPUError, match="truncated"):
        Q.read_star_quipu(bad)
    with pytest.raises(capi.ViraGPUError, match="cannot open"):
        Q.read_star_quipu(tmp_path / "missing.qsc")


# ---- polar stereographic levels: the inverse of level_positions against the
# geometric projection (from the far pole onto the plane at the near pole)

def _ps_level(R, n, px, south=False, lat=None, k=1.0, lon0=0.0):
    p = Q.DEMProjection(resolution=(n, n), pixel_scale=(px, -px), is_ogr=True,
                        proj_ref=Q.polar_stereographic_wkt(R, south=south, central_meridian=lon0,
                                                           latitude_of_origin=lat, scale_factor=k),
                        model_tie_point=(-(n / 2) * px, (n / 2) * px))
    return Q.DEMLevel(np.zeros((n, n), np.float32), p)


def _grid_xy(lv):
    p = lv.projection
    n = lv.heights.shape[0]
    c = np.arange(n) + 0.5
    X = p.model_tie_point[0] + c[None, :] * p.pixel_scale[0]
    Y = p.model_tie_point[1] + c[:, None] * p.pixel_scale[1]
    return np.broadcast_arrays(X, Y)


@pytest.mark.parametrize("south,k,lon0", [(False, 1.0, 0.0), (True, 1.0, 0.0), (False, 0.994, 0.0),
                                          (True, 1.0, 45.0)])
def test_polar_stereographic_matches_geometric_projection(south, k, lon0):
    R = 1737400.0
    lv = _ps_level(R, 33, 20000.0, south=south, k=k, lon0=lon0)
    P = Q.level_positions(lv)
    u = P / np.linalg.norm(P, axis=2, keepdims=True)
    assert np.allclose(np.linalg.norm(P, axis=2), R, rtol=1e-14)
    # rotate the central meridian onto +x, then project from the far pole
    t = np.radians(lon0)
    ux = np.cos(t) * u[..., 0] + np.sin(t) * u[..., 1]
    uy = -np.sin(t) * u[..., 0] + np.cos(t) * u[..., 1]
    s = -1.0 if south else 1.0
    q = 2.0 * k * R / (1.0 + s * u[..., 2])
    X, Y = _grid_xy(lv)
    np.testing.assert_allclose(q * uy, X, atol=1e-6)
    np.testing.assert_allclose(-s * q * ux, Y, atol=1e-6)


def test_polar_stereographic_known_points():
    R = 1000.0
    # 2 R k0 from the pole is the equator (k0 = 1); the pole is the origin
    lv = _ps_level(R, 2, 2 * R)      # pixel centres at (+-R, +-R): rho = R sqrt 2
    P = Q.level_positions(lv)
    colat = np.degrees(np.arccos(P[..., 2] / R))
    np.testing.assert_allclose(colat, np.degrees(2 * np.arctan(np.sqrt(2) / 2)), rtol=1e-12)
    lv = _ps_level(R, 2, 4 * R / np.sqrt(2))   # rho = 2 R: on the equator
    assert np.allclose(Q.level_positions(lv)[..., 2], 0.0, atol=1e-9)
    lv = Q.DEMLevel(np.zeros((1, 1), np.float32),
                    Q.DEMProjection(resolution=(1, 1), pixel_scale=(2.0, -2.0), is_ogr=True,
                                    proj_ref=Q.polar_stereographic_wkt(R, south=True),
                                    model_tie_point=(-1.0, 1.0)))
    np.testing.assert_allclose(Q.level_positions(lv)[0, 0], [0, 0, -R], atol=1e-9)


def test_polar_stereographic_true_scale_latitude():
    # a latitude of origin other than +-90 is the latitude of true scale: there,
    # one pixel along the parallel spans one pixel's length on the sphere
    R, px = 1737400.0, 10.0
    for lat in (71.0, -80.0):
        lv = _ps_level(R, 2, px, south=lat < 0, lat=lat)
        pole, akm1 = Q._polar_params(lv.projection.proj_ref)
        assert pole == np.sign(lat) and akm1 == pytest.approx(1 + np.sin(np.radians(abs(lat))), rel=1e-14)
        # shift the grid to the true-scale parallel on the central meridian
        rho = R * akm1 * np.tan(np.radians(90 - abs(lat)) / 2)
        p = lv.projection
        p.model_tie_point = (-px, (-rho if lat > 0 else rho) + px)
        P = Q.level_positions(lv)
        lat_got = np.degrees(np.arcsin(P[..., 2] / R))
        assert abs(abs(lat_got).mean() - abs(lat)) < 1e-3
        ds = np.linalg.norm(P[0, 1] - P[0, 0])
        assert ds == pytest.approx(px, rel=1e-4)

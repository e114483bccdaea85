"""Pins the CPU oracle before it is trusted (CPU only).

The reference ships no golden vectors for this path (its tests cover quipu
I/O, units and Vulkan plumbing only, tests/CMakeLists.txt:1-15), so each
oracle function is pinned by what the mathematics fixes: closed forms,
invariants, special cases, brute force on small inputs, and an independent
vectorised restatement where the function is a definition of this project's
synthetic inputs.
"""
import hashlib
import json
import math
from pathlib import Path

import numpy as np
import pytest

import pyoracle as O
from vira_amd import scene as S
from vira_amd import workloads as W

GOLDEN = Path(__file__).resolve().parent / "golden"


# ---------------------------------------------------------------- synthetic DEM (definition)
def _lowbias32(x):
    x = x.astype(np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def _fbm_numpy(W_, H_, seed, amplitude=500.0, octaves=6, base=64.0):
    """Independent vectorised restatement of the fBm definition."""
    i = np.arange(W_, dtype=np.float64)[None, :]
    j = np.arange(H_, dtype=np.float64)[:, None]
    total = np.zeros((H_, W_))
    norm, amp, scale = 0.0, 1.0, 1.0 / base
    for o in range(octaves):
        x, y = i * scale, j * scale
        x0, y0 = np.floor(x), np.floor(y)
        fx, fy = x - x0, y - y0

        def lat(ix, iy):
            h = _lowbias32(np.full(np.broadcast(ix, iy).shape, seed, np.uint64))
            h = _lowbias32(np.uint64(o) ^ h)
            h = _lowbias32(iy.astype(np.int64).astype(np.uint64) & 0xFFFFFFFF ^ h)
            h = _lowbias32(ix.astype(np.int64).astype(np.uint64) & 0xFFFFFFFF ^ h)
            return (h >> 8).astype(np.float64) * (1.0 / 8388608.0) - 1.0

        X0 = np.broadcast_to(x0, (H_, W_))
        Y0 = np.broadcast_to(y0, (H_, W_))
        v00, v10, v01, v11 = lat(X0, Y0), lat(X0 + 1, Y0), lat(X0, Y0 + 1), lat(X0 + 1, Y0 + 1)
        fade = lambda t: t * t * t * (t * (t * 6.0 - 15.0) + 10.0)
        sx, sy = np.broadcast_to(fade(fx), (H_, W_)), np.broadcast_to(fade(fy), (H_, W_))
        lerp = lambda a, b, t: a + t * (b - a)
        n = lerp(lerp(v00, v10, sx), lerp(v01, v11, sx), sy)
        total = total + amp * n
        norm = norm + amp
        amp *= 0.5
        scale *= 2.0
    return (amplitude * (total / norm)).astype(np.float32)


def test_fbm_matches_independent_restatement():
    for (w, h, s) in [(257, 257, 0), (65, 33, 2), (100, 7, 12345)]:
        a = O.fbm_heights(w, h, s)
        b = _fbm_numpy(w, h, s)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (w, h, s)
        assert np.abs(a).max() <= 500.0


def test_fbm_golden_checksums():
    """Checksums of the benchmark DEMs, pinned across rounds."""
    fix = json.loads((GOLDEN / "fbm_checksums.json").read_text())
    for key, want in fix.items():
        n, seed = map(int, key.split("_"))
        got = hashlib.sha256(O.fbm_heights(n, n, seed).tobytes()).hexdigest()
        assert got == want, key


def test_vertex_albedo_range():
    a = O.vertex_albedo(100000, 0)
    assert a.min() >= 0.10 and a.max() <= 0.14 and abs(a.mean() - 0.12) < 1e-3


# ---------------------------------------------------------------- tessellation
def test_index_buffer_closed_form():
    h = np.zeros((4, 5), np.float32)  # H=4 rows (j), W=5 columns (i)
    idx = O.dem_indices(h)
    Wd, Hd = 5, 4
    want = []
    for i in range(Wd - 1):
        for j in range(Hd - 1):
            want.append([i + (j + 1) * Wd, i + 1 + j * Wd, i + j * Wd])
            want.append([i + 1 + j * Wd, i + (j + 1) * Wd, i + 1 + (j + 1) * Wd])
    assert np.array_equal(idx, np.array(want, np.uint32))
    # triangle id = 2 * (i * (H - 1) + j) + {0, 1}
    assert idx.shape[0] == 2 * (Wd - 1) * (Hd - 1)


def test_index_buffer_drops_holes():
    h = np.zeros((4, 5), np.float32)
    h[1, 2] = np.inf
    idx = O.dem_indices(h)
    hole = 1 * 5 + 2
    assert not (idx == hole).any()
    assert idx.shape[0] == 24 - 6  # an interior vertex touches 6 triangles


def test_flat_and_tilted_normals():
    h = np.zeros((6, 7), np.float32)
    pos, _ = O.dem_positions(h, O.make_projection(0, 30.0))
    idx = O.dem_indices(h)
    n = O.vertex_normals(pos, idx)
    assert np.array_equal(n, np.tile(np.array([0, 0, 1], np.float32), (42, 1)))
    # plane z = a*x: normal (-a, 0, 1)/sqrt(1+a^2)
    a = 0.25
    ii = np.arange(7, dtype=np.float32)[None, :].repeat(6, 0)
    h2 = (a * (ii - 3) * 30.0).astype(np.float32)
    pos2, _ = O.dem_positions(h2, O.make_projection(0, 30.0))
    n2 = O.vertex_normals(pos2, idx)
    want = np.array([-a, 0, 1]) / math.sqrt(1 + a * a)
    assert np.abs(n2 - want).max() < 1e-6


def test_sphere_drape_radius():
    h = np.full((9, 9), 100.0, np.float32)
    R = 1737.4e3
    _, p64 = O.dem_positions(h, O.make_projection(1, 30.0, R))
    r = np.linalg.norm(p64, axis=1)
    assert np.abs(r - (R + 100.0)).max() < 1e-6
    assert np.allclose(p64[4 * 9 + 4], [R + 100.0, 0, 0])


# ---------------------------------------------------------------- watertight intersection
def _hit(ray, tri):
    t, b1, b2 = np.zeros(1, np.float32), np.zeros(1, np.float32), np.zeros(1, np.float32)
    r = np.ascontiguousarray(ray, np.float32)
    tr = np.ascontiguousarray(tri, np.float32)
    ok = O.lib().orc_intersect_tri(r.ctypes.data, tr.ctypes.data, t.ctypes.data, b1.ctypes.data, b2.ctypes.data)
    return ok, float(t[0]), float(b1[0]), float(b2[0])


TRI = [0, 0, 0, 1, 0, 0, 0, 1, 0]


def test_intersect_known_answers():
    ok, t, b1, b2 = _hit([0.25, 0.25, 1, 0, 0, -1], TRI)
    assert ok and t == 1.0 and b1 == 0.25 and b2 == 0.25
    assert not _hit([0.75, 0.75, 1, 0, 0, -1], TRI)[0]          # outside
    assert _hit([0.25, 0.25, -1, 0, 0, 1], TRI)[0]              # back face: no culling
    assert not _hit([0.25, 0.25, 1, 0, 0, 1], TRI)[0]           # behind the origin
    assert not _hit([-1, 0.25, 0, 1, 0, 0], TRI)[0]             # in the triangle's plane
    ok, t, _, _ = _hit([0.2, 0.3, 5, 0, 0, -2], TRI)            # unnormalised direction: t in its units
    assert ok and t == 2.5


def test_watertight_shared_edges_and_vertices():
    # two triangles sharing the diagonal x + y = 1, and a fan around (0.5, 0.5)
    quad = [[0, 0, 0, 1, 0, 0, 0, 1, 0], [1, 0, 0, 1, 1, 0, 0, 1, 0]]
    for k in range(65):
        x = k / 64.0
        ray = [x, 1 - x, 3.0, 0, 0, -1]
        assert _hit(ray, quad[0])[0] or _hit(ray, quad[1])[0], x
    c = [0.5, 0.5, 0]
    ring = [[math.cos(a), math.sin(a), 0] for a in np.linspace(0, 2 * math.pi, 7)[:-1]]
    fan = [c + ring[i] + ring[(i + 1) % 6] for i in range(6)]
    for d in ([0, 0, -1], [0.3, -0.2, -1], [-0.01, 0.0, -1]):
        o = np.array(c) - 4 * np.array(d, np.float64)
        ray = list(o) + list(d)
        assert any(_hit(ray, f)[0] for f in fan), d


def _random_soup(n, rng):
    c = rng.uniform(-10, 10, (n, 1, 3))
    v = c + rng.normal(0, 0.6, (n, 3, 3))
    pos = v.reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    return pos, idx


def _random_rays(n, rng, scale=15.0):
    o = rng.uniform(-scale, scale, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


def test_bvh_closest_hit_equals_brute_force():
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 300, 3000):
        pos, idx = _random_soup(n, rng)
        m = O.Mesh(pos, idx)
        assert m.check()[0] == 0
        rays = _random_rays(4000, rng)
        t, i, b1, b2, _ = m.trace(rays)
        tb, ib, b1b, b2b, _ = m.trace(rays, brute=True)
        assert np.array_equal(i, ib)
        assert np.array_equal(t.view(np.uint32), tb.view(np.uint32))
        assert np.array_equal(b1.view(np.uint32), b1b.view(np.uint32))
        tm = np.where(np.isfinite(tb), tb * 1.01, 1e30).astype(np.float32)
        _, ia, _, _, _ = m.trace(rays, any_hit=True, tmax=tm)
        assert np.array_equal(ia != 0xFFFFFFFF, ib != 0xFFFFFFFF)


def test_bvh_dem_grid_brute_force_and_invariants():
    h = O.fbm_heights(33, 29, 4)
    pos, _ = O.dem_positions(h, O.make_projection(0, 30.0))
    idx = O.dem_indices(h)
    m = O.Mesh(pos, idx)
    code, depth, leaf = m.check()
    assert code == 0 and leaf <= 2  # ViraBLAS leaves hold <= 2 triangles (vira_blas.ipp:136)
    rng = np.random.default_rng(3)
    n = 5000
    o = np.zeros((n, 3))
    o[:, :2] = rng.uniform(-500, 500, (n, 2))
    o[:, 2] = 3000
    tgt = np.zeros((n, 3))
    tgt[:, 0] = rng.uniform(-470, 470, n)
    tgt[:, 1] = rng.uniform(-410, 410, n)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    # vertex- and edge-aimed nadir rays: exact zeros in the direction, origins on slab planes
    vx = pos[:, :2][::7]
    nad = np.concatenate([vx, np.full((len(vx), 1), 3000), np.tile([0, 0, -1], (len(vx), 1))], 1)
    vert = np.concatenate([rng.uniform(-479, 479, (2000, 1)), rng.uniform(-419, 419, (2000, 1)),
                           np.full((2000, 1), 3000), np.tile([0, 0, -1], (2000, 1))], 1)
    nv = len(vx) + len(vert)
    rays = np.concatenate([rays, nad.astype(np.float32), vert.astype(np.float32)])
    t, i, _, _, _ = m.trace(rays)
    tb, ib, _, _, _ = m.trace(rays, brute=True)
    assert np.array_equal(i, ib) and np.array_equal(t.view(np.uint32), tb.view(np.uint32))
    assert (i[-nv:] != 0xFFFFFFFF).all()  # vertical rays inside the grid always hit: no cracks


def test_bvh_sah_split_by_hand():
    # two clusters of 4 small triangles at x = 0 and x = 10: the first split separates them
    tris = []
    for cx in (0.0, 10.0):
        for k in range(4):
            y = 2.0 * k
            tris.append([[cx, y, 0], [cx + 0.1, y, 0], [cx, y + 0.1, 0]])
    pos = np.array(tris, np.float32).reshape(-1, 3)
    idx = np.arange(24, dtype=np.uint32).reshape(-1, 3)
    m = O.Mesh(pos, idx)
    lohi, left, cnt, tri = m.nodes()
    assert cnt[0] == 0
    l, r = left[0], left[0] + 1
    assert lohi[l][3] < 1.0 and lohi[r][0] > 9.0  # children boxes split the clusters
    assert m.check()[0] == 0


def test_bvh_degenerate_identical_triangles():
    pos = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (50, 1))
    idx = np.arange(150, dtype=np.uint32).reshape(-1, 3)
    m = O.Mesh(pos, idx)
    assert m.check()[0] == 0
    t, i, _, _, _ = m.trace(np.array([[0.2, 0.2, 1, 0, 0, -1]], np.float32))
    assert i[0] == 0 and t[0] == 1.0  # ties resolve to the lowest id


def test_empty_mesh():
    m = O.Mesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    t, i, _, _, _ = m.trace(np.array([[0, 0, 1, 0, 0, -1]], np.float32))
    assert i[0] == 0xFFFFFFFF and np.isinf(t[0])


# ---------------------------------------------------------------- offset
def test_offset_intersection_integer_ulp():
    p = np.array([1.0, -2.0, 0.01], np.float32)
    n = np.array([0.6, -0.8, 0.0], np.float32)
    r = O.offset_point(p, n)
    # x: +int(256*0.6)=153 ulps (positive x); y: p<0 so subtract int(256*-0.8)=-204 -> +204 ulps away from 0
    assert r[0].view(np.int32) == p[0].view(np.int32) + 153
    assert r[1].view(np.int32) == p[1].view(np.int32) + 204
    # z: |z| < 1/32 -> float offset z + n.z/65536
    assert r[2] == np.float32(0.01) + np.float32(1 / 65536.0) * np.float32(0.0)


# ---------------------------------------------------------------- camera
def test_intrinsics_closed_form():
    cam = S.Camera(512, 384, 40.0)
    sw, sh = cam.sensor_size()
    k = S.intrinsics_inverse(512, 384, cam.focal_length, sw, sh)
    fx = cam.focal_length * 512 / sw
    # h = K^-1 (x, y, 1): ((x - W/2)/fx, (y - H/2)/fy)
    for (x, y) in [(0, 0), (256, 192), (511.5, 10.25)]:
        hx = (k[0] * x + k[2] * y) + k[4]
        hy = (k[1] * x + k[3] * y) + k[5]
        assert abs(hx - (x - 256) / fx) < 1e-6 and abs(hy - (y - 192) / fx) < 1e-6
    assert abs(math.degrees(2 * math.atan(256 / fx)) - 40.0) < 1e-9


def test_look_in_direction_nadir():
    R = S.look_in_direction((0, 0, -1))
    assert np.allclose(R, [[1, 0, 0], [0, -1, 0], [0, 0, -1]])
    R = S.look_in_direction((-1, 0, 0), (0, 0, 1))
    assert np.allclose(R[:, 2], [-1, 0, 0]) and np.allclose(R[:, 1], [0, 0, -1])


# ---------------------------------------------------------------- lights
def test_planck_bands_against_quadrature():
    from scipy.integrate import quad
    h, c, kb = 6.62607015e-34, 299792458.0, 1.380649e-23
    rad = S.planck_band_radiance(5800.0)
    for b in range(8):
        lo, hi = S.BAND_EDGES_NM[b] * 1e-9, S.BAND_EDGES_NM[b + 1] * 1e-9
        f = lambda l: (2 * h * c * c / l**5) / math.expm1(h * c / (l * kb * 5800.0))
        want = quad(f, lo, hi, epsrel=1e-12)[0]
        assert abs(rad[b] / want - 1) < 1e-6


# ---------------------------------------------------------------- shading closed form
def _flat_scene(n=17, spacing=10.0, alb=0.3):
    h = np.zeros((n, n), np.float32)
    pos, pos64 = O.dem_positions(h, O.make_projection(0, spacing))
    idx = O.dem_indices(h)
    nrm = O.vertex_normals(pos, idx)
    a = np.full((n * n, 8), alb, np.float32)
    return O.Mesh(pos, idx, nrm, a, O.dem_uv(n, n), pos64)


@pytest.mark.parametrize("precision", [32, 64])
def test_point_light_closed_form(precision):
    """L = rho/pi * cos^2(theta) * E / (1 + (cos(theta)/pi)^2): the cos^2 and the
    power-heuristic weight of a delta light (cpu_path_tracer.ipp:415-425,472-478)."""
    alb = 0.3
    m = _flat_scene(alb=alb)
    cam = S.Camera(32, 32, 30.0, (0.0, 0.0, 200.0)).look_in_direction((0, 0, -1))
    sun = S.directional_sun(35.0, 20.0, 2.0, distance=1e6)
    I = sun.intensity[0]
    out = m.render(O.camera_from(cam.to_c()), [O.light_from(sun.to_c())], O.material(1.0), O.options(),
                   precision=precision)
    hit = out["tri_id"] != 0xFFFFFFFF
    assert hit.all()
    # hit points from depth along the pixel-corner rays
    oc = cam.to_c()
    k = np.array(oc.kinv, np.float64)
    R = np.array(oc.rot, np.float64).reshape(3, 3)
    jj, ii = np.mgrid[0:32, 0:32]
    hx = (k[0] * ii + k[2] * jj) + k[4]
    hy = (k[1] * ii + k[3] * jj) + k[5]
    d = np.stack([hx, hy, np.ones_like(hx)], -1) @ R.T
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    P = np.array([0, 0, 200.0]) + d.reshape(-1, 3) * out["depth"][:, None].astype(np.float64)
    L = np.array(sun.position) - P
    dist = np.linalg.norm(L, axis=1)
    cth = L[:, 2] / dist
    E = I / dist**2
    want = alb / math.pi * cth**2 * E / (1 + (cth / math.pi) ** 2)
    got = out["radiance"][:, 0].astype(np.float64)
    tol = 2e-5 if precision == 32 else 1.5e-7  # FP64 result stored as FP32
    assert np.abs(got / want - 1).max() < tol
    assert np.allclose(out["direct"], out["radiance"]) and np.abs(out["indirect"]).max() == 0
    assert np.allclose(out["albedo"], alb) and np.allclose(out["normal"], [0, 0, 1])


def test_dem257_fp32_vs_fp64():
    sc = W.DEM257
    h = O.fbm_heights(sc.dem, sc.dem, sc.seed)
    p = W.make_projection(sc)
    proj = O.make_projection(p.kind, p.spacing, p.radius, tuple(p.center), tuple(p.offset))
    m = O.dem_mesh(h, proj, O.vertex_albedo(sc.dem * sc.dem, sc.seed), with_pos64=True)
    cam = O.camera_from(W.make_camera(sc).to_c())
    lights = [O.light_from(L.to_c()) for L in W.make_lights(sc)]
    o32 = m.render(cam, lights, O.material(), O.options())
    o64 = m.render(cam, lights, O.material(), O.options(), precision=64)
    same = o32["tri_id"] == o64["tri_id"]
    assert same.mean() > 0.999
    hit = (o32["tri_id"] != 0xFFFFFFFF) & same
    ref = o64["radiance"][hit]
    lit = ref[:, 0] > 0
    rel = np.abs(o32["radiance"][hit][lit] - ref[lit]) / ref[lit]
    assert np.median(rel) < 2e-6 and np.quantile(rel, 0.999) <= 1e-5 and np.mean(rel <= 1e-5) >= 0.999
    assert o32["counts"]["primary"] == sc.image ** 2


def test_rng_uniform_range_and_independence():
    u = np.array([O.lib().orc_rng_u01(0, p, s, d) for p in range(40) for s in range(5) for d in range(5)])
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.05
    assert len(np.unique(u)) == len(u)


def _rect_solid_angle(x0, x1, y0, y1):
    """Solid angle of the rectangle [x0,x1] x [y0,y1] on the plane z = 1 seen from
    the origin: F(x1,y1) - F(x0,y1) - F(x1,y0) + F(x0,y0), F = atan(xy / sqrt(1+x^2+y^2))."""
    F = lambda x, y: np.arctan(x * y / np.sqrt(1.0 + x * x + y * y))
    return F(x1, y1) - F(x0, y1) - F(x1, y0) + F(x0, y0)


@pytest.mark.parametrize("W,H,fov", [(64, 48, 40.0), (31, 31, 90.0)])
def test_pixel_solid_angle_against_closed_form(W, H, fov):
    cam = S.Camera(W, H, fov, (0.0, 0.0, 0.0))
    c = O.camera_from(cam.to_c())
    om = O.pixel_solid_angles(c).astype(np.float64)
    k = [np.float64(c.kinv[i]) for i in range(6)]
    i = np.arange(W + 1, dtype=np.float64)
    j = np.arange(H + 1, dtype=np.float64)
    xs = k[0] * i + k[4]                 # K^-1 has no skew: x depends on the column only
    ys = k[3] * j + k[5]
    ref = _rect_solid_angle(xs[None, :-1], xs[None, 1:], ys[:-1, None], ys[1:, None])
    ref = np.abs(ref)
    assert np.max(np.abs(om - ref) / ref) < 2e-6
    # the pixels tile the sensor: their sum is the solid angle of the whole sensor
    whole = abs(_rect_solid_angle(xs[0], xs[-1], ys[0], ys[-1]))
    assert abs(om.sum() - whole) / whole < 1e-5

"""The oracle's unresolved-source renderer (CPUUnresolvedRenderer,
cpu_unresolved_renderer.ipp:25-200) pinned against what does not depend on it:

* a source projected onto a pixel centre deposits power x kernel, bit for bit,
  at the kernel's place (orientation, offset and footprint);
* at any sub-pixel position the deposit equals scipy's bilinear resampling of
  the kernel (map_coordinates, order 1, clamped at the upper edge and zero
  below the lower one, which is what addImage's branches do), across the
  reference's 512-pixel ROI seams;
* the kernel chosen is the first whose edge maximum x power is below the
  minimum power, with edge maxima found by brute force;
* the minimum-power cut, the no-PSF pixel assignment (truncation, last source
  wins), sources behind the camera (mirrored unless culled) and additivity.

The library-side host helpers (edge maxima, minimum detectable power,
intrinsic matrix) are pinned here too."""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates

import pyoracle as O
from vira_amd import sensor as SN
from vira_amd import scene as S

NB = 8


def _camera(W=96, H=80, fx=64.0, fy=64.0):
    # K as GLM mat3x2 [col][row]: (fx, kyx, kxy, fy, px, py)
    return np.array([fx, 0.0, 0.0, fy, W / 2.0, H / 2.0], np.float32)


def _kernel(K, seed):
    rng = np.random.default_rng(seed)
    k = rng.random((K, K, NB)).astype(np.float32)
    return k / k.sum(axis=(0, 1), keepdims=True)


def _dir_at(X, Y, intr):
    """Camera-frame (unnormalised) direction projecting to (X, Y) with z = 1."""
    return np.array([(X - intr[4]) / intr[0], (Y - intr[5]) / intr[3], 1.0])


def test_pixel_centre_deposits_power_times_kernel_exactly():
    W, H = 96, 80
    intr = _camera(W, H)
    k = _kernel(9, 1)
    # X = 37.5 = 32 + 64 * 0.0859375: every step exact in FP32
    d = _dir_at(37.5, 21.5, intr)
    assert np.float32(d[0]) == d[0] and np.float32(d[1]) == d[1]
    P = np.linspace(1.0, 2.0, NB).astype(np.float32)
    out = O.render_stars(np.eye(3), W, H, 1.0, 1.0, [k], 0.0, intr, d[None], P[None])
    ix, iy, h = 37, 21, 4
    expect = np.zeros_like(out)
    expect[iy - h:iy + h + 1, ix - h:ix + h + 1] = P * k  # kernel[y][x] -> image rows y
    assert np.array_equal(out, expect)


@pytest.mark.parametrize("X,Y", [(30.25, 40.8), (511.3, 256.6), (512.0, 511.999), (3.7, 2.2), (1021.9, 700.1)])
def test_subpixel_deposit_equals_bilinear_resampling(X, Y):
    W, H = 1024, 768  # two 512-pixel ROI seams in x, one in y
    intr = _camera(W, H, 512.0, 512.0)
    K = 27
    k = _kernel(K, 2)
    d = _dir_at(X, Y, intr)
    P = np.full(NB, 3.0, np.float32)
    out = O.render_stars(np.eye(3), W, H, 1.0, 1.0, [k], 0.0, intr, d[None], P[None]).astype(np.float64)
    # the projected position in FP64 (the FP32 one differs by ~1e-5 px)
    Xp, Yp = intr[0] * d[0] / d[2] + intr[4], intr[3] * d[1] / d[2] + intr[5]
    ix, iy, h = int(np.floor(Xp)), int(np.floor(Yp)), K // 2
    expect = np.zeros_like(out)
    for j in range(max(0, iy - h), min(H, iy + h + 1)):
        for i in range(max(0, ix - h), min(W, ix + h + 1)):
            kx, ky = i - (Xp - K / 2), j - (Yp - K / 2)
            if np.floor(kx) < 0 or np.floor(ky) < 0 or np.floor(kx) >= K or np.floor(ky) >= K:
                continue
            for b in range(NB):
                expect[j, i, b] = P[b] * map_coordinates(k[:, :, b].astype(np.float64), [[ky], [kx]], order=1,
                                                         mode="nearest")[0]
    assert np.abs(out - expect).max() <= 1e-4 * expect.max()
    assert (out > 0).sum() == (expect > 0).sum()


def test_kernel_choice_by_edge_maxima():
    W, H = 256, 256
    intr = _camera(W, H, 128.0, 128.0)
    sizes = [3, 9, 27, 81]
    # band-flat Gaussians: the edge maximum of each is a single value
    ks = []
    for K in sizes:
        y, x = np.mgrid[0:K, 0:K] - K // 2
        g = np.exp(-(x * x + y * y) / (2 * 4.0 ** 2))
        ks.append(np.repeat((g / g.sum())[:, :, None], NB, 2).astype(np.float32))
    edge = [max(np.abs(k[0]).max(), np.abs(k[-1]).max(), np.abs(k[:, 0]).max(), np.abs(k[:, -1]).max()) for k in ks]
    pmin = 1e-3
    d = _dir_at(128.5, 128.5, intr)
    for scale in [2e-3, 1e-1, 1e1, 1e3, 1e6]:
        P = np.full(NB, scale / np.sqrt(NB), np.float32)
        mag = np.sqrt(np.sum(P.astype(np.float64) ** 2))
        want = next((K for K, e in zip(sizes, edge) if e * mag < pmin), sizes[-1])
        out = O.render_stars(np.eye(3), W, H, 1.0, 1.0, ks, pmin, intr, d[None], P[None])
        ys, xs = np.nonzero(out[:, :, 0])
        assert xs.max() - xs.min() + 1 == want and ys.max() - ys.min() + 1 == want, (scale, want)


def test_minimum_power_cut_and_no_psf_assignment():
    W, H = 64, 48
    intr = _camera(W, H, 32.0, 32.0)
    k = _kernel(3, 3)
    pmin = 1.0
    d = np.array([_dir_at(10.5, 10.5, intr), _dir_at(20.5, 20.5, intr)])
    E = np.zeros((2, NB), np.float32)
    E[0] = 0.999 / np.sqrt(NB)  # magnitude just under the cut
    E[1] = 1.001 / np.sqrt(NB)
    out = O.render_stars(np.eye(3), W, H, 1.0, 1.0, [k], pmin, intr, d, E)
    assert out[8:13, 8:13].max() == 0 and out[19:22, 19:22].max() > 0
    # no PSF: power assigned to pixel ((int) X, (int) Y); truncation puts X in (-1, 0) on column 0;
    # of two sources on one pixel the later one wins
    d = np.array([_dir_at(5.2, 7.9, intr), _dir_at(-0.5, 3.5, intr), _dir_at(5.9, 7.1, intr)])
    E = np.stack([np.full(NB, v, np.float32) for v in (2.0, 3.0, 5.0)])
    out = O.render_stars(np.eye(3), W, H, 1.0, 1.0, [], 0.0, intr, d, E)
    assert np.all(out[7, 5] == 5.0) and np.all(out[3, 0] == 3.0)
    assert np.count_nonzero(out[:, :, 0]) == 2


def test_behind_camera_mirrored_unless_culled_and_additivity():
    W, H = 64, 64
    intr = _camera(W, H, 40.0, 40.0)
    k = _kernel(9, 4)
    d = _dir_at(20.3, 30.7, intr)
    d = d / np.linalg.norm(d)
    E = np.full((1, NB), 2.0, np.float32)
    front = O.render_stars(np.eye(3), W, H, 1.0, 1.0, [k], 0.0, intr, d[None], E)
    back = O.render_stars(np.eye(3), W, H, 1.0, 1.0, [k], 0.0, intr, -d[None], E)
    assert front.max() > 0 and np.array_equal(front, back)
    culled = O.render_stars(np.eye(3), W, H, 1.0, 1.0, [k], 0.0, intr, -d[None], E, cull_behind=True)
    assert culled.max() == 0
    d2 = _dir_at(23.9, 28.2, intr)
    both = O.render_stars(np.eye(3), W, H, 1.0, 1.0, [k], 0.0, intr, np.stack([d, d2]), np.concatenate([E, 3 * E]))
    second = O.render_stars(np.eye(3), W, H, 1.0, 1.0, [k], 0.0, intr, d2[None], 3 * E)
    assert np.array_equal(both, front + second)
    # a rotated camera: R^T d takes the global direction to the camera frame
    c, s = np.cos(0.3), np.sin(0.3)
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    rot = O.render_stars(R, W, H, 1.0, 1.0, [k], 0.0, intr, (R @ d)[None], E)
    assert np.abs(rot - front).max() <= 1e-4 * front.max()


def test_edge_max_matches_brute_force():
    rng = np.random.default_rng(5)
    for K in (3, 9, 27):
        k = rng.random((K, K, NB)).astype(np.float32)
        border = np.ones((K, K), bool)
        border[1:-1, 1:-1] = False
        cand = k[border]
        want = cand[np.argmax(np.linalg.norm(cand.astype(np.float64), axis=1))]
        assert np.array_equal(SN.edge_max(k), want)


def test_minimum_detectable_power_closed_form():
    h, c = 6.62607015e-34, 299792458.0
    area = S.Camera(64, 64, focal_length=0.1, f_stop=4.0).aperture_area()
    lam = (400.0 + 75.0 * (np.arange(NB) + 0.5)) * 1e-9
    want = 10.0 * np.sum(h * c / lam * 75e-9) / area
    got = SN.minimum_detectable_power(area, 0.01)
    assert abs(got - want) <= 1e-5 * want


def test_intrinsic_matrix_inverts_kinv():
    cam = S.Camera(640, 480, fov_deg=30.0, focal_length=0.08)
    K = SN.intrinsic_matrix(cam).astype(np.float64)
    sw, sh = cam.sensor_size()
    kinv = S.intrinsics_inverse(cam.width, cam.height, cam.focal_length, sw, sh, cam.z_dir).astype(np.float64)
    Km = np.array([[K[0], K[2], K[4]], [K[1], K[3], K[5]], [0, 0, 1]])
    Ki = np.array([[kinv[0], kinv[2], kinv[4]], [kinv[1], kinv[3], kinv[5]], [0, 0, 1]])
    assert np.allclose(Km @ Ki, np.eye(3), atol=1e-3)  # FP32 entries up to 320 px
    assert abs(K[0] - cam.focal_length * cam.width / sw) <= 1e-4 * K[0]


# ---- star catalogue light (StarCatalogue::makeStarLight)

def _band_integral_series(T, lo, hi, terms=400):
    """Planck's law integrated over [lo, hi] metres in closed form: with
    x = hc / (lambda kB T), the integral is 2 kB^4 T^4 / (h^3 c^2) times
    G(x(hi)) - G(x(lo)), G(a) = sum_n e^(-n a) (a^3/n + 3a^2/n^2 + 6a/n^3 + 6/n^4)
    (the blackbody fractional-emission series)."""
    c, h, kB = 299792458.0, 6.62607015e-34, 1.380649e-23

    def G(a):
        n = np.arange(1, terms + 1, dtype=np.float64)
        return float(np.sum(np.exp(-n * a) * (a ** 3 / n + 3 * a ** 2 / n ** 2 + 6 * a / n ** 3 + 6 / n ** 4)))

    x = lambda lam: h * c / (lam * kB * T)
    return 2 * kB ** 4 * T ** 4 / (h ** 3 * c ** 2) * (G(x(hi)) - G(x(lo)))


def test_series_matches_stefan_boltzmann_at_the_limits():
    # G(0) = pi^4 / 15: the whole spectrum gives sigma T^4 / pi
    T = 5772.0
    sigma = 5.670374419e-8
    whole = _band_integral_series(T, 1e-12, 1e3, terms=4000)
    assert whole == pytest.approx(sigma * T ** 4 / np.pi, rel=1e-6)


@pytest.mark.parametrize("T", [2500.0, 5772.0, 9600.0, 30000.0])
def test_black_body_bands_match_closed_form_band_integrals(T):
    bands = O.black_body_bands(T)[0]
    fine = O.black_body_bands(T, steps=1000)[0]
    e = O.band_edges_m()
    for k in range(O.NB):
        exact = _band_integral_series(T, e[k], e[k + 1])
        # trapezoid error (b - a) h^2 f''/12: below 1e-4 at 100 points over a
        # 75 nm band (the cool star's Wien tail is the steepest), and O(h^2), so
        # 1000 points land within FP32 rounding of the closed form
        assert bands[k] == pytest.approx(exact, rel=1e-4), (T, k)
        assert fine[k] == pytest.approx(exact, rel=1e-6), (T, k)


def test_band_edges_are_the_fp32_spectrum_bounds():
    e = O.band_edges_m()
    assert e.dtype == np.float64 and e.shape == (O.NB + 1,)
    nm = np.array([400, 475, 550, 625, 700, 775, 850, 925, 1000], np.float32)
    assert np.array_equal(e, (nm * np.float32(1e-9)).astype(np.float64))
    assert e[0] != 400e-9   # FP32 edges, widened: not the FP64 decimal


def test_star_directions_closed_forms_and_proper_motion():
    from vira_amd import quipu as Q
    st = Q.make_stars([0.0, np.pi / 2, 1.0, 0.3], [0.0, 0.0, np.pi / 2, -0.4], 1.0, 5772.0, 1e-10)
    d, _ = O.star_light(st, 0.0)
    np.testing.assert_allclose(d[0], [1, 0, 0], atol=1e-7)
    np.testing.assert_allclose(d[1], [0, 1, 0], atol=1e-7)
    np.testing.assert_allclose(d[2], [0, 0, 1], atol=1e-7)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    # RA 0.3, DE -0.4: z = sin(DE), atan2(y, x) = RA
    assert d[3, 2] == pytest.approx(np.sin(-0.4), abs=1e-7)
    assert np.arctan2(d[3, 1], d[3, 0]) == pytest.approx(0.3, abs=1e-6)
    # proper motion in rad/yr over et seconds: the reference adds the RA rate to
    # the declination and the DE rate to the right ascension (star.ipp:27-28)
    year = 365.25 * 86400.0
    moving = Q.make_stars([0.2, 0.2], [0.1, 0.1], 1.0, 5772.0, 1e-10, pm_ra=[0.05, 0.0], pm_de=[0.0, 0.05])
    d2, _ = O.star_light(moving, 2 * year)
    d_de, _ = O.star_light(Q.make_stars(0.2, 0.1 + 0.1, 1.0, 5772.0, 1e-10), 0.0)
    d_ra, _ = O.star_light(Q.make_stars(0.2 + 0.1, 0.1, 1.0, 5772.0, 1e-10), 0.0)
    np.testing.assert_array_equal(d2[0], d_de[0])
    np.testing.assert_array_equal(d2[1], d_ra[0])


def test_star_irradiance_is_fp32_omega_times_band_radiance():
    from vira_amd import quipu as Q
    om = np.array([1e-12, 3.3e-9, 2.0])
    st = Q.make_stars([0.0] * 3, [0.0] * 3, 1.0, [3000.0, 6000.0, 12000.0], om)
    _, irr = O.star_light(st, 0.0)
    assert irr.dtype == np.float32
    for i in range(3):
        expect = np.float32(om[i]) * O.black_body_bands(st["temperature"][i])[0]
        np.testing.assert_array_equal(irr[i], expect)
    # hotter stars peak bluer: 12000 K brightest in the first band, 3000 K in the last
    assert np.argmax(irr[2]) == 0 and np.argmax(irr[0]) == O.NB - 1

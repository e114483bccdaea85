"""CPU pins of the sensor-tail oracle (oracle_core: psf_kernel,
convolve_spatial) and of the host band table: closed forms, not retyped
formulas."""
import numpy as np
import pytest
from scipy import signal, special

import pyoracle as O
from vira_amd import sensor as SN


def _gauss_psf(sx, sy=None, angle=0.0, ss=20):
    return O.psf_from(SN.GaussianPSF(sx, sy, angle, ss).to_c())


def test_band_table_and_photon_energies():
    lo, hi, c, E = SN.band_table()
    assert np.allclose(lo * 1e9, 400 + 75 * np.arange(8), rtol=1e-6)
    assert np.allclose(c * 1e9, 437.5 + 75 * np.arange(8), rtol=1e-6)
    ref = 6.62607015e-34 * 299792458.0 / (c.astype(np.float64))
    assert np.max(np.abs(E - ref) / ref) < 3e-7
    qe = SN.default_quantum_efficiency()
    assert abs(np.mean(qe) - 0.5) < 1e-6 and np.all(np.diff(qe) < 0)  # proportional to photon energy


@pytest.mark.parametrize("sx,sy", [(1.3, 1.3), (0.8, 2.1)])
def test_gaussian_kernel_against_pixel_integrals(sx, sy):
    K = 15
    k = O.psf_kernel(_gauss_psf(sx, sy), K)
    c = (K - 1) / 2.0
    e = np.arange(K + 1) - 0.5 - c
    Fx = 0.5 * (1 + special.erf(e / (np.sqrt(2) * sx)))
    Fy = 0.5 * (1 + special.erf(e / (np.sqrt(2) * sy)))
    ref = np.outer(np.diff(Fy), np.diff(Fx))  # [y, x]
    got = k[:, :, 0].astype(np.float64)
    assert np.max(np.abs(got - ref)) <= 2e-4 * ref.max()  # 20 x 20 midpoint rule per pixel
    assert abs(got.sum() - 1.0) < 1e-3
    # every band carries the same sigma here
    assert np.array_equal(k[:, :, 0], k[:, :, 5])


def test_gaussian_rotation_by_90_degrees_transposes():
    K = 11
    a = O.psf_kernel(_gauss_psf(0.9, 2.0, 0.0), K)[:, :, 0]
    b = O.psf_kernel(_gauss_psf(0.9, 2.0, 90.0), K)[:, :, 0]
    assert np.max(np.abs(a - b.T)) <= 1e-5 * a.max()


def test_airy_kernel_against_bessel_closed_form():
    p = SN.AiryDiskPSF(focal_length=0.05, aperture_diameter=0.05 / 2.8, pixel_size=(4.8e-6, 4.8e-6), supersample=6)
    K = 9
    k = O.psf_kernel(O.psf_from(p.to_c()), K)
    _, _, wl, _ = SN.band_table()
    c = (K - 1) / 2.0
    s = (np.arange(6) + 0.5) / 6.0
    coef = np.pi * np.float64(np.float32(np.float32(p.aperture_diameter) / np.float32(p.focal_length)))
    for band in (0, 7):
        ref = np.zeros((K, K))
        for j in range(K):
            for i in range(K):
                px = (i + s[:, None]) - 0.5 - c
                py = (j + s[None, :]) - 0.5 - c
                r = np.hypot(px * 4.8e-6, py * 4.8e-6)
                x = coef * r / np.float64(wl[band])
                with np.errstate(invalid="ignore", divide="ignore"):
                    v = np.where(r < 1e-20, 1.0, (2 * special.j1(x) / x) ** 2)
                ref[j, i] = v.mean()
        got = k[:, :, band].astype(np.float64)
        assert np.max(np.abs(got - ref)) <= 1e-5 * ref.max()
    assert np.allclose(k[:, :, 0], k[:, :, 0][::-1, ::-1], rtol=1e-5, atol=1e-9)  # point symmetric


def test_convolution_against_scipy_and_identity():
    rng = np.random.default_rng(0)
    img = rng.random((23, 31, 8)).astype(np.float32)
    kern = O.psf_kernel(_gauss_psf(1.1, 1.7, 30.0), 7)
    got = O.convolve(img, kern)
    for b in (0, 4):
        ref = signal.correlate(img[:, :, b].astype(np.float64), kern[:, :, b].astype(np.float64), mode="same")
        assert np.max(np.abs(got[:, :, b] - ref)) <= 1e-5 * np.abs(ref).max()
    delta = np.zeros((5, 5, 8), np.float32)
    delta[2, 2, :] = 1.0
    assert np.array_equal(O.convolve(img, delta), img)

"""Camera sensor tail: spectral band table, point-spread functions and the
photosite/noise configuration, mirroring vira's camera components.

* band table and photon energies: ``initializeSpectralBands`` /
  ``PhotonEnergy`` (spectral_data.ipp:64-132, math.ipp:11-14), FP32 as there;
* ``GaussianPSF`` (gaussian_psf.ipp), ``AiryDiskPSF`` (airy_disk.ipp),
  kernels by ``PointSpreadFunction::makeKernel`` (psf.ipp);
* ``PhotositeConfig`` (photosite.hpp:19-42, photosite.ipp) and
  ``NoiseModelConfig`` (noise_model.hpp:16-25).
All compute runs in the HIP library (``Context.psf_kernel``, ``convolve``,
``simulate_sensor``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi

NUM_BINS = capi.NUM_BINS
PLANCK = np.float32(6.62607015e-34)      # math.hpp:35
LIGHT_SPEED = np.float32(299792458.0)    # math.hpp:32


def band_table(lo_nm: float = 400.0, hi_nm: float = 1000.0, n: int = NUM_BINS):
    """(min, max, centre) wavelengths in metres and photon energies in joules,
    computed in FP32 as initializeSpectralBands<N, UNIFORM_RANGE, lo, hi>."""
    f = np.float32
    step = (f(hi_nm) - f(lo_nm)) / f(n)
    bounds = np.array([f(lo_nm) + f(i) * step for i in range(n + 1)], np.float32)
    lo = bounds[:-1] * f(1e-9)
    hi = bounds[1:] * f(1e-9)
    centre = (lo + hi) / f(2)
    energy = (PLANCK * LIGHT_SPEED) / centre
    return lo, hi, centre.astype(np.float32), energy.astype(np.float32)


def default_quantum_efficiency() -> np.ndarray:
    """PhotositeConfig::defaultQE (photosite.ipp): 0.5 x E_k / mean(E)."""
    _, _, _, E = band_table()
    total = np.float32(0)
    for e in E:
        total = np.float32(total + e / np.float32(NUM_BINS))
    return (np.float32(0.5) * E / total).astype(np.float32)


@dataclass
class GaussianPSF:
    sigma_x: object = 1.0   # pixels, scalar or per band
    sigma_y: object = None
    angle_deg: float = 0.0
    supersample: int = 20

    def to_c(self) -> capi.PSF:
        p = capi.PSF()
        p.kind = 0
        sx = np.broadcast_to(np.asarray(self.sigma_x, np.float32), (NUM_BINS,))
        sy = sx if self.sigma_y is None else np.broadcast_to(np.asarray(self.sigma_y, np.float32), (NUM_BINS,))
        for k in range(NUM_BINS):
            p.sigma_x[k], p.sigma_y[k] = float(sx[k]), float(sy[k])
        p.angle_deg = float(self.angle_deg)
        p.supersample = int(self.supersample)
        return p


@dataclass
class AiryDiskPSF:
    focal_length: float        # metres
    aperture_diameter: float   # metres
    pixel_size: tuple          # metres (x, y)
    supersample: int = 20

    def to_c(self) -> capi.PSF:
        p = capi.PSF()
        p.kind = 1
        p.focal_length = float(self.focal_length)
        p.aperture_diameter = float(self.aperture_diameter)
        p.pixel_size[0], p.pixel_size[1] = float(self.pixel_size[0]), float(self.pixel_size[1])
        _, _, centre, _ = band_table()
        for k in range(NUM_BINS):
            p.wavelength[k] = float(centre[k])
        p.supersample = int(self.supersample)
        return p


@dataclass
class SensorConfig:
    """Photosite + noise model + exposure (defaults as the reference)."""
    exposure_time: float = 1.0
    quantum_efficiency: np.ndarray = field(default_factory=default_quantum_efficiency)
    gain: float = 0.333333         # ADU / electron (photosite.hpp:25)
    bit_depth: int = 8             # photosite.hpp:22
    linear_scale_factor: float = 1.0
    photon_noise: bool = False
    noise_model: bool = False
    dark_current: float = 0.0
    readout_noise_mean: float = 0.0
    readout_noise_std: float = 0.0
    horizontal_scale: float = 0.0
    vertical_scale: float = 0.0
    horizontal_pattern_period: int = 0
    vertical_pattern_period: int = 0
    seed: int = 0

    def max_adu(self) -> float:
        f = np.float32
        return float(f(self.linear_scale_factor) * (f(2.0) ** f(self.bit_depth) - f(1.0)))

    def to_c(self) -> capi.SensorSim:
        c = capi.SensorSim()
        _, _, _, E = band_table()
        qe = np.broadcast_to(np.asarray(self.quantum_efficiency, np.float32), (NUM_BINS,))
        c.exposure_time = float(self.exposure_time)
        for k in range(NUM_BINS):
            c.photon_energy[k] = float(E[k])
            c.quantum_efficiency[k] = float(qe[k])
        c.gain, c.max_adu = float(self.gain), self.max_adu()
        c.photon_noise, c.noise_model = int(self.photon_noise), int(self.noise_model)
        c.dark_current = float(self.dark_current)
        c.readout_noise_mean, c.readout_noise_std = float(self.readout_noise_mean), float(self.readout_noise_std)
        c.horizontal_scale, c.vertical_scale = float(self.horizontal_scale), float(self.vertical_scale)
        c.horizontal_pattern_period = int(self.horizontal_pattern_period)
        c.vertical_pattern_period = int(self.vertical_pattern_period)
        c.seed = int(self.seed)
        return c


def intrinsic_matrix(camera) -> np.ndarray:
    """K as Camera::initializeIntrinsicMatrix builds it (camera.ipp:1526-1557), FP32,
    GLM mat3x2 [col][row]: (fx, kyx, kxy, fy, px, py)."""
    f = np.float32
    sw, sh = camera.sensor_size()
    kx = f(camera.width) / f(sw)
    ky_ = f(camera.height) / f(sh)
    zd = f(camera.z_dir)
    ky = zd * ky_
    px = zd * (f(camera.width) / f(2))
    py = zd * (f(camera.height) / f(2))
    kxy = zd * f(0)
    kyx = zd * f(0)
    fx = f(camera.focal_length) * kx
    fy = f(camera.focal_length) * ky
    return np.array([fx, kyx, kxy, fy, px, py], np.float32)


def edge_max(kernel: np.ndarray) -> np.ndarray:
    """PointSpreadFunction::initKernels edge maximum (psf.ipp:143-159) of a kernel
    indexed [y][x]: the border pixel of largest spectral magnitude, visited as
    kernel(0, x), kernel(K-1, x), then kernel(y, 0), kernel(y, K-1) (Image(i, j)
    is pixel x = i, y = j); std::max keeps the first of equal magnitudes."""
    k = np.asarray(kernel, np.float32)
    K = k.shape[0]

    def mag(v):
        q = np.float32(0)
        for x in v:
            q = np.float32(q + np.float32(x * x))
        return np.float32(np.sqrt(q))

    best = np.zeros(NUM_BINS, np.float32)
    for x in range(K):
        for v in (k[x, 0], k[x, K - 1]):
            if mag(best) < mag(v):
                best = v.copy()
    for y in range(1, K - 1):
        for v in (k[0, y], k[K - 1, y]):
            if mag(best) < mag(v):
                best = v.copy()
    return best


def minimum_detectable_power(aperture_area: float, exposure_time: float = 0.01,
                             minimum_electrons: float = 10.0) -> float:
    """The unresolved renderer's cut (cpu_unresolved_renderer.ipp:56-57):
    Camera::computeMinimumDetectableIrradiance (camera.ipp:1201-1211), 10 electrons
    x the photon energies integrated over the bands / (area x exposure), times the
    exposure again, in FP32 as the reference."""
    f = np.float32
    lo, hi, _, E = band_table()
    total = f(0)
    for k in range(NUM_BINS):
        total = f(total + f(f(hi[k] - lo[k]) * E[k]))
    energy = f(f(minimum_electrons) * total)
    irradiance = f(energy / f(f(aperture_area) * f(exposure_time)))
    return float(f(irradiance * f(exposure_time)))


def bayer_mosaic(width: int, height: int, red, green, blue) -> np.ndarray:
    """bayerFilter (filter_arrays.ipp:22-50): (H, W, 8) float32 colour filter array;
    pixel (i, j) (column i, row j) takes red where i and j are even, blue where both
    are odd and green elsewhere.  red/green/blue: 8-band transmissions."""
    r, g, b = (np.broadcast_to(np.asarray(x, np.float32), (NUM_BINS,)) for x in (red, green, blue))
    i = np.arange(width)[None, :]
    j = np.arange(height)[:, None]
    even_i, even_j = (i % 2) == 0, (j % 2) == 0
    out = np.where((even_i & even_j)[..., None], r, np.where((~even_i & ~even_j)[..., None], b, g))
    return np.ascontiguousarray(out, np.float32)


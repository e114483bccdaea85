// sensor.hip -- received power per pixel (the last step of the render):
// Camera::calculateReceivedPower (camera.ipp:1175-1178)
//   P_k(i, j) = ((eta_k * A) * Omega_ij) * L_k(i, j)
// with Omega_ij the pixel's solid angle (Camera::initializePixelSolidAngle,
// camera.ipp:1594-1666): the four pixel-corner directions K^-1 (x, y, 1) in
// FP64 (the FP32 K^-1 widened, as camera.ipp:1566 does), normalised, and the
// two spherical triangles' areas by Girard's theorem (camera.ipp:1578-1592),
// rounded to FP32.
#include "context.hpp"
#include "scene_params.hpp"
#include "shading.hpp"

namespace vg {
namespace {

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 dcross(D3 a, D3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ double ddot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ D3 dnormalize(D3 a) {
    const double s = __ddiv_rn(1.0, __dsqrt_rn(ddot(a, a)));
    return {a.x * s, a.y * s, a.z * s};
}
// unit tangent at p0 of the great circle towards p1
__device__ __forceinline__ D3 tangent(D3 p0, D3 p1) { return dnormalize(dcross(dcross(p0, dsub(p1, p0)), p0)); }

__device__ __forceinline__ double triangle_solid_angle(D3 c0, D3 c1, D3 c2) {
    const double a0 = acos(ddot(tangent(c0, c1), tangent(c0, c2)));
    const double a1 = acos(ddot(tangent(c1, c0), tangent(c1, c2)));
    const double a2 = acos(ddot(tangent(c2, c0), tangent(c2, c1)));
    return ((a0 + a1) + a2) - 3.14159265358979323846;
}

__device__ __forceinline__ D3 corner(const double k[6], double zdir, double x, double y) {
    const double hx = (k[0] * x + k[2] * y) + k[4];
    const double hy = (k[1] * x + k[3] * y) + k[5];
    return dnormalize(D3{hx, zdir * hy, zdir});
}

// Pixel solid angles of a camera (Camera::initializePixelSolidAngle): computed
// once per camera geometry and kept on the context, as the reference keeps
// them on the camera after initialisation.
__global__ void k_pixel_solid_angle(int W, int H, CamF cam, float* __restrict__ omega_out) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (size_t)W * H) return;
    const int i = (int)(p % (size_t)W), j = (int)(p / (size_t)W);
    double k[6];
    for (int m = 0; m < 6; ++m) k[m] = (double)cam.kinv[m];
    const double zd = (double)cam.zdir;
    const D3 c0 = corner(k, zd, i, j), c1 = corner(k, zd, i + 1, j);
    const D3 c2 = corner(k, zd, i + 1, j + 1), c3 = corner(k, zd, i, j + 1);
    omega_out[p] = (float)(triangle_solid_angle(c0, c1, c2) + triangle_solid_angle(c0, c2, c3));
}

// P = ((eta * A) * Omega) * L per band (camera.ipp:1175-1178), a thread per
// pixel; VEC: the 8 bands as two float4 loads and stores (16-byte aligned images).
template <bool VEC>
__global__ void k_received_power(size_t n, vira_gpu_sensor sensor, const float* __restrict__ omega,
                                 const float* __restrict__ radiance, float* __restrict__ power) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float om = omega[p];
    float r[NB];
    if (VEC) {
        const float4 r0 = reinterpret_cast<const float4*>(radiance)[2 * p];
        const float4 r1 = reinterpret_cast<const float4*>(radiance)[2 * p + 1];
        r[0] = r0.x; r[1] = r0.y; r[2] = r0.z; r[3] = r0.w;
        r[4] = r1.x; r[5] = r1.y; r[6] = r1.z; r[7] = r1.w;
    } else {
#pragma unroll
        for (int b = 0; b < NB; ++b) r[b] = radiance[NB * p + b];
    }
    float o[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) o[b] = ((sensor.optical_efficiency[b] * sensor.aperture_area) * om) * r[b];
    if (VEC) {
        reinterpret_cast<float4*>(power)[2 * p] = make_float4(o[0], o[1], o[2], o[3]);
        reinterpret_cast<float4*>(power)[2 * p + 1] = make_float4(o[4], o[5], o[6], o[7]);
    } else {
#pragma unroll
        for (int b = 0; b < NB; ++b) power[NB * p + b] = o[b];
    }
}

// ------------------------------------------------------------ point-spread functions
// GaussianPSF::evaluate (gaussian_psf.ipp) / AiryDiskPSF::evaluate (airy_disk.ipp)
// at a kernel-plane offset (pixels) from the kernel centre.
__device__ __forceinline__ void psf_eval(const vira_gpu_psf& P, float cr, float sr, float px, float py,
                                         float v[NB]) {
    if (P.kind == 0) {
        const float rx = cr * px + sr * py;  // rotation_ * point, rotation_ = [[c, -s], [s, c]] (column-major)
        const float ry = -sr * px + cr * py;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const float xn = __fdiv_rn(rx, P.sigma_x[k]), yn = __fdiv_rn(ry, P.sigma_y[k]);
            const float ex = -0.5f * ((xn * xn) + (yn * yn));
            const float coef = __fdiv_rn(1.0f, ((2.0f * kPi) * P.sigma_x[k]) * P.sigma_y[k]);
            v[k] = coef * expf(ex);
        }
        return;
    }
    const float qx = px * P.pixel_size[0], qy = py * P.pixel_size[1];
    const double radius = (double)__builtin_sqrtf(qx * qx + qy * qy);  // length() of the FP32 vector
    if (radius < 1e-20) {
#pragma unroll
        for (int k = 0; k < NB; ++k) v[k] = 1.0f;
        return;
    }
    const double coefficient = 3.141592653589793 * (double)__fdiv_rn(P.aperture_diameter, P.focal_length);
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const double x = coefficient * radius / (double)P.wavelength[k];
        const float a = (float)(2.0 * j1(x) / x);
        v[k] = a * a;
    }
}

// PointSpreadFunction::makeKernel (psf.ipp): kernel pixel (x, y) is the mean of
// S x S sub-pixel evaluations, sample_i outer and sample_j inner.
__global__ void k_psf_kernel(vira_gpu_psf P, int K, float* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K * K) return;
    const int i = t % K, j = t / K;
    const int S = P.supersample;
    const float center = ((float)K - 1.0f) / 2.0f;
    const float ang = (3.14159265358979323846f / 180.0f) * P.angle_deg;
    const float cr = cosf(ang), sr = sinf(ang);
    float acc[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k] = 0.0f;
    for (int si = 0; si < S; ++si) {
        for (int sj = 0; sj < S; ++sj) {
            const float ox = __fdiv_rn((float)si + 0.5f, (float)S), oy = __fdiv_rn((float)sj + 0.5f, (float)S);
            const float spx = ((float)i + ox) - 0.5f, spy = ((float)j + oy) - 0.5f;
            float v[NB];
            psf_eval(P, cr, sr, spx - center, spy - center, v);
#pragma unroll
            for (int k = 0; k < NB; ++k) acc[k] = acc[k] + v[k];
        }
    }
    const float n = (float)(S * S);
#pragma unroll
    for (int k = 0; k < NB; ++k) out[(size_t)t * NB + k] = __fdiv_rn(acc[k], n);
}

// Image::convolveSpatial (image.ipp:1267-1330): zero padding, per-band sum in
// kernel order ky outer, kx inner.  The kernel row being summed is staged in
// LDS for the block.
constexpr int kConvBS = 256;
__global__ void __launch_bounds__(kConvBS) k_convolve(const float* __restrict__ in, int W, int H,
                                                     const float* __restrict__ kern, int K, float* __restrict__ out) {
    __shared__ float s_row[256 * NB];
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < (size_t)W * H;
    const int x = valid ? (int)(p % (size_t)W) : 0, y = valid ? (int)(p / (size_t)W) : 0;
    const int h = K / 2;
    float sum[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) sum[k] = 0.0f;
    for (int ky = 0; ky < K; ++ky) {
        for (int kx0 = 0; kx0 < K; kx0 += 256) {
            const int n = min(256, K - kx0);
            __syncthreads();
            for (int e = threadIdx.x; e < n * NB; e += kConvBS) s_row[e] = kern[((size_t)ky * K + kx0) * NB + e];
            __syncthreads();
            const int iy = y + ky - h;
            if (!valid || iy < 0 || iy >= H) continue;
            for (int kx = 0; kx < n; ++kx) {
                const int ix = x + kx0 + kx - h;
                if (ix < 0 || ix >= W) continue;
                const float* src = in + ((size_t)iy * W + ix) * NB;
#pragma unroll
                for (int k = 0; k < NB; ++k) sum[k] = sum[k] + src[k] * s_row[kx * NB + k];
            }
        }
    }
    if (!valid) return;
#pragma unroll
    for (int k = 0; k < NB; ++k) out[p * NB + k] = sum[k];
}

// Kernels of 256 pixels or more: the reference switches to an FFT there
// (Image::convolve, image.ipp:1006-1065), which is the same zero-padded
// correlation up to the transform's rounding.  Here it is computed directly
// with fused multiply-adds, tiled for the CU:
//   * one block per 64 x 32 output tile of one band (grid z = band);
//   * the kernel is walked in sub-blocks of at most 48 x 48 taps; for each,
//     the block stages the input window (64 + kb - 1) x (32 + kb - 1) of its
//     band in LDS (zeros outside the image) and the sub-block's taps;
//   * a thread owns 8 outputs of one column; it walks the taps column by
//     column and, down a column, keeps the 8 input values it needs in a
//     ring of registers, so each LDS read feeds 8 FMAs and a wave's reads
//     are 64 consecutive floats (no bank conflicts).
constexpr int kCvTW = 64, kCvTH = 32, kCvR = 8, kCvKB = 48;
constexpr int kCvRows = kCvTH / kCvR;  // 4 thread rows: 256 threads
constexpr int kCvSW = kCvTW + kCvKB - 1, kCvSH = kCvTH + kCvKB - 1;

__global__ void __launch_bounds__(kCvTW* kCvRows) k_convolve_tiled(const float* __restrict__ in, int W, int H,
                                                                  const float* __restrict__ kern, int K,
                                                                  float* __restrict__ out) {
    __shared__ float tile[kCvSH * kCvSW];
    __shared__ float taps[kCvKB * kCvKB];
    const int band = blockIdx.z;
    const int x0 = blockIdx.x * kCvTW, y0 = blockIdx.y * kCvTH;
    const int tx = threadIdx.x % kCvTW, ty = threadIdx.x / kCvTW;
    const int h = K / 2;
    float acc[kCvR];
#pragma unroll
    for (int r = 0; r < kCvR; ++r) acc[r] = 0.0f;
    for (int by = 0; by < K; by += kCvKB) {
        const int kbh = min(kCvKB, K - by);
        for (int bx = 0; bx < K; bx += kCvKB) {
            const int kbw = min(kCvKB, K - bx);
            const int sw = kCvTW + kbw - 1, sh = kCvTH + kbh - 1;
            __syncthreads();
            // input window: rows y0 + by - h .., columns x0 + bx - h ..
            for (int e = threadIdx.x; e < sw * sh; e += kCvTW * kCvRows) {
                const int r = e / sw, q = e - r * sw;
                const int iy = y0 + by - h + r, ix = x0 + bx - h + q;
                float v = 0.0f;
                if (iy >= 0 && iy < H && ix >= 0 && ix < W) v = in[((size_t)iy * W + ix) * NB + band];
                tile[r * kCvSW + q] = v;
            }
            for (int e = threadIdx.x; e < kbw * kbh; e += kCvTW * kCvRows) {
                const int r = e / kbw, q = e - r * kbw;
                taps[q * kCvKB + r] = kern[((size_t)(by + r) * K + (bx + q)) * NB + band];  // column-major
            }
            __syncthreads();
            const int row0 = ty * kCvR;
            for (int kx = 0; kx < kbw; ++kx) {
                const float* col = tile + tx + kx;
                const float* tcol = taps + kx * kCvKB;
                float ring[kCvR];
#pragma unroll
                for (int r = 0; r < kCvR; ++r) ring[r] = col[(row0 + r) * kCvSW];
                int ky = 0;
                for (; ky + kCvR <= kbh; ky += kCvR) {
#pragma unroll
                    for (int s = 0; s < kCvR; ++s) {
                        const float kv = tcol[ky + s];
#pragma unroll
                        for (int r = 0; r < kCvR; ++r) acc[r] = fmaf(ring[(r + s) % kCvR], kv, acc[r]);
                        // slot s held row row0 + ky + s, which only r = 0 needed: refill it with the row
                        // r = 7 needs at the next tap (none after the column's last tap)
                        if (ky + s + 1 < kbh) ring[s] = col[(row0 + kCvR + ky + s) * kCvSW];
                    }
                }
                // the last kbh % 8 taps of the column (the ring is rotated by ky % 8 == 0 here)
                for (int s = 0; ky + s < kbh; ++s) {
                    const float kv = tcol[ky + s];
#pragma unroll
                    for (int r = 0; r < kCvR; ++r) acc[r] = fmaf(col[(row0 + r + ky + s) * kCvSW], kv, acc[r]);
                }
            }
        }
    }
    const int ox = x0 + tx;
    if (ox >= W) return;
#pragma unroll
    for (int r = 0; r < kCvR; ++r) {
        const int oy = y0 + ty * kCvR + r;
        if (oy < H) out[((size_t)oy * W + ox) * NB + band] = acc[r];
    }
}

// ------------------------------------------------------------ sensor
// Counter-based draws: stream (seed, pixel, iteration, dim).
constexpr uint32_t kDimPhoton = 64, kDimDark = 80, kDimReadout = 96;

// Poisson(lambda): multiplication method below 10, PTRS (Hoermann 1993,
// "The transformed rejection method for generating Poisson random
// variables") at and above (camera.ipp:1306-1316 draws from
// std::poisson_distribution); at most 64 proposals (the mean is returned in
// the vanishingly rare case that none is accepted).
//
// The acceptance test compares log(V alpha / (a / us^2 + b)) with the
// log-probability -lambda + k log(lambda) - lgamma(k + 1), whose three terms
// reach ~1e7 at large means: summed in FP32 they keep only ~1 of their value.
// For k >= 10 it is evaluated in the Stirling form
//   -D - log(2 pi k) / 2 - 1 / (12 k) + 1 / (360 k^3)
// (the series' next term is below 1e-9 for k >= 10), with the deviance
//   D = k log1p(x) - (k - lambda) = lambda phi(x),  x = (k - lambda) / lambda,
//   phi(x) = (1 + x) log1p(x) - x = sum_{n>=2} (-1)^n x^n / (n (n - 1)),
// k - lambda exact (two floats within a factor of 2) and phi taken from its
// series to x^9 for |x| < 0.1 (relative truncation below 1e-9), where the
// closed form would cancel; FP32 then holds the log-probability to ~2e-6
// within 3 standard deviations of any mean from 10 to 1e7 and ~1e-4 beyond.
// Below 10 the terms are small and lgammaf is used.  The proposal's floor
// takes lambda in FP64, where FP32's spacing (1/16 at 1e6) would move it.
__device__ __forceinline__ float poisson(float lambda, uint32_t seed, uint32_t pix, uint32_t dim) {
    if (!(lambda > 0.0f)) return 0.0f;
    if (lambda < 10.0f) {
        const float L = expf(-lambda);
        float prod = 1.0f;
        for (uint32_t k = 0; k < 256u; ++k) {
            prod = prod * rng_u01(seed, pix, k, dim);
            if (prod <= L) return (float)k;
        }
        return lambda;
    }
    const float slam = sqrtf(lambda);
    const float b = 0.931f + 2.53f * slam;
    const float a = -0.059f + 0.02483f * b;
    const float vr = 0.9277f - __fdiv_rn(3.6224f, b - 2.0f);
    for (uint32_t it = 0; it < 64u; ++it) {
        const float U = rng_u01(seed, pix, 2u * it, dim) - 0.5f;
        const float V = rng_u01(seed, pix, 2u * it + 1u, dim);
        const float us = 0.5f - fabsf(U);
        const float k = (float)floor((double)((__fdiv_rn(2.0f * a, us) + b) * U) + (double)lambda + 0.43);
        if (us >= 0.07f && V <= vr) return k;
        if (k < 0.0f || (us < 0.013f && V > us)) continue;
        const float invalpha = 1.1239f + __fdiv_rn(1.1328f, b - 3.4f);
        const float lhs = (logf(V) + logf(invalpha)) - logf(__fdiv_rn(a, us * us) + b);
        float rhs;
        if (k < 10.0f) {
            rhs = (-lambda + k * logf(lambda)) - lgammaf(k + 1.0f);
        } else {
            const float dk = k - lambda;
            const float x = __fdiv_rn(dk, lambda);
            float D;
            if (fabsf(x) < 0.1f) {
                float p = -1.0f / 72.0f;
                p = fmaf(p, x, 1.0f / 56.0f);
                p = fmaf(p, x, -1.0f / 42.0f);
                p = fmaf(p, x, 1.0f / 30.0f);
                p = fmaf(p, x, -1.0f / 20.0f);
                p = fmaf(p, x, 1.0f / 12.0f);
                p = fmaf(p, x, -1.0f / 6.0f);
                p = fmaf(p, x, 0.5f);
                D = lambda * ((x * x) * p);
            } else {
                D = k * log1pf(x) - dk;
            }
            const float ik = __fdiv_rn(1.0f, k);
            rhs = ((-D - 0.5f * logf(6.283185307179586f * k)) - ik * (1.0f / 12.0f)) + (ik * ik * ik) * (1.0f / 360.0f);
        }
        if (lhs <= rhs) return k;
    }
    return lambda;
}

// Normal(mean, std) by Box-Muller on two counter draws.
__device__ __forceinline__ float normal(float mean, float sd, uint32_t seed, uint32_t pix, uint32_t dim) {
    const float u1 = 1.0f - rng_u01(seed, pix, 0u, dim);  // (0, 1]
    const float u2 = rng_u01(seed, pix, 1u, dim);
    return mean + sd * (sqrtf(-2.0f * logf(u1)) * cosf((2.0f * kPi) * u2));
}

__global__ void k_simulate_sensor(vira_gpu_sensor_sim c, const float* __restrict__ power, int W, int H,
                                  float* __restrict__ photons, float* __restrict__ image) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (size_t)W * H) return;
    const uint32_t pix = (uint32_t)p;
    // getPhotonCounts (camera.ipp:1281-1330): energy = t P (x the filter mosaic), photons = energy / E_photon
    float cnt[NB];
    float total = 0.0f;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        float energy = c.exposure_time * power[p * NB + k];
        if (c.filter_mosaic_dev) energy = energy * c.filter_mosaic_dev[p * NB + k];  // camera.ipp:1296
        cnt[k] = __fdiv_rn(energy, c.photon_energy[k]);
        total = total + cnt[k];
    }
    if (c.photon_noise && total >= 1.0f) {
#pragma unroll
        for (int k = 0; k < NB; ++k) cnt[k] = poisson(cnt[k], c.seed, pix, kDimPhoton + (uint32_t)k);
    }
    if (photons) {
#pragma unroll
        for (int k = 0; k < NB; ++k) photons[p * NB + k] = cnt[k];
    }
    // NoiseModel::simulate (noise_model.ipp): fixed pattern x (dark + readout)
    float noise = 0.0f;
    if (c.noise_model) {
        const int i = (int)(p % (size_t)W), j = (int)(p / (size_t)W);
        const float dark = poisson(c.dark_current, c.seed, pix, kDimDark) * c.exposure_time;
        const float readout = normal(c.readout_noise_mean, c.readout_noise_std, c.seed, pix, kDimReadout);
        float scale = 1.0f;
        if (c.horizontal_pattern_period != 0)
            scale = scale * ((1.0f - c.horizontal_scale) +
                             c.horizontal_scale * (1.0f + sinf(__fdiv_rn((2.0f * kPi) * (float)i,
                                                                          2.0f * (float)c.horizontal_pattern_period))));
        if (c.vertical_pattern_period != 0)
            scale = scale * ((1.0f - c.vertical_scale) +
                             c.vertical_scale * (1.0f + sinf(__fdiv_rn((2.0f * kPi) * (float)j,
                                                                        2.0f * (float)c.vertical_pattern_period))));
        noise = scale * (dark + readout);
    }
    // Photosite::exposePixel (photosite.ipp): electrons = sum QE photons + noise
    float e = 0.0f;
#pragma unroll
    for (int k = 0; k < NB; ++k) e = e + c.quantum_efficiency[k] * cnt[k];
    e = e + noise;
    const float amp = fminf(c.gain * e, c.max_adu);
    if (image) image[p] = __fdiv_rn(amp, c.max_adu);
}

// ------------------------------------------------------------ unresolved point sources
// Deterministic splat: the reference adds each source's response to the image
// in source order (cpu_unresolved_renderer.ipp:176-199), so a pixel's value is
// ((0 + v_s1) + v_s2) + ... over its sources s1 < s2 < ...  Here the sources
// are projected once (k_star_project), binned to 32x32-pixel tiles
// (k_star_bin), and each tile sorts its list by source index in LDS and sums
// per pixel in that order (k_star_splat): the same additions in the same order.
constexpr int kStarTile = 32;
constexpr int kStarSplatBS = 256;
constexpr int kStarListCap = 4096;  // sources per tile sorted in LDS; busier tiles scan every source
constexpr int kStarAssign = 4;      // StarRec::sel of a source assigned to one pixel (camera without PSF)

struct StarRec {
    float X, Y;          // projected position (pixels)
    int32_t sel;         // chosen kernel, kStarAssign without PSF, -1 if the source adds nothing
    int32_t x0, y0, x1, y1;  // footprint clipped to the image, [x0, x1) x [y0, y1)
    int32_t pad;
    float P[NB];         // received power
};
static_assert(sizeof(StarRec) == 64, "StarRec is one 64-byte record");

__device__ __forceinline__ float magnitude8(const float v[NB]) {
    float q = 0.0f;
#pragma unroll
    for (int k = 0; k < NB; ++k) q = q + v[k] * v[k];
    return __fsqrt_rn(q);
}

// One thread per source: camera frame, projection, received power, kernel
// choice and footprint; counts the 32x32 tiles the footprint touches.
__global__ void __launch_bounds__(256) k_star_project(CamF cam, vira_gpu_sensor sen, vira_gpu_star_psf psf,
                                                      const float* __restrict__ dirs, const float* __restrict__ irr,
                                                      uint32_t n, int tiles_x, StarRec* __restrict__ recs,
                                                      uint32_t* __restrict__ tile_count) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    StarRec r;
    r.X = r.Y = 0.0f;
    r.sel = -1;
    r.x0 = r.y0 = r.x1 = r.y1 = 0;
    r.pad = 0;
#pragma unroll
    for (int k = 0; k < NB; ++k) r.P[k] = 0.0f;
    const float dx = dirs[3 * (size_t)s], dy = dirs[3 * (size_t)s + 1], dz = dirs[3 * (size_t)s + 2];
    // camera_view_normal * direction (cpu_unresolved_renderer.ipp:104): R^T d
    const float cx = (cam.rot[0] * dx + cam.rot[3] * dy) + cam.rot[6] * dz;
    const float cy = (cam.rot[1] * dx + cam.rot[4] * dy) + cam.rot[7] * dz;
    const float cz = (cam.rot[2] * dx + cam.rot[5] * dy) + cam.rot[8] * dz;
    bool live = !(psf.cull_behind && !(cz > 0.0f));
    // Camera::projectCameraPoint (camera.ipp:1062-1081): K * (x/z, y/z, 1)
    const float hx = __fdiv_rn(cx, cz), hy = __fdiv_rn(cy, cz);
    const float X = (psf.intrinsic[0] * hx + psf.intrinsic[2] * hy) + psf.intrinsic[4];
    const float Y = (psf.intrinsic[1] * hx + psf.intrinsic[3] * hy) + psf.intrinsic[5];
    const float radius = 41.0f;  // findPointsInRegion (cpu_unresolved_renderer.ipp:165)
    live = live && (X + radius) >= 0.0f && (X - radius) < (float)cam.W && (Y + radius) >= 0.0f &&
           (Y - radius) < (float)cam.H;
    if (live) {
        float P[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) P[k] = (sen.optical_efficiency[k] * sen.aperture_area) * irr[(size_t)NB * s + k];
        live = magnitude8(P) > psf.minimum_power;
        if (live) {
            // no PSF: the source's power is assigned to the pixel (int) X, (int) Y
            // (processRegion, cpu_unresolved_renderer.ipp:191-196); else the first
            // kernel whose edge response is below the minimum (PointSpreadFunction::
            // getKernel, psf.ipp:89-110) and Image::addImage's footprint
            // (image.ipp:1082-1098)
            int sel = kStarAssign, x0, y0, x1, y1;
            if (psf.count == 0) {
                x0 = (int)X;
                y0 = (int)Y;
                x1 = x0 + 1;
                y1 = y0 + 1;
                live = x0 >= 0 && x0 < cam.W && y0 >= 0 && y0 < cam.H;
            } else {
                sel = 0;
                for (int i = 0; i < psf.count; ++i) {
                    sel = i;
                    float e[NB];
#pragma unroll
                    for (int k = 0; k < NB; ++k) e[k] = psf.edge_max[i][k] * P[k];
                    if (magnitude8(e) < psf.minimum_power) break;
                }
                const int K = psf.size[sel];
                const int ix = (int)floorf(X), iy = (int)floorf(Y);
                const int h = (int)((float)K / 2.0f);
                x0 = max(0, ix - h);
                y0 = max(0, iy - h);
                x1 = min(cam.W, ix + h + 1);
                y1 = min(cam.H, iy + h + 1);
                live = x0 < x1 && y0 < y1;
            }
            if (live) {
                r.X = X;
                r.Y = Y;
                r.sel = sel;
                r.x0 = x0;
                r.y0 = y0;
                r.x1 = x1;
                r.y1 = y1;
#pragma unroll
                for (int k = 0; k < NB; ++k) r.P[k] = P[k];
                for (int ty = y0 / kStarTile; ty <= (y1 - 1) / kStarTile; ++ty)
                    for (int tx = x0 / kStarTile; tx <= (x1 - 1) / kStarTile; ++tx)
                        atomicAdd(tile_count + (size_t)ty * tiles_x + tx, 1u);
            }
        }
    }
    recs[s] = r;
}

// One thread per source: its index into the lists of the tiles it touches
// (order within a tile is arbitrary here and fixed by the splat's sort).
__global__ void __launch_bounds__(256) k_star_bin(const StarRec* __restrict__ recs, uint32_t n, int tiles_x,
                                                  const uint32_t* __restrict__ tile_start,
                                                  uint32_t* __restrict__ tile_fill, uint32_t* __restrict__ list) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const StarRec& r = recs[s];
    if (r.sel < 0) return;
    for (int ty = r.y0 / kStarTile; ty <= (r.y1 - 1) / kStarTile; ++ty)
        for (int tx = r.x0 / kStarTile; tx <= (r.x1 - 1) / kStarTile; ++tx) {
            const size_t t = (size_t)ty * tiles_x + tx;
            const uint32_t slot = atomicAdd(tile_fill + t, 1u);
            list[tile_start[t] + slot] = s;
        }
}

// The bilinear-sampled response of one source at pixel (i, j), added to acc
// (Image::addImage, image.ipp:1099-1150, with received_power * kernel formed
// per kernel pixel as PointSpreadFunction::getResponse does).
__device__ __forceinline__ void star_add(const StarRec& r, const vira_gpu_star_psf& psf, int i, int j,
                                         float acc[NB]) {
    const int K = psf.size[r.sel];
    const float* __restrict__ kern = psf.kernel_dev[r.sel];
    const float rx = (float)K / 2.0f, ry = (float)K / 2.0f;
    const float kx = (float)i - (r.X - rx), ky = (float)j - (r.Y - ry);
    const int kx0 = (int)floorf(kx), ky0 = (int)floorf(ky);
    if (!(kx0 >= 0 && kx0 < K && ky0 >= 0 && ky0 < K)) return;
    const int kx1 = kx0 + 1, ky1 = ky0 + 1;
    const float wx1 = kx - (float)kx0, wy1 = ky - (float)ky0;
    const float wx0 = 1.0f - wx1, wy0 = 1.0f - wy1;
    const float* k00 = kern + ((size_t)kx0 + (size_t)ky0 * K) * NB;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const float P = r.P[k];
        float v = 0.0f;
        if (kx1 < K && ky1 < K) {
            v = v + ((P * k00[k]) * wx0) * wy0;
            v = v + ((P * k00[NB + k]) * wx1) * wy0;
            v = v + ((P * k00[(size_t)K * NB + k]) * wx0) * wy1;
            v = v + ((P * k00[(size_t)(K + 1) * NB + k]) * wx1) * wy1;
        } else if (kx1 < K) {
            v = v + (P * k00[k]) * wx0;
            v = v + (P * k00[NB + k]) * wx1;
        } else if (ky1 < K) {
            v = v + (P * k00[k]) * wy0;
            v = v + (P * k00[(size_t)K * NB + k]) * wy1;
        } else {
            v = P * k00[k];
        }
        acc[k] = acc[k] + v;
    }
}

__device__ __forceinline__ bool star_covers(const StarRec& r, int i, int j) {
    return r.sel >= 0 && i >= r.x0 && i < r.x1 && j >= r.y0 && j < r.y1;
}

// One block per 32x32 tile: sort the tile's sources by index (bitonic, in
// LDS), then each pixel adds its sources' responses in that order.
__global__ void __launch_bounds__(kStarSplatBS) k_star_splat(vira_gpu_star_psf psf, const StarRec* __restrict__ recs,
                                                             uint32_t n, int W, int H, int tiles_x,
                                                             const uint32_t* __restrict__ tile_count,
                                                             const uint32_t* __restrict__ tile_start,
                                                             const uint32_t* __restrict__ list,
                                                             float* __restrict__ img) {
    __shared__ uint32_t keys[kStarListCap];
    const uint32_t t = blockIdx.x;
    const uint32_t cnt = tile_count[t];
    if (cnt == 0) return;
    const int tx = (int)(t % (uint32_t)tiles_x), ty = (int)(t / (uint32_t)tiles_x);
    const bool sorted = cnt <= (uint32_t)kStarListCap;
    if (sorted) {
        uint32_t m = 1;
        while (m < cnt) m <<= 1;
        for (uint32_t q = threadIdx.x; q < m; q += kStarSplatBS)
            keys[q] = q < cnt ? list[tile_start[t] + q] : 0xffffffffu;
        __syncthreads();
        for (uint32_t size = 2; size <= m; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t q = threadIdx.x; q < m; q += kStarSplatBS) {
                    const uint32_t o = q ^ stride;
                    if (o > q) {
                        const uint32_t a = keys[q], b = keys[o];
                        const bool up = (q & size) == 0;
                        if ((a > b) == up) {
                            keys[q] = b;
                            keys[o] = a;
                        }
                    }
                }
                __syncthreads();
            }
    }
    for (int lp = threadIdx.x; lp < kStarTile * kStarTile; lp += kStarSplatBS) {
        const int i = tx * kStarTile + (lp & (kStarTile - 1)), j = ty * kStarTile + lp / kStarTile;
        if (i >= W || j >= H) continue;
        float acc[NB];
        bool touched = false;
        float* dst = img + ((size_t)j * W + i) * NB;
        const uint32_t m = sorted ? cnt : n;
        for (uint32_t e = 0; e < m; ++e) {
            const uint32_t s = sorted ? keys[e] : e;
            const StarRec& r = recs[s];
            if (!star_covers(r, i, j)) continue;
            if (!touched) {
#pragma unroll
                for (int k = 0; k < NB; ++k) acc[k] = dst[k];
                touched = true;
            }
            if (r.sel == kStarAssign) {
#pragma unroll
                for (int k = 0; k < NB; ++k) acc[k] = r.P[k];
            } else {
                star_add(r, psf, i, j, acc);
            }
        }
        if (touched) {
#pragma unroll
            for (int k = 0; k < NB; ++k) dst[k] = acc[k];
        }
    }
}

}  // namespace

void render_stars(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_sensor& sen, const vira_gpu_star_psf& psf,
                  const float* dirs, const float* irr, size_t n, float* img) {
    VG_CHECK(cam.width > 0 && cam.height > 0, VIRA_GPU_ERR_INVALID, "render_stars: empty image");
    VG_CHECK(img && (n == 0 || (dirs && irr)), VIRA_GPU_ERR_INVALID, "render_stars: null buffers");
    VG_CHECK(psf.count >= 0 && psf.count <= 4, VIRA_GPU_ERR_INVALID, "render_stars: 0..4 PSF kernels");
    for (int i = 0; i < psf.count; ++i)
        VG_CHECK(psf.size[i] >= 1 && psf.size[i] % 2 == 1 && psf.kernel_dev[i], VIRA_GPU_ERR_INVALID,
                 "render_stars: kernels need odd sizes and device data");
    VG_CHECK(psf.cull_behind == 0 || psf.cull_behind == 1, VIRA_GPU_ERR_INVALID, "render_stars: cull_behind 0 or 1");
    VG_CHECK(n < (1u << 30), VIRA_GPU_ERR_INVALID, "render_stars: too many sources");
    if (n == 0) return;
    const CamF cf = cam_params(cam);
    const int tiles_x = (cam.width + kStarTile - 1) / kStarTile, tiles_y = (cam.height + kStarTile - 1) / kStarTile;
    const size_t tiles = (size_t)tiles_x * tiles_y;
    c.st_rec.alloc(n * sizeof(StarRec));
    c.st_count.alloc(tiles * sizeof(uint32_t));
    c.st_start.alloc(tiles * sizeof(uint32_t));
    c.st_fill.alloc(tiles * sizeof(uint32_t));
    auto* recs = (StarRec*)c.st_rec.p;
    auto* cnt = (uint32_t*)c.st_count.p;
    auto* start = (uint32_t*)c.st_start.p;
    auto* fill = (uint32_t*)c.st_fill.p;
    VG_HIP(hipMemsetAsync(cnt, 0, tiles * sizeof(uint32_t), c.stream));
    VG_HIP(hipMemsetAsync(fill, 0, tiles * sizeof(uint32_t), c.stream));
    k_star_project<<<blocks_for(n, 256), 256, 0, c.stream>>>(cf, sen, psf, dirs, irr, (uint32_t)n, tiles_x, recs,
                                                             cnt);
    VG_HIP(hipGetLastError());
    const uint32_t total = scan_exclusive_u32(c, cnt, start, tiles);
    if (total == 0) return;
    c.st_list.alloc((size_t)total * sizeof(uint32_t));
    auto* list = (uint32_t*)c.st_list.p;
    k_star_bin<<<blocks_for(n, 256), 256, 0, c.stream>>>(recs, (uint32_t)n, tiles_x, start, fill, list);
    VG_HIP(hipGetLastError());
    k_star_splat<<<(unsigned)tiles, kStarSplatBS, 0, c.stream>>>(psf, recs, (uint32_t)n, cam.width, cam.height,
                                                                 tiles_x, cnt, start, list, img);
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(c.stream));
}

void psf_kernel(Ctx& c, const vira_gpu_psf& P, int K, float* out) {
    VG_CHECK(K >= 3 && K % 2 == 1 && K <= 1023, VIRA_GPU_ERR_INVALID, "psf_kernel: size must be odd, 3..1023");
    VG_CHECK(P.kind == 0 || P.kind == 1, VIRA_GPU_ERR_INVALID, "psf_kernel: kind 0 (Gaussian) or 1 (Airy disk)");
    VG_CHECK(P.supersample >= 1 && P.supersample <= 64, VIRA_GPU_ERR_INVALID, "psf_kernel: supersample 1..64");
    VG_CHECK(out, VIRA_GPU_ERR_INVALID, "psf_kernel: null output");
    if (P.kind == 0)
        for (int k = 0; k < NB; ++k)
            VG_CHECK(P.sigma_x[k] > 0 && P.sigma_y[k] > 0, VIRA_GPU_ERR_INVALID, "psf_kernel: sigmas must be positive");
    else
        VG_CHECK(P.focal_length > 0 && P.aperture_diameter > 0 && P.pixel_size[0] > 0 && P.pixel_size[1] > 0,
                 VIRA_GPU_ERR_INVALID, "psf_kernel: Airy disk needs focal length, aperture and pixel size");
    k_psf_kernel<<<blocks_for((size_t)K * K, 64), 64, 0, c.stream>>>(P, K, out);
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(c.stream));
}

void convolve(Ctx& c, const float* in, int W, int H, const float* kern, int K, float* out) {
    VG_CHECK(in && kern && out && in != out, VIRA_GPU_ERR_INVALID, "convolve: null or aliased buffers");
    VG_CHECK(W > 0 && H > 0 && K >= 1 && K % 2 == 1, VIRA_GPU_ERR_INVALID, "convolve: bad sizes");
    const size_t n = (size_t)W * H;
    if ((size_t)K * K < 256) {
        // the reference's spatial branch: per band, ky outer, kx inner, separate multiply and add
        k_convolve<<<blocks_for(n, kConvBS), kConvBS, 0, c.stream>>>(in, W, H, kern, K, out);
    } else if (fft_convolve_fits(W, H, K)) {
        fft_convolve(c, in, W, H, kern, K, out);  // the reference's FFT branch (image.ipp:1014-1065)
        return;
    } else {
        // images past the longest transform: the direct sums
        const dim3 grid((W + kCvTW - 1) / kCvTW, (H + kCvTH - 1) / kCvTH, NB);
        k_convolve_tiled<<<grid, kCvTW * kCvRows, 0, c.stream>>>(in, W, H, kern, K, out);
    }
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(c.stream));
}

void simulate_sensor(Ctx& c, const vira_gpu_sensor_sim& cfg, const float* power, int W, int H, float* photons,
                     float* image) {
    VG_CHECK(power && W > 0 && H > 0, VIRA_GPU_ERR_INVALID, "simulate_sensor: bad arguments");
    VG_CHECK(cfg.max_adu > 0 && cfg.exposure_time >= 0, VIRA_GPU_ERR_INVALID, "simulate_sensor: bad sensor config");
    for (int k = 0; k < NB; ++k)
        VG_CHECK(cfg.photon_energy[k] > 0, VIRA_GPU_ERR_INVALID, "simulate_sensor: photon energies must be positive");
    const size_t n = (size_t)W * H;
    k_simulate_sensor<<<blocks_for(n, 256), 256, 0, c.stream>>>(cfg, power, W, H, photons, image);
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(c.stream));
}

void received_power(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_sensor& s, const float* radiance,
                    float* power, float* omega) {
    VG_CHECK(cam.width > 0 && cam.height > 0, VIRA_GPU_ERR_INVALID, "received_power: empty image");
    VG_CHECK(!power || radiance, VIRA_GPU_ERR_INVALID, "received_power: power needs the radiance image");
    const size_t n = (size_t)cam.width * cam.height;
    CamF cf;
    cf.W = cam.width;
    cf.H = cam.height;
    for (int k = 0; k < 6; ++k) cf.kinv[k] = cam.kinv[k];
    cf.zdir = cam.z_dir;
    // the solid angles depend only on the size, K^-1 and the z sign: recomputed when those change
    std::vector<float> key = {(float)cam.width, (float)cam.height, cam.z_dir};
    key.insert(key.end(), cam.kinv, cam.kinv + 6);
    if (!c.rp_omega.p || c.rp_key != key) {
        c.rp_key.clear();
        c.rp_omega.alloc(n * sizeof(float));
        k_pixel_solid_angle<<<blocks_for(n, 256), 256, 0, c.stream>>>(cam.width, cam.height, cf,
                                                                        c.rp_omega.as<float>());
        VG_HIP(hipGetLastError());
        c.rp_key = key;
    }
    if (omega)
        VG_HIP(hipMemcpyAsync(omega, c.rp_omega.p, n * sizeof(float), hipMemcpyDeviceToDevice, c.stream));
    if (power) {
        if ((((uintptr_t)power | (uintptr_t)radiance) & 15) == 0)
            k_received_power<true><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, s, c.rp_omega.as<float>(), radiance,
                                                                             power);
        else
            k_received_power<false><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, s, c.rp_omega.as<float>(),
                                                                              radiance, power);
        VG_HIP(hipGetLastError());
    }
    VG_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace vg

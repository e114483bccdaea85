// catalogue.hip -- star catalogues: StarCatalogue::makeStarLight
// (star_catalogue.ipp:55-78) on the GPU, one thread per star, and the star
// quipu (.qsc) reader (StarQuipu::read, star_quipu.ipp:20-29).
//
// Star::getUnitVector (star.ipp:19-32): dt = et / (365.25 * 86400);
//   delta = (float)(DEm + pmRA dt), alpha = (float)(RAm + pmDE dt)
// -- the reference adds the right-ascension rate to the declination and the
// declination rate to the right ascension; that is kept, as every caller of
// the reference sees it -- and (cos d cos a, cos d sin a, sin d) in FP32.
// Star::getIrradiance (star.ipp:13-16): (float)omega * blackBodyRadiance(T, 100)
// per band, where blackBodyRadiance (spectral_data.ipp:613-629) integrates
// Planck's law (math.ipp:45-59, FP64) over the band's [min, max] wavelength
// (the FP32 band edges widened) at linspace(min, max, 100) points
// (math.ipp:33-42) with the trapezoid rule (math.ipp:131-172), summed in
// order and rounded to FP32.
#include <cstdio>
#include <cstring>

#include "context.hpp"

namespace vg {
namespace {

constexpr int kPlanckSteps = 100;

// the spectrum's band edges: linspace of the nm bounds in FP32, then * 1e-9f
// (spectral_data.ipp:64-125)
__host__ __device__ inline float band_edge_m(int i) {
    const float lo = 400.0f, hi = 1000.0f;
    const float step = (hi - lo) / (float)NB;
    return (lo + (float)i * step) * 1e-9f;
}

constexpr double kC = 299792458.0, kH = 6.62607015e-34, kKB = 1.380649e-23;
constexpr int kBlock = 256;                // 32 stars x 8 bands
constexpr int kPoints = NB * kPlanckSteps;

// One thread per (star, band).  The block first tabulates, per band and
// integration point, the wavelength x, Planck's star-independent factor
// 2hc^2 / pow(x, 5) and x kB in LDS (the same FP64 operations, in the same
// order, as plancksLaw, math.ipp:45-59); each thread then evaluates
// coeff * 1. / (exp(hc / (x kB T)) - 1.) at its band's 100 points and sums the
// trapezoids in order.
__global__ __launch_bounds__(kBlock) void k_star_light(const vira_gpu_star* __restrict__ stars, size_t n, double et,
                                                     float* __restrict__ dirs, float* __restrict__ irr) {
    __shared__ double s_x[kPoints], s_coeff[kPoints], s_lk[kPoints];
    for (int j = threadIdx.x; j < kPoints; j += kBlock) {
        const int k = j / kPlanckSteps, i = j - k * kPlanckSteps;
        const double lo = (double)band_edge_m(k), hi = (double)band_edge_m(k + 1);
        const double step = (hi - lo) / (double)(kPlanckSteps - 1);
        const double x = lo + ((double)i * step);
        s_x[j] = x;
        s_coeff[j] = 2 * kH * kC * kC / pow(x, 5.0);
        s_lk[j] = x * kKB;
    }
    __syncthreads();
    const size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t s = t / NB;
    const int k = (int)(t % NB);
    if (s >= n) return;
    const double T = stars[s].temperature;
    if (k == 0) {
        const double dt = et / (365.25 * 86400.0);
        const float delta = (float)(stars[s].de + stars[s].pm_ra * dt);
        const float alpha = (float)(stars[s].ra + stars[s].pm_de * dt);
        dirs[3 * s] = cosf(delta) * cosf(alpha);
        dirs[3 * s + 1] = cosf(delta) * sinf(alpha);
        dirs[3 * s + 2] = sinf(delta);
    }
    const int base = k * kPlanckSteps;
    double x0 = s_x[base];
    double y0 = s_coeff[base] * 1. / (exp(kH * kC / (s_lk[base] * T)) - 1.);
    double area = 0.0;
    for (int i = 1; i < kPlanckSteps; ++i) {
        const double x1 = s_x[base + i];
        const double y1 = s_coeff[base + i] * 1. / (exp(kH * kC / (s_lk[base + i] * T)) - 1.);
        area += (x1 - x0) * (y0 + y1) / 2.;
        x0 = x1;
        y0 = y1;
    }
    irr[t] = (float)stars[s].omega * (float)area;
}

}  // namespace

void star_light(Ctx& c, const vira_gpu_star* stars, size_t n, double et, float* dirs, float* irr) {
    VG_CHECK(n == 0 || (stars && dirs && irr), VIRA_GPU_ERR_INVALID, "star_light: null buffers");
    if (n == 0) return;
    VG_CHECK(((uintptr_t)stars & 7) == 0 && ((uintptr_t)dirs & 3) == 0 && ((uintptr_t)irr & 3) == 0,
             VIRA_GPU_ERR_INVALID, "star_light: star records must be 8-byte aligned, outputs 4-byte aligned");
    VG_CHECK(n <= ((size_t)1 << 36), VIRA_GPU_ERR_INVALID, "star_light: too many stars for one launch");
    k_star_light<<<blocks_for(n * NB, kBlock), kBlock, 0, c.stream>>>(stars, n, et, dirs, irr);
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(c.stream));
}

// "QUIPU", then readVector (quipu_io.ipp:118-128): uint32 count, uint32 byte
// size, the raw records.
void qsc_read(const char* path, vira_gpu_star* out, size_t capacity, size_t* count) {
    static_assert(sizeof(vira_gpu_star) == 64, "a star quipu record is 64 bytes");
    VG_CHECK(path && count, VIRA_GPU_ERR_INVALID, "qsc_read: null argument");
    FILE* f = std::fopen(path, "rb");
    VG_CHECK(f, VIRA_GPU_ERR_INVALID, std::string("qsc_read: cannot open ") + path);
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    char ident[5];
    uint32_t hdr[2];
    VG_CHECK(std::fread(ident, 1, 5, f) == 5 && std::memcmp(ident, "QUIPU", 5) == 0, VIRA_GPU_ERR_INVALID,
             std::string(path) + " is not a valid Quipu");
    VG_CHECK(std::fread(hdr, 4, 2, f) == 2, VIRA_GPU_ERR_INVALID, "qsc_read: truncated header");
    VG_CHECK((uint64_t)hdr[0] * sizeof(vira_gpu_star) == hdr[1], VIRA_GPU_ERR_INVALID,
             "qsc_read: byte size " + std::to_string(hdr[1]) + " does not hold " + std::to_string(hdr[0]) +
                 " 64-byte star records");
    *count = hdr[0];
    if (!out) return;
    const size_t take = std::min<size_t>(capacity, hdr[0]);
    VG_CHECK(std::fread(out, sizeof(vira_gpu_star), take, f) == take, VIRA_GPU_ERR_INVALID,
             "qsc_read: truncated star records");
}

}  // namespace vg

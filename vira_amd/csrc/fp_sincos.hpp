// fp_sincos.hpp -- FP32 sine and cosine written out in +, -, * only, so the
// GPU kernels and the CPU restatement that checks them round identically.
//
// The path tracer draws directions with sin/cos (the sphere light's cone,
// sphere_light.ipp:67-125, and the concentric-disk map, sampling.ipp:53-84).
// libm's sinf/cosf and the device's differ in the last ulp, and one ulp in a
// bounce direction sends a path elsewhere; with this one routine on both sides
// bounce renders agree bit for bit.
//
// Method (Cody & Waite, "Software Manual for the Elementary Functions", 1980):
//   j = rint(x * 2/pi);  r = ((x - j*P1) - j*P2) - j*P3,  P1 + P2 + P3 = pi/2,
//   P1 and P2 carry 12 significant bits, so j*P1 and j*P2 are exact for
//   |j| < 2^12 (|x| < ~6400) and r is the reduced argument in [-pi/4, pi/4];
//   sin r = r + r^3 S(r^2),  cos r = (1 - r^2/2) + r^4 C(r^2),
//   S and C of degree 2, fitted by the Remez exchange on z = r^2 in [0, (pi/4)^2]
//   (relative error of the fits 7.6e-9 and 2.6e-10, below FP32's 6e-8), with
//   the coefficients rounded to FP32;
//   the quadrant j mod 4 selects +-sin r / +-cos r.
// Measured against FP64 sin/cos over [-50, 50]: at most 1.6 ulp; where the
// result is below 1e-6 (next to a zero) the bound is absolute, 2e-14, from the
// FP32 tail P3 of pi/2 (the tests pin both).
// Compiled with FP contraction off on both sides (no FMA), in the order written.
#pragma once

#if defined(__HIPCC__)
#define VG_FP_INLINE __host__ __device__ __forceinline__
#else
#define VG_FP_INLINE inline
#endif

namespace vgm {

constexpr float kTwoOverPi = 0.6366197466850281f;
constexpr float kPio2P1 = 1.5703125f;                 // 0x3fc90000
constexpr float kPio2P2 = 4.837512969970703125e-4f;   // 0x39fda000
constexpr float kPio2P3 = 7.549790126404332e-08f;     // 0x33a22169
constexpr float kSinS1 = -0.1666666567325592f;        // 0xbe2aaaaa
constexpr float kSinS2 = 0.008332689292728901f;       // 0x3c0885d5
constexpr float kSinS3 = -0.0001957274944288656f;     // 0xb94d3c33
constexpr float kCosC1 = 0.041666653007268906f;       // 0x3d2aaaa7
constexpr float kCosC2 = -0.0013887654058635235f;     // 0xbab6073c
constexpr float kCosC3 = 2.446383768983651e-05f;      // 0x37cd37b1

// s = sin(x), c = cos(x) for finite x.
VG_FP_INLINE void sincos_f32(float x, float& s, float& c) {
    const float j = __builtin_rintf(x * kTwoOverPi);
    const float r = ((x - j * kPio2P1) - j * kPio2P2) - j * kPio2P3;
    const float z = r * r;
    const float sr = r + (r * z) * (kSinS1 + z * (kSinS2 + z * kSinS3));
    const float cr = (1.0f - 0.5f * z) + (z * z) * (kCosC1 + z * (kCosC2 + z * kCosC3));
    const int q = ((int)j) & 3;
    const float ss = (q & 1) ? cr : sr;
    const float cc = (q & 1) ? sr : cr;
    s = (q & 2) ? -ss : ss;
    c = ((q + 1) & 2) ? -cc : cc;
}

}  // namespace vgm

// fp_trig.hpp -- FP32 sine, cosine and atan2 written out in +, -, *, / only,
// so the GPU kernels and the CPU restatement that checks them round
// identically.
//
// The path tracer draws directions with sin/cos (the sphere light's cone,
// sphere_light.ipp:67-125, and the concentric-disk map, sampling.ipp:53-84)
// and looks up the background by the direction's atan2 angles
// (cpu_path_tracer.ipp:257-270).
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

// atan2(y, x) in (-pi, pi] for finite y, x (atan2(+-0, +0) = +-0 and
// atan2(+-0, -0 or x < 0) = +-pi, as C's atan2).
// Method: t = min(|x|,|y|) / max(|x|,|y|) in [0, 1]; above tan(pi/8) the
// identity atan t = pi/4 + atan((t - 1)/(t + 1)) brings the argument to
// [-tan(pi/8), tan(pi/8)], where atan u = u + u^3 A(u^2) with A of degree 4
// fitted by the Remez exchange (relative error 1.1e-9); then the octant
// reflections pi/2 - a, pi - a and the sign of y, each constant split in two
// FP32 parts.  Measured against FP64 atan2: at most 2.2 ulp (tests pin 2.5).
constexpr float kAtanA1 = -0.3333333432674408f;   // 0xbeaaaaab
constexpr float kAtanA2 = 0.19999337196350098f;   // 0x3e4ccb10
constexpr float kAtanA3 = -0.1425725370645523f;   // 0xbe11fe89
constexpr float kAtanA4 = 0.10682182759046555f;   // 0x3ddac567
constexpr float kAtanA5 = -0.06278196722269058f;  // 0xbd8093d5
constexpr float kTanPi8 = 0.4142135679721832f;
constexpr float kPio4Hi = 0.7853981852531433f, kPio4Lo = -2.1855695e-08f;
constexpr float kPio2Hi = 1.5707963705062866f, kPio2Lo = -4.371139e-08f;
constexpr float kPiHi = 3.1415927410125732f, kPiLo = -8.742278e-08f;

VG_FP_INLINE float atan2_f32(float y, float x) {
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    const bool xneg = __builtin_signbit(x) != 0, yneg = __builtin_signbit(y) != 0;
    if (ax == 0.0f && ay == 0.0f) {
        const float a = xneg ? kPiHi : 0.0f;
        return yneg ? -a : a;
    }
    const bool swap = ay > ax;
    const float t = swap ? ax / ay : ay / ax;  // in [0, 1]
    const bool hi = t > kTanPi8;
    const float u = hi ? (t - 1.0f) / (t + 1.0f) : t;
    const float z = u * u;
    const float p = u + (u * z) * (kAtanA1 + z * (kAtanA2 + z * (kAtanA3 + z * (kAtanA4 + z * kAtanA5))));
    float a = hi ? (kPio4Hi + p) + kPio4Lo : p;
    if (swap) a = (kPio2Hi - a) + kPio2Lo;
    if (xneg) a = (kPiHi - a) + kPiLo;
    return yneg ? -a : a;
}

}  // namespace vgm

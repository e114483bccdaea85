// fftconv.hip -- zero-padded correlation of an 8-band image with a K x K
// kernel through 2-D FFTs: the branch Image::convolve takes for kernels of
// 256 pixels or more (image.ipp:1006-1065, convolveChannelFFT :1160-1264),
//   out(x, y) = sum_{kx, ky} in(x + kx - K/2, y + ky - K/2) * kernel(kx, ky),
// the same result as the spatial branch up to the transforms' rounding.
//
// Sizes: Lx = 2^a >= max(W + K/2, K), Ly = 2^b >= max(H + K/2, K).  A tap
// that reaches past the image then lands in the zero padding and never wraps
// onto a pixel, and no two taps share a residue, so the circular correlation
// of size Lx x Ly restricted to the W x H pixels is the zero-padded one.
//
// Three launches per image, each a batch of in-LDS Stockham FFTs (radix 16,
// then one radix 2/4/8 stage; twiddles from an FP64-computed table):
//   rows:    per image row, the two bands 2p, 2p+1 packed as one complex row
//            (re, im), forward FFT of length Lx, split into the two bands'
//            half spectra (fx = 0 .. Lx/2; a real row's spectrum is
//            Hermitian), written to the spectrum buffer in strips of CW
//            columns: spec[band][strip][row][c];
//   columns: per (band, strip), the CW columns of length Ly (rows past the
//            image are zero), forward FFT, times the kernel's spectrum,
//            inverse FFT, the H image rows written back in place;
//   inverse: per image row, the band pairs' half spectra extended by
//            Hermitian symmetry, packed (a + i b), inverse FFT, real and
//            imaginary parts written to the two bands.
// The kernel's spectrum comes from the first two steps run on the kernel
// placed at offsets (kx - K/2, ky - K/2) mod (Lx, Ly), conjugated (a
// correlation) and scaled by 1 / (4 Lx Ly) (the inverse transforms are
// unnormalised and both operands come out of the row split doubled).  It is
// kept on the context and recomputed only when the kernel's values or the
// padded sizes change.
#include "context.hpp"

namespace vg {
namespace {

constexpr int kFcLogMax = 13;
constexpr int kFcMax = 1 << kFcLogMax;  // longest transform (8192)

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 mul_negi(float2 a) { return make_float2(a.y, -a.x); }  // a * (-i)

// cos and sin of 2 pi m / 16
constexpr double kC1 = 0.92387953251128674, kS1 = 0.38268343236508977, kR2 = 0.70710678118654752;
constexpr double kCos16[16] = {1, kC1, kR2, kS1, 0, -kS1, -kR2, -kC1, -1, -kC1, -kR2, -kS1, 0, kS1, kR2, kC1};
constexpr double kSin16[16] = {0, kS1, kR2, kC1, 1, kC1, kR2, kS1, 0, -kS1, -kR2, -kC1, -1, -kC1, -kR2, -kS1};

// v * exp(-2 pi i m / 16), m known at compile time
template <int M>
__device__ __forceinline__ float2 w16(float2 v) {
    constexpr int m = M & 15;
    if constexpr (m == 0) return v;
    else if constexpr (m == 4) return mul_negi(v);
    else if constexpr (m == 8) return make_float2(-v.x, -v.y);
    else if constexpr (m == 12) return make_float2(-v.y, v.x);
    else return cmul(v, make_float2((float)kCos16[m], (float)-kSin16[m]));
}

__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
    const float2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = mul_negi(csub(v1, v3));
    v0 = cadd(a0, a2);
    v1 = cadd(a1, a3);
    v2 = csub(a0, a2);
    v3 = csub(a1, a3);
}

// In-register DFT of R points (R = 2, 4, 8, 16): R = 4 * N2 splits into N2
// 4-point DFTs over stride-N2 inputs, twiddles w_R^(n2 k1), then 4 N2-point
// DFTs; outputs in natural order.
template <int R>
__device__ __forceinline__ void dft(float2* v) {
    if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3]);
    } else {
        constexpr int N2 = R / 4;
        float2 y[N2][4];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) {
            y[n2][0] = v[n2];
            y[n2][1] = v[N2 + n2];
            y[n2][2] = v[2 * N2 + n2];
            y[n2][3] = v[3 * N2 + n2];
            dft4(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
        }
        // w_R^(n2 k1) = w_16^(n2 k1 16 / R)
        if constexpr (N2 >= 2) {
            y[1][1] = w16<1 * 16 / R>(y[1][1]);
            y[1][2] = w16<2 * 16 / R>(y[1][2]);
            y[1][3] = w16<3 * 16 / R>(y[1][3]);
        }
        if constexpr (N2 == 4) {
            y[2][1] = w16<2>(y[2][1]);
            y[2][2] = w16<4>(y[2][2]);
            y[2][3] = w16<6>(y[2][3]);
            y[3][1] = w16<3>(y[3][1]);
            y[3][2] = w16<6>(y[3][2]);
            y[3][3] = w16<9>(y[3][3]);
        }
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            float2 t[N2];
#pragma unroll
            for (int n2 = 0; n2 < N2; ++n2) t[n2] = y[n2][k1];
            dft<N2>(t);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) v[k1 + 4 * k2] = t[k2];
        }
    }
}

// LDS slot of logical element i: for transforms of 256 points or more, one
// unused slot after every 16, so that the first stage's writes (16
// consecutive values per thread) and the other stages' strided accesses fall
// in distinct banks.  The strides the stages step by are then multiples of
// 16, so slot(i + m) = slot(i) + m + m/16 for them.
__device__ __forceinline__ int lp(int i, bool pad) { return pad ? i + (i >> 4) : i; }

// Twiddles exp(-2 pi i m / 8192) as the product of two correctly rounded
// table entries, exp(-2 pi i 64 (m >> 6) / 8192) (128 coarse) and
// exp(-2 pi i (m & 63) / 8192) (64 fine): 192 entries, copied to LDS by every
// block, instead of one 8192-entry table read through the caches.
constexpr int kFcTwCoarse = 128, kFcTwFine = 64, kFcTw = kFcTwCoarse + kFcTwFine;
__device__ __forceinline__ float2 twiddle(const float2* twl, unsigned m) {
    return cmul(twl[m >> 6], twl[kFcTwCoarse + (m & 63)]);
}
// LDS of a block: the twiddle tables, then the transforms' slots
__device__ __forceinline__ float2* load_twiddles(float2* lds, const float2* __restrict__ tw) {
    for (int i = threadIdx.x; i < kFcTw; i += blockDim.x) lds[i] = tw[i];
    return lds + kFcTw;
}
__host__ inline size_t lds_bytes(size_t n) { return (kFcTw + n + n / 16 + 1) * sizeof(float2); }

// One radix-R Stockham stage over nb sequences of length L = 2^logL in LDS
// (logical element i of sequence q at slot lp(q L + i, L >= 256)), after stages whose
// radices multiply to Ns:
// butterfly j of a sequence reads z[j + r L/R], twiddles by
// exp(-2 pi i r (j mod Ns) / (Ns R)), and writes its DFT to
// z[(j - j mod Ns) R + j mod Ns + r Ns].  A thread holds at most 16 values
// (nb L / blockDim <= 16), read before the barrier and written after it.
template <int R>
__device__ __forceinline__ void fft_stage(float2* z, int logL, int nb, int logNs, const float2* twl) {
    constexpr int kQ = 16 / R;
    constexpr int logR = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
    const int logB = logL - logR, B = 1 << logB;  // butterflies per sequence
    const int Ns = 1 << logNs, total = nb * B;
    const int twshift = kFcLogMax - (logNs + logR);
    const bool pad = logL >= 8;
    // slot steps between a butterfly's reads (B apart) and writes (Ns apart;
    // the first stage's 16 writes are consecutive and start on a multiple of 16)
    const int rstep = pad ? B + (B >> 4) : B;
    const int wstep = pad ? (Ns >= 16 ? Ns + (Ns >> 4) : 1) : Ns;
    float2 v[kQ][R];
    int dst[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int jj = (int)threadIdx.x + q * (int)blockDim.x;
        dst[q] = -1;
        if (jj < total) {
            const int s = jj >> logB, j = jj & (B - 1), k = j & (Ns - 1);
            const int base = lp((s << logL) + j, pad);
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = z[base + r * rstep];
            if (logNs > 0) {
                // w^r for w = exp(-2 pi i k / (Ns R)): w, w^2, w^4, w^8 from the tables, the other
                // powers as one product of two of those (at most three roundings deep)
                float2 w[R];
                const unsigned t = (unsigned)k << twshift;
#pragma unroll
                for (int r = 1; r < R; r <<= 1) w[r] = twiddle(twl, t * r);
#pragma unroll
                for (int r = 3; r < R; ++r) {
                    const int hb = r >= 8 ? 8 : r >= 4 ? 4 : 2;
                    if (r != hb) w[r] = cmul(w[hb], w[r - hb]);
                }
#pragma unroll
                for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], w[r]);
            }
            dft<R>(v[q]);
            dst[q] = lp((s << logL) + ((j - k) << logR) + k, pad);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        if (dst[q] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) z[dst[q] + r * wstep] = v[q][r];
        }
    }
    __syncthreads();
}

// Forward FFT (unnormalised, natural order in and out) of nb sequences of
// length 2^logL in LDS (slots lp(q L + i, L >= 256)); the block must be synchronised on
// entry and is on exit.
__device__ void fft_lds(float2* z, int logL, int nb, const float2* twl) {
    int logNs = 0;
    for (; logL - logNs >= 4; logNs += 4) fft_stage<16>(z, logL, nb, logNs, twl);
    const int rem = logL - logNs;
    if (rem == 3) fft_stage<8>(z, logL, nb, logNs, twl);
    else if (rem == 2) fft_stage<4>(z, logL, nb, logNs, twl);
    else if (rem == 1) fft_stage<2>(z, logL, nb, logNs, twl);
}

struct FcGeom {
    int logLx, logLy;
    int cw, ns;  // strip width (columns) and strips per band
    int nc;      // columns per column-transform block (1 or 2)
    int logcw;   // log2(cw)
};

// Workgroups reach the 8 XCDs round robin (block b on XCD b mod 8).  The
// logical index of block b numbers each XCD's blocks consecutively, so
// neighbouring logical blocks (rows or half-strips sharing cache lines) run
// on one XCD, through one L2, at about the same time.
__device__ __forceinline__ int xcd_block(int b, int n) {
    const int q = n >> 3, rem = n & 7, x = b & 7, i = b >> 3;
    return x < rem ? x * (q + 1) + i : rem * (q + 1) + (x - rem) * q + i;
}

__host__ __device__ inline size_t spec_at(const FcGeom& g, int nrows, int band, int fx, int row) {
    return ((((size_t)band * g.ns + (fx >> g.logcw)) * nrows + row) << g.logcw) + (fx & (g.cw - 1));
}

// Row transform: row `row` of an AoS [rows][width][8] image, each band pair
// packed as one complex row placed at x - shift (mod Lx), forward FFT, the
// two half spectra (doubled) into spec (nrows rows per strip).
__global__ void __launch_bounds__(1024) k_fc_rows(const float* __restrict__ src, int width, int shift, FcGeom g,
                                                  int nrows, float2* __restrict__ spec, const float2* __restrict__ tw,
                                                  const uint32_t* __restrict__ only_if) {
    if (only_if && *only_if == 0u) return;
    extern __shared__ float2 fc_lds[];
    float2* z = load_twiddles(fc_lds, tw);
    const float2* twl = fc_lds;
    const int row = xcd_block(blockIdx.x, gridDim.x), Lx = 1 << g.logLx, T = blockDim.x;
    const bool pad = g.logLx >= 8;
    const float* in = src + (size_t)row * width * 8;
    for (int p = 0; p < 4; ++p) {
        for (int i = threadIdx.x; i < Lx; i += T) {
            const int x = (i + shift) & (Lx - 1);
            float2 v = make_float2(0.f, 0.f);
            if (x < width) v = *reinterpret_cast<const float2*>(in + (size_t)x * 8 + 2 * p);
            z[lp(i, pad)] = v;
        }
        __syncthreads();
        fft_lds(z, g.logLx, 1, twl);
        // Z = A + iB with A, B the spectra of the real rows: 2A = Z(f) + conj Z(-f), 2B = -i (Z(f) - conj Z(-f))
        const int ncols = g.ns * g.cw;
        for (int f = threadIdx.x; f < ncols; f += T) {
            float2 a = make_float2(0.f, 0.f), b = a;
            if (f <= Lx / 2) {
                const float2 zk = z[lp(f, pad)], zm = cconj(z[lp((Lx - f) & (Lx - 1), pad)]);
                a = cadd(zk, zm);
                b = mul_negi(csub(zk, zm));
            }
            spec[spec_at(g, nrows, 2 * p, f, row)] = a;
            spec[spec_at(g, nrows, 2 * p + 1, f, row)] = b;
        }
        __syncthreads();
    }
}

// Column transform of nc columns of one (band, strip): SPECTRUM = the kernel's
// rows (nrows = K, row ky at (ky - shift) mod Ly), forward FFT, stored
// conjugated and scaled as all Ly rows of the kernel spectrum; otherwise the
// image rows (nrows = H from row 0), forward FFT, times the kernel spectrum,
// inverse FFT, the H rows written back.  A strip's column blocks are
// consecutive after xcd_block, so they read the strip's lines through one L2.
template <bool SPECTRUM>
__global__ void __launch_bounds__(1024) k_fc_cols(float2* __restrict__ spec, int nrows, int shift, FcGeom g,
                                                  float2* __restrict__ kspec, float scale,
                                                  const float2* __restrict__ tw, const uint32_t* __restrict__ only_if) {
    if (only_if && *only_if == 0u) return;
    extern __shared__ float2 fc_lds[];
    const int parts = g.cw / g.nc, lognc = g.nc == 2 ? 1 : 0;
    const int L = xcd_block(blockIdx.x, gridDim.x);
    const int band = L / (g.ns * parts), rest = L - band * g.ns * parts;
    const int strip = rest / parts, c0 = (rest - strip * parts) * g.nc;
    const int Ly = 1 << g.logLy, cw = g.cw, nc = g.nc, T = blockDim.x;
    const bool pad = g.logLy >= 8;
    float2* z = load_twiddles(fc_lds, tw);
    const float2* twl = fc_lds;
    const size_t sbase = ((size_t)band * g.ns + strip) * (size_t)nrows * cw + c0;
    const size_t kbase = ((size_t)band * g.ns + strip) * (size_t)Ly * cw + c0;
    for (int e = threadIdx.x; e < Ly * nc; e += T) {
        const int r = e >> lognc, c = e & (nc - 1);
        const int src = SPECTRUM ? ((r + shift) & (Ly - 1)) : r;
        z[lp(c * Ly + r, pad)] = src < nrows ? spec[sbase + (size_t)src * cw + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lds(z, g.logLy, nc, twl);
    if (SPECTRUM) {
        for (int e = threadIdx.x; e < Ly * nc; e += T) {
            const int r = e >> lognc, c = e & (nc - 1);
            const float2 v = z[lp(c * Ly + r, pad)];
            kspec[kbase + (size_t)r * cw + c] = make_float2(v.x * scale, -v.y * scale);
        }
        return;
    }
    // inverse through the forward transform: ifft(X) = conj(fft(conj X)) (unnormalised)
    for (int e = threadIdx.x; e < Ly * nc; e += T) {
        const int r = e >> lognc, c = e & (nc - 1);
        const int i = lp(c * Ly + r, pad);
        z[i] = cconj(cmul(z[i], kspec[kbase + (size_t)r * cw + c]));
    }
    __syncthreads();
    fft_lds(z, g.logLy, nc, twl);
    for (int e = threadIdx.x; e < nrows * nc; e += T) {
        const int r = e >> lognc, c = e & (nc - 1);
        spec[sbase + (size_t)r * cw + c] = cconj(z[lp(c * Ly + r, pad)]);
    }
}

// Inverse row transform: row `row`, per band pair the half spectra extended by
// Hermitian symmetry and packed a + i b, inverse FFT, written to the bands.
__global__ void __launch_bounds__(1024) k_fc_rows_inv(const float2* __restrict__ spec, int width, FcGeom g,
                                                      int nrows, float* __restrict__ out,
                                                      const float2* __restrict__ tw) {
    extern __shared__ float2 fc_lds[];
    float2* z = load_twiddles(fc_lds, tw);
    const float2* twl = fc_lds;
    const int row = xcd_block(blockIdx.x, gridDim.x), Lx = 1 << g.logLx, T = blockDim.x;
    const bool pad = g.logLx >= 8;
    float* o = out + (size_t)row * width * 8;
    for (int p = 0; p < 4; ++p) {
        // each stored frequency m fills slots m and Lx - m (A(-m) = conj A(m) for both bands); the
        // slot holds conj(a + i b), so the forward FFT computes the inverse
        for (int m = threadIdx.x; m <= Lx / 2; m += T) {
            const float2 a = spec[spec_at(g, nrows, 2 * p, m, row)], b = spec[spec_at(g, nrows, 2 * p + 1, m, row)];
            z[lp(m, pad)] = make_float2(a.x - b.y, -(a.y + b.x));
            if (m > 0 && m < Lx / 2) z[lp(Lx - m, pad)] = make_float2(a.x + b.y, a.y - b.x);
        }
        __syncthreads();
        fft_lds(z, g.logLx, 1, twl);
        for (int x = threadIdx.x; x < width; x += T) {
            const float2 v = z[lp(x, pad)];  // conj of (a + i b)(x)
            *reinterpret_cast<float2*>(o + (size_t)x * 8 + 2 * p) = make_float2(v.x, -v.y);
        }
        __syncthreads();
    }
}

// One block: *changed = whether the kernel's bits differ from the copy kept
// with its spectrum (or force); if so the copy is replaced.
__global__ void __launch_bounds__(1024) k_fc_kernel_changed(const float* __restrict__ kern, float* __restrict__ kept,
                                                            int n, int force, uint32_t* __restrict__ changed) {
    __shared__ uint32_t diff;
    if (threadIdx.x == 0) diff = force ? 1u : 0u;
    __syncthreads();
    // n is a multiple of 8 and both buffers are 8-byte aligned: compare and copy in pairs
    const uint2* a = reinterpret_cast<const uint2*>(kern);
    const uint2* b = reinterpret_cast<const uint2*>(kept);
    const int n2 = n / 2;
    uint32_t d = 0u;
    if (!force) {
#pragma unroll 4
        for (int i = threadIdx.x; i < n2; i += blockDim.x) {
            const uint2 x = a[i], y = b[i];
            d |= (x.x ^ y.x) | (x.y ^ y.y);
        }
    }
    if (d) atomicOr(&diff, 1u);
    __syncthreads();
    const bool replace = diff != 0u;
    if (replace)
        for (int i = threadIdx.x; i < n2; i += blockDim.x) reinterpret_cast<uint2*>(kept)[i] = a[i];
    if (threadIdx.x == 0) *changed = replace ? 1u : 0u;
}

int log2_ceil(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

// threads for nb transforms of length 2^logL: at most 16 values per thread
unsigned fc_threads(int logL, int nb) {
    const long n = (long)nb << logL;
    return (unsigned)std::min<long>(1024, std::max<long>(64, n / 16));
}

}  // namespace

bool fft_convolve_fits(int W, int H, int K) {
    const int h = K / 2;
    return std::max(W + h, K) <= kFcMax && std::max(H + h, K) <= kFcMax;
}

void fft_convolve(Ctx& c, const float* in, int W, int H, const float* kern, int K, float* out) {
    VG_CHECK(fft_convolve_fits(W, H, K), VIRA_GPU_ERR_INVALID, "fft_convolve: image too large for one transform");
    VG_CHECK((((uintptr_t)in | (uintptr_t)out | (uintptr_t)kern) & 7) == 0, VIRA_GPU_ERR_INVALID,
             "convolve: image, kernel and output must be 8-byte aligned for kernels of 256 taps or more");
    const int h = K / 2;
    FcGeom g;
    g.logLx = std::max(1, log2_ceil(std::max(W + h, K)));
    g.logLy = std::max(1, log2_ceil(std::max(H + h, K)));
    g.cw = g.logLy <= 12 ? 4 : 2;  // strip width: a half strip of Ly rows fills at most 70 KiB of LDS
    g.logcw = g.cw == 4 ? 2 : 1;
    const int Lx = 1 << g.logLx, Ly = 1 << g.logLy;
    g.ns = (Lx / 2 + 1 + g.cw - 1) / g.cw;
    g.nc = g.cw / 2;  // two column blocks per strip: 70 KiB of LDS each at most, two blocks per CU
    const size_t row_lds = lds_bytes(Lx), col_lds = lds_bytes((size_t)g.nc * Ly);
    const unsigned col_blocks = 8u * g.ns * (g.cw / g.nc);
    if (!c.fc_attrs) {  // the blocks take up to 70 KiB of LDS
        for (const void* f : {(const void*)k_fc_rows, (const void*)k_fc_rows_inv, (const void*)k_fc_cols<true>,
                              (const void*)k_fc_cols<false>})
            VG_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        c.fc_attrs = true;
    }
    if (!c.fc_tw.p) {
        std::vector<float2> t(kFcTw);
        for (int i = 0; i < kFcTw; ++i) {
            const int m = i < kFcTwCoarse ? 64 * i : i - kFcTwCoarse;
            const double a = -2.0 * 3.14159265358979323846 * (double)m / (double)kFcMax;
            t[i] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        c.fc_tw.alloc(t.size() * sizeof(float2));
        VG_HIP(hipMemcpy(c.fc_tw.p, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
    const float2* tw = c.fc_tw.as<float2>();
    // The kernel's spectrum: recomputed when the sizes changed or, decided on the
    // device without a round trip, when the kernel's values differ from the
    // copy kept with it (the spectrum launches then return at once).
    const bool same_sizes = c.fc_kvalid && c.fc_K == K && c.fc_logLx == g.logLx && c.fc_logLy == g.logLy;
    const size_t nk = (size_t)K * K * 8;
    c.fc_kvalid = false;
    c.fc_kept.alloc(nk * sizeof(float));
    c.fc_flag.alloc(sizeof(uint32_t));
    c.fc_kspec.alloc((size_t)8 * g.ns * Ly * g.cw * sizeof(float2));
    c.fc_spec.alloc((size_t)8 * g.ns * std::max(H, K) * g.cw * sizeof(float2));
    float2* spec = c.fc_spec.as<float2>();
    const uint32_t* changed = c.fc_flag.as<uint32_t>();
    k_fc_kernel_changed<<<1, 1024, 0, c.stream>>>(kern, c.fc_kept.as<float>(), (int)nk, same_sizes ? 0 : 1,
                                                  c.fc_flag.as<uint32_t>());
    VG_HIP(hipGetLastError());
    k_fc_rows<<<K, fc_threads(g.logLx, 1), row_lds, c.stream>>>(kern, K, h, g, K, spec, tw, changed);
    VG_HIP(hipGetLastError());
    k_fc_cols<true><<<col_blocks, fc_threads(g.logLy, g.nc), col_lds, c.stream>>>(
        spec, K, h, g, c.fc_kspec.as<float2>(), 1.0f / (4.0f * (float)Lx * (float)Ly), tw, changed);
    VG_HIP(hipGetLastError());
    c.fc_K = K;
    c.fc_logLx = g.logLx;
    c.fc_logLy = g.logLy;
    c.fc_kvalid = true;
    k_fc_rows<<<H, fc_threads(g.logLx, 1), row_lds, c.stream>>>(in, W, 0, g, H, spec, tw, nullptr);
    VG_HIP(hipGetLastError());
    k_fc_cols<false><<<col_blocks, fc_threads(g.logLy, g.nc), col_lds, c.stream>>>(
        spec, H, 0, g, c.fc_kspec.as<float2>(), 0.0f, tw, nullptr);
    VG_HIP(hipGetLastError());
    k_fc_rows_inv<<<H, fc_threads(g.logLx, 1), row_lds, c.stream>>>(spec, W, g, H, out, tw);
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace vg

// denoise.hip -- edge-avoiding a-trous wavelet denoiser of the direct and
// indirect radiance passes (denoiseSpectralRadianceEATWT, cpu_denoise.ipp:10-187).
//
// The reference demodulates both passes by max(albedo, eps), filters each
// with levels of a separable 5-tap B3 kernel (1 4 6 4 1)/16 at step 2^level,
// each tap weighted by a depth-ratio and a normal-similarity term, then
// remodulates.  Its images are indexed image(y, x) = data[y + x*W]: the first
// pass of a level moves along x in that index, i.e. along the image's rows
// index j for a square image, and the second along i.  This file keeps that
// order for square images and, for any W x H, filters first along j (taps
// clamped to [0, H-1]) and then along i (clamped to [0, W-1]).
//
// Both passes share every weight (the weights depend only on depth and
// normal), so one thread filters a pixel of both images: 16 channels per
// tap.  The first pass divides by the albedo as it reads, the last one
// multiplies it back and writes the caller's buffers (and total = direct +
// indirect), so the demodulated images never go to HBM on their own.
#include "context.hpp"

namespace vg {
namespace {

constexpr int kDnBX = 64, kDnBY = 4;

struct DnImage {
    const float* src;   // W*H*8: the caller's pass (first pass) or a work image
    float* dst;         // W*H*8: a work image, or the caller's pass on the last pass
    int first, last, active;
};

struct DnParams {
    int W, H, step, vertical;
    float depth_threshold, normal_threshold, min_normal_weight, eps;
    const float* albedo;  // W*H*8
    const float* depth;   // W*H
    const float* normal;  // W*H*3
    float* total;         // W*H*8 or null: written by the pass that finishes the later image
    const float* total_other;  // the already finished image's final pass when only one is active
};

// std::max / std::min as the reference evaluates them (NaN operands included).
__device__ __forceinline__ float ref_max(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float ref_min(float a, float b) { return (b < a) ? b : a; }

__device__ __forceinline__ void load8(const float* p, float v[NB]) {
    const float4 a = *reinterpret_cast<const float4*>(p), b = *reinterpret_cast<const float4*>(p + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void store8(float* p, const float v[NB]) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<float4*>(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
}

// One separable pass of one level over both images.
__global__ void __launch_bounds__(kDnBX* kDnBY) k_eatwt_pass(DnParams P, DnImage A, DnImage B) {
    const int i = blockIdx.x * kDnBX + threadIdx.x, j = blockIdx.y * kDnBY + threadIdx.y;
    if (i >= P.W || j >= P.H) return;
    const size_t c = (size_t)j * P.W + i;
    const float kern[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};
    const float cd = P.depth[c];
    const float cnx = P.normal[3 * c], cny = P.normal[3 * c + 1], cnz = P.normal[3 * c + 2];
    float alb[NB];
    const bool need_alb = (A.active && (A.first || A.last)) || (B.active && (B.first || B.last));
    if (need_alb) load8(P.albedo + NB * c, alb);
    float sa[NB], sb[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) sa[k] = sb[k] = 0.0f;
    float ws = 0.0f;
    for (int t = -2; t <= 2; ++t) {
        int x = i, y = j;
        if (P.vertical) y = min(max(j + t * P.step, 0), P.H - 1);
        else x = min(max(i + t * P.step, 0), P.W - 1);
        const size_t s = (size_t)y * P.W + x;
        const float sd = P.depth[s];
        const float ratio = ref_max(cd, sd) / (ref_min(cd, sd) + P.eps);
        const float dw = ratio < (1.0f + P.depth_threshold) ? 1.0f : expf(-(ratio - 1.0f) / P.depth_threshold);
        const float dot = (cnx * P.normal[3 * s] + cny * P.normal[3 * s + 1]) + cnz * P.normal[3 * s + 2];
        const float sim = ref_max(0.0f, dot);
        const float u = 1.0f - sim;
        const float nw = ref_max(P.min_normal_weight, expf(-(u * u) / P.normal_threshold));
        const float w = (kern[t + 2] * dw) * nw;
        float sal[NB];
        if (A.first || B.first) load8(P.albedo + NB * s, sal);
        if (A.active) {
            float v[NB];
            load8(A.src + NB * s, v);
            if (A.first) {
#pragma unroll
                for (int k = 0; k < NB; ++k) v[k] = v[k] / ref_max(sal[k], P.eps);
            }
#pragma unroll
            for (int k = 0; k < NB; ++k) sa[k] = sa[k] + v[k] * w;
        }
        if (B.active) {
            float v[NB];
            load8(B.src + NB * s, v);
            if (B.first) {
#pragma unroll
                for (int k = 0; k < NB; ++k) v[k] = v[k] / ref_max(sal[k], P.eps);
            }
#pragma unroll
            for (int k = 0; k < NB; ++k) sb[k] = sb[k] + v[k] * w;
        }
        ws = ws + w;
    }
    // normalise, or keep the centre value when the weights vanish (or are NaN)
    const bool norm = ws > P.eps;
    float fa[NB], fb[NB];
    auto finish = [&](const DnImage& I, float* sum, float* out) {
        if (norm) {
#pragma unroll
            for (int k = 0; k < NB; ++k) out[k] = sum[k] / ws;
        } else {
            load8(I.src + NB * c, out);
            if (I.first) {
#pragma unroll
                for (int k = 0; k < NB; ++k) out[k] = out[k] / ref_max(alb[k], P.eps);
            }
        }
        if (I.last) {
#pragma unroll
            for (int k = 0; k < NB; ++k) out[k] = out[k] * alb[k];
        }
        store8(I.dst + NB * c, out);
    };
    if (A.active) finish(A, sa, fa);
    if (B.active) finish(B, sb, fb);
    if (P.total) {
        // total = direct + indirect once both are final (cpu_path_tracer.ipp:133)
        float da[NB], db[NB];
        if (A.active) {
#pragma unroll
            for (int k = 0; k < NB; ++k) da[k] = fa[k];
        } else {
            load8(P.total_other + NB * c, da);
        }
        if (B.active) {
#pragma unroll
            for (int k = 0; k < NB; ++k) db[k] = fb[k];
        } else {
            load8(P.total_other + NB * c, db);
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) da[k] = da[k] + db[k];
        store8(P.total + NB * c, da);
    }
}

// An image filtered with zero levels still goes through demodulation and
// remodulation: v -> (v / max(albedo, eps)) * albedo.
__global__ void k_eatwt_identity(size_t n, const float* __restrict__ albedo, float eps, float* __restrict__ img) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float a = albedo[q];
    img[q] = (img[q] / ref_max(a, eps)) * a;
}

__global__ void k_add(size_t n, const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ out) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    out[q] = a[q] + b[q];
}

}  // namespace

void denoise(Ctx& c, const vira_gpu_denoise_opts& o, int W, int H, float* direct, float* indirect,
             const float* albedo, const float* depth, const float* normal, float* total) {
    VG_CHECK(W > 0 && H > 0, VIRA_GPU_ERR_INVALID, "denoise: empty image");
    VG_CHECK(direct && indirect && albedo && depth && normal, VIRA_GPU_ERR_INVALID,
             "denoise: direct, indirect, albedo, depth and normal are required");
    VG_CHECK(o.max_levels_direct >= 0 && o.max_levels_direct <= 16 && o.max_levels_indirect >= 0 &&
                 o.max_levels_indirect <= 16,
             VIRA_GPU_ERR_INVALID, "denoise: levels must be 0..16");
    VG_CHECK(o.depth_threshold > 0.0f && o.normal_threshold > 0.0f, VIRA_GPU_ERR_INVALID,
             "denoise: thresholds must be positive");
    VG_CHECK(((uintptr_t)direct | (uintptr_t)indirect | (uintptr_t)albedo | (uintptr_t)total) % 16 == 0,
             VIRA_GPU_ERR_INVALID, "denoise: spectral images must be 16-byte aligned");
    const size_t npx = (size_t)W * H, bytes = npx * NB * sizeof(float);
    const int La = o.max_levels_direct, Lb = o.max_levels_indirect, L = La > Lb ? La : Lb;
    const size_t n8 = npx * NB;
    if (La == 0) k_eatwt_identity<<<blocks_for(n8, 256), 256, 0, c.stream>>>(n8, albedo, o.epsilon, direct);
    if (Lb == 0) k_eatwt_identity<<<blocks_for(n8, 256), 256, 0, c.stream>>>(n8, albedo, o.epsilon, indirect);
    VG_HIP(hipGetLastError());
    if (L > 0) {
        for (auto* b : {&c.dn_work[0], &c.dn_work[1], &c.dn_work[2], &c.dn_work[3]}) b->alloc(bytes);
        float* wa[2] = {(float*)c.dn_work[0].p, (float*)c.dn_work[1].p};
        float* wb[2] = {(float*)c.dn_work[2].p, (float*)c.dn_work[3].p};
        DnParams P;
        P.W = W;
        P.H = H;
        P.depth_threshold = o.depth_threshold;
        P.normal_threshold = o.normal_threshold;
        P.min_normal_weight = o.min_normal_weight;
        P.eps = o.epsilon;
        P.albedo = albedo;
        P.depth = depth;
        P.normal = normal;
        const dim3 grid((W + kDnBX - 1) / kDnBX, (H + kDnBY - 1) / kDnBY), block(kDnBX, kDnBY);
        const int passes = 2 * L;
        for (int p = 0; p < passes; ++p) {
            P.step = 1 << (p / 2);
            P.vertical = (p % 2) == 0;
            auto image = [&](int levels, float* user, float* const* work) {
                DnImage I{};
                I.active = p < 2 * levels;
                if (!I.active) return I;
                I.first = p == 0;
                I.last = p == 2 * levels - 1;
                I.src = I.first ? user : work[(p + 1) % 2];
                I.dst = I.last ? user : work[p % 2];
                return I;
            };
            const DnImage A = image(La, direct, wa), B = image(Lb, indirect, wb);
            // total is written by the last pass overall, reading the image finished earlier
            P.total = (total && p == passes - 1) ? total : nullptr;
            P.total_other = !A.active ? direct : (!B.active ? indirect : nullptr);
            k_eatwt_pass<<<grid, block, 0, c.stream>>>(P, A, B);
            VG_HIP(hipGetLastError());
        }
    }
    if (total && L == 0) {
        k_add<<<blocks_for(n8, 256), 256, 0, c.stream>>>(n8, direct, indirect, total);
        VG_HIP(hipGetLastError());
    }
    VG_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace vg

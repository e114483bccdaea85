// fused.hpp -- the per-pixel renderer's tile routine, shared by the
// per-request launch (render.hip) and the batched persistent launch
// (batch.hip): one wave renders one 8x8 pixel tile, one lane per pixel.
//
// CPUPathTracer::unidirectional / simulatePath / processIntersection
// (cpu_path_tracer.ipp:169-511): each lane walks its pixel's samples and
// bounces as a small state machine; every loop iteration traces exactly one
// ray per lane (camera, shadow or bounce ray, with a per-lane any-hit flag),
// so lanes whose path ended regenerate the next sample instead of idling.
// Path throughput, path radiance and the Kahan sums' inputs stay in
// registers; the per-pixel sums live in a tile-blocked accumulator block.
#pragma once

#include "instance.hpp"
#include "shading.hpp"
#include "traverse.hpp"

namespace vg {

constexpr int kFusedBS = 64;   // fused renderer: one wave = one 8x8 tile
// Waves per SIMD the per-pixel kernels are compiled for: 4 caps them at 128
// VGPRs (a few dwords spill to scratch), which measured 15 % faster on the
// 33.5M-triangle frame than the unconstrained 143 VGPRs at 3 waves.
constexpr int kFusedWavesPerSimd = 4;

// Per-pixel accumulator fields.  Layout is tile-blocked SoA: field f of
// pixel q lives at acc[((q >> 6) * A_FIELDS + f) * 64 + (q & 63)], so a wave
// owning a tile reads or writes each field as 256 contiguous bytes from one
// uniform base address.
enum : int {
    A_RAD_S = 0,
    A_RAD_C = 8,
    A_DIR_S = 16,
    A_DIR_C = 24,
    A_ALB = 32,
    A_NRM = 40,
    A_DEPTH = 43,
    A_COUNT = 44,
    A_TRI = 45,
    A_FIELDS = 46
};
// Per-path fields of the wavefront schedule (one path per pixel per sample pass).
// Accumulator of pixel q: base of its tile block; field f is base[f * 64 + (q & 63)].
__device__ __forceinline__ float* acc_tile(float* acc, uint32_t q) { return acc + (size_t)(q >> 6) * (A_FIELDS * 64); }
__device__ __forceinline__ const float* acc_tile(const float* acc, uint32_t q) {
    return acc + (size_t)(q >> 6) * (A_FIELDS * 64);
}
// The empty asm makes the lane offset opaque at each use, so the compiler does
// not hoist 46 per-field addresses out of the render loop into live registers.
__device__ __forceinline__ float& acc_ref(float* base, int f, uint32_t lane) {
    uint32_t l = lane;
    asm volatile("" : "+v"(l));
    return base[f * 64 + l];
}
__device__ __forceinline__ const float& acc_ref(const float* base, int f, uint32_t lane) {
    uint32_t l = lane;
    asm volatile("" : "+v"(l));
    return base[f * 64 + l];
}
#define ACC(base, f, lane) acc_ref((base), (f), (lane))

// First-hit passes at bounce 0 (cpu_path_tracer.ipp:352-381).
__device__ __forceinline__ void accumulate_first_hit(float* __restrict__ a, uint32_t lane, uint32_t id, float t,
                                                     const Surface& S) {
    if (__float_as_uint(ACC(a, A_TRI, lane)) == 0xffffffffu) ACC(a, A_TRI, lane) = __uint_as_float(id);
#pragma unroll
    for (int k = 0; k < NB; ++k) ACC(a, A_ALB + k, lane) = ACC(a, A_ALB + k, lane) + S.alb[k];
    ACC(a, A_NRM + 0, lane) = ACC(a, A_NRM + 0, lane) + S.N.x;
    ACC(a, A_NRM + 1, lane) = ACC(a, A_NRM + 1, lane) + S.N.y;
    ACC(a, A_NRM + 2, lane) = ACC(a, A_NRM + 2, lane) + S.N.z;
    ACC(a, A_DEPTH, lane) = ACC(a, A_DEPTH, lane) + t;
    ACC(a, A_COUNT, lane) = ACC(a, A_COUNT, lane) + 1.0f;
}

struct OutPtrs {
    float *radiance, *direct, *indirect, *albedo, *normal, *depth, *alpha;
    uint32_t* tri;
    uint32_t* inst;
};

// The first-hit id field holds instance << 27 | triangle (triangles < 2^27).
constexpr uint32_t kTriBits = 27;

// Averages and passes (cpu_path_tracer.ipp:229-239, passes.ipp:81-127).
__device__ __forceinline__ void write_pixel(const OutPtrs& o, size_t p, int spp, const float* __restrict__ a,
                                            uint32_t lane) {
    float sf = (float)spp;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        float rad = __fdiv_rn(ACC(a, A_RAD_S + k, lane), sf);
        float dr = __fdiv_rn(ACC(a, A_DIR_S + k, lane), sf);
        if (o.radiance) o.radiance[NB * p + k] = rad;
        if (o.direct) o.direct[NB * p + k] = dr;
        if (o.indirect) o.indirect[NB * p + k] = rad - dr;
        if (o.albedo) o.albedo[NB * p + k] = __fdiv_rn(ACC(a, A_ALB + k, lane), sf);
    }
    float cnt = ACC(a, A_COUNT, lane);
    if (o.normal) {
        f3 n = mk3(ACC(a, A_NRM, lane), ACC(a, A_NRM + 1, lane), ACC(a, A_NRM + 2, lane));
        if (cnt > 0.0f) n = normalize(n);
        o.normal[3 * p] = n.x;
        o.normal[3 * p + 1] = n.y;
        o.normal[3 * p + 2] = n.z;
    }
    if (o.depth) o.depth[p] = cnt > 0.0f ? __fdiv_rn(ACC(a, A_DEPTH, lane), cnt) : kInf;
    if (o.alpha) o.alpha[p] = __fdiv_rn(cnt, sf);
    const uint32_t first = __float_as_uint(ACC(a, A_TRI, lane));
    if (o.tri) o.tri[p] = first == 0xffffffffu ? first : (first & ((1u << kTriBits) - 1u));
    if (o.inst) o.inst[p] = first == 0xffffffffu ? first : (first >> kTriBits);
}

// Kahan-compensated add of v[] (or zeros) into the (sum, comp) fields fs, fc.
__device__ __forceinline__ void kahan_acc(float* __restrict__ a, uint32_t lane, int fs, int fc, const float* v,
                                          bool zero) {
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        float sv = ACC(a, fs + k, lane), cv = ACC(a, fc + k, lane);
        kahan_add(sv, cv, zero ? 0.0f : v[k]);
        ACC(a, fs + k, lane) = sv;
        ACC(a, fc + k, lane) = cv;
    }
}

__device__ __forceinline__ void acc_clear(float* __restrict__ a, uint32_t lane) {
    for (int f = 0; f < A_FIELDS; ++f) ACC(a, f, lane) = f == A_TRI ? __uint_as_float(0xffffffffu) : 0.0f;
}

enum : int { M_NEW = 0, M_CLOSEST = 1, M_SHADOW = 2, M_DONE = 3 };

// Everything one render request needs per tile (kernel argument of the
// per-request launch; device-memory descriptor of the batched launch).
struct TileJob {
    const void* nodes;
    const float4* tris;
    const GVertex* verts;
    const uint32_t* idx;
    MatF mat;
    Lights lights;
    CamF cam;
    uint32_t seed;
    int spp, max_bounces, rr, rr_start;
    uint32_t tiles_x;
    OutPtrs out;
    const InstG* inst;  // instanced scenes: the camera-relative instance table
    int ninst;
};

// One wave renders the 8x8 tile `tile` of job J: lane = pixel; accumulators at
// a (this wave's A_FIELDS x 64 block); LDS slices of this wave.  Adds the
// wave's ray counts to n[0..2] (per lane).
template <int W, bool INST = false>
__device__ __forceinline__ void render_tile(const TileJob& J, uint32_t tile, float* __restrict__ a,
                                            uint32_t* s_ref, float* s_dist, float* s_contrib, uint32_t lane,
                                            uint32_t& n_prim, uint32_t& n_shadow, uint32_t& n_bounce,
                                            uint32_t& n_hit) {
    const void* __restrict__ nodes = J.nodes;
    const float4* __restrict__ tris = J.tris;
    const GVertex* __restrict__ verts = J.verts;
    const uint32_t* __restrict__ idx = J.idx;
    const MatF& mat = J.mat;
    const Lights& lights = J.lights;
    const CamF& cam = J.cam;
    const uint32_t seed = J.seed;
    const int spp = J.spp, max_bounces = J.max_bounces, rr = J.rr, rr_start = J.rr_start;
    const uint32_t tiles_x = J.tiles_x;
    const uint32_t q = tile * kFusedBS + lane;
    uint32_t x, y;
    tile_xy(q, tiles_x, x, y);
    const bool valid = x < (uint32_t)cam.W && y < (uint32_t)cam.H;
    const uint32_t pix = x + y * (uint32_t)cam.W;
    if (valid) acc_clear(a, lane);
    float beta[NB], path[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) beta[k] = path[k] = 0.0f;
    f3 o = mk3(cam.ox, cam.oy, cam.oz), d = mk3(0.f, 0.f, 1.f), nd = mk3(0.f, 0.f, 1.f);
    float tmax = kInf;
    uint32_t s = 0, b = 0;
    int mode = valid ? M_NEW : M_DONE;
    bool has_next = false;
    const bool one_light = lights.n > 0;
    while (true) {
        if (mode == M_NEW) {
            if ((int)s < spp) {
                o = mk3(cam.ox, cam.oy, cam.oz);
                d = camera_dir(cam, x, y, pix, seed, s);
                tmax = kInf;
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    beta[k] = 1.0f;
                    path[k] = 0.0f;
                }
                b = 0;
                mode = M_CLOSEST;
                n_prim++;
            } else {
                mode = M_DONE;
            }
        }
        if (__ballot(mode != M_DONE) == 0) break;
        HitRec h;
        h.id = 0xffffffffu;
        h.t = kInf;
        h.b1 = h.b2 = 0.f;
        uint32_t hinst = 0;
        if (mode != M_DONE) {
            if (INST)
                h = trace_instances<W>(J.inst, J.ninst, o, d, tmax, mode == M_SHADOW, s_ref + lane, s_dist + lane,
                                       kFusedBS, hinst);
            else
                h = traverse<W>(nodes, tris, o, d, tmax, mode == M_SHADOW, s_ref + lane, s_dist + lane, kFusedBS);
        }
        bool vertex_done = false, lit = false;
        if (mode == M_CLOSEST) {
            if (h.id == 0xffffffffu) {
                // miss: the path ends (no background, cpu_path_tracer.ipp:256-270)
                if (b == 0) kahan_acc(a, lane, A_DIR_S, A_DIR_C, path, true);
                kahan_acc(a, lane, A_RAD_S, A_RAD_C, path, false);
                s++;
                mode = M_NEW;
            } else {
                Surface S = INST ? surface_instance(J.inst[hinst], h.id, h.b1, h.b2)
                                 : surface_at(verts, idx, mat, h.id, h.b1, h.b2);
                n_hit++;
                if (b == 0) accumulate_first_hit(a, lane, h.id | (hinst << kTriBits), h.t, S);
                f3 ld = mk3(0.f, 0.f, 1.f);
                float dist = 0.f, c[NB];
                bool sh = one_light &&
                          light_contrib(S, lights.l[0], beta, seed, pix, s, (b << 8) | SLOT_LIGHT0, ld, dist, c);
                if (sh) {
#pragma unroll
                    for (int k = 0; k < NB; ++k) s_contrib[k * kFusedBS + lane] = c[k];
                }
                has_next = (int)b < max_bounces && sample_bounce(S, lights, beta, seed, pix, s, b, rr, rr_start, nd);
                o = S.org;
                if (sh) {
                    d = ld;
                    tmax = dist;
                    mode = M_SHADOW;
                    n_shadow++;
                } else {
                    vertex_done = true;
                }
            }
        } else if (mode == M_SHADOW) {
            lit = h.id == 0xffffffffu;
            vertex_done = true;
        }
        if (vertex_done) {
            // radiance = 0 + lit contribution; pathRadiance += radiance (cpu_path_tracer.ipp:272,384-428);
            // the direct pass is the bounce-0 radiance (:463-465)
            float v[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                v[k] = 0.0f + (lit ? s_contrib[k * kFusedBS + lane] : 0.0f);
                path[k] = path[k] + v[k];
            }
            if (b == 0) kahan_acc(a, lane, A_DIR_S, A_DIR_C, v, false);
            if (has_next) {
                d = nd;
                tmax = kInf;
                b++;
                mode = M_CLOSEST;
                n_bounce++;
            } else {
                kahan_acc(a, lane, A_RAD_S, A_RAD_C, path, false);
                s++;
                mode = M_NEW;
            }
        }
    }
    if (valid) write_pixel(J.out, pix, spp, a, lane);
}

}  // namespace vg

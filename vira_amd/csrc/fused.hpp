// fused.hpp -- the per-pixel renderer's tile routine, shared by the
// per-request launch (render.hip) and the batched persistent launch
// (batch.hip): one wave renders one 8x8 pixel tile, one lane per pixel.
//
// CPUPathTracer::unidirectional / simulatePath / processIntersection
// (cpu_path_tracer.ipp:169-511): each lane walks its pixel's samples and
// bounces as a small state machine; every loop iteration traces exactly one
// ray per lane (camera, shadow or bounce ray, with a per-lane any-hit flag),
// so lanes whose path ended regenerate the next sample instead of idling.
// Path throughput and path radiance stay in registers, the radiance Kahan
// sums in LDS; the other per-pixel sums live in a tile-blocked accumulator
// block in global memory, the first-hit sums updated by atomic adds that
// return nothing, so a lane finishing a sample never makes its wave wait on
// global memory except for the direct pass's Kahan pair.
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
// uniform base address.  The per-pixel renderer keeps the radiance pair
// (A_RAD_S, A_RAD_C) in LDS while it renders and stores it here at the end of
// the tile; A_ALB..A_COUNT are only ever changed by relaxed device-scope
// atomics (cleared by atomic stores, summed by atomic adds, read by atomic
// loads).  One lane adds all of a pixel's samples, in sample order, and
// accesses of one work-item to one address are coherence-ordered, so each sum
// is the sequential FP32 sum.
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
    A_S1 = 46,  // adaptive sampling: sum of the samples' radiance magnitudes
    A_S2 = 47,  //                    and of their squares
    A_INST = 48,  // first-hit instance
    A_FIELDS = 49
};
// Accumulator of pixel q: base of its tile block; field f is base[f * 64 + (q & 63)].
__device__ __forceinline__ float* acc_tile(float* acc, uint32_t q) { return acc + (size_t)(q >> 6) * (A_FIELDS * 64); }
__device__ __forceinline__ const float* acc_tile(const float* acc, uint32_t q) {
    return acc + (size_t)(q >> 6) * (A_FIELDS * 64);
}
// The empty asm makes the lane offset opaque at each use, so the compiler does
// not hoist 48 per-field addresses out of the render loop into live registers.
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

// The atomically summed fields A_ALB..A_COUNT.
__device__ __forceinline__ void acc_sum_add(float* base, int f, uint32_t lane, float v) {
    __hip_atomic_fetch_add(&acc_ref(base, f, lane), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float acc_sum_get(const float* base, int f, uint32_t lane) {
    return __hip_atomic_load(&acc_ref(base, f, lane), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void acc_sum_set(float* base, int f, uint32_t lane, float v) {
    __hip_atomic_store(&acc_ref(base, f, lane), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// First-hit sums at bounce 0 (cpu_path_tracer.ipp:352-381): albedo, normal,
// depth and the hit count.
__device__ __forceinline__ void accumulate_first_hit_sums(float* __restrict__ a, uint32_t lane, float t,
                                                          const Surface& S) {
#pragma unroll
    for (int k = 0; k < NB; ++k) acc_sum_add(a, A_ALB + k, lane, S.alb[k]);
    acc_sum_add(a, A_NRM + 0, lane, S.N.x);
    acc_sum_add(a, A_NRM + 1, lane, S.N.y);
    acc_sum_add(a, A_NRM + 2, lane, S.N.z);
    acc_sum_add(a, A_DEPTH, lane, t);
    acc_sum_add(a, A_COUNT, lane, 1.0f);
}

// The same plus the first hit's triangle and instance, found by reading the
// triangle field (the wavefront schedule, whose samples are separate launches).
__device__ __forceinline__ void accumulate_first_hit(float* __restrict__ a, uint32_t lane, uint32_t id,
                                                     uint32_t inst, float t, const Surface& S) {
    if (__float_as_uint(ACC(a, A_TRI, lane)) == 0xffffffffu) {
        ACC(a, A_TRI, lane) = __uint_as_float(id);
        ACC(a, A_INST, lane) = __uint_as_float(inst);
    }
    accumulate_first_hit_sums(a, lane, t, S);
}

struct OutPtrs {
    float *radiance, *direct, *indirect, *albedo, *normal, *depth, *alpha;
    uint32_t* tri;
    uint32_t* inst;
    float* normal_camera;
    uint32_t* mesh_id;
    uint32_t* material_id;
};

// Per-request values the passes need besides the sums: the view normal
// matrix (row-major) and the mesh/material ids of an uninstanced scene.
struct PassInfo {
    float vn[9];
    uint32_t mesh_id, material_id;
};

// Averages and passes (cpu_path_tracer.ipp:229-239, passes.ipp:81-127); ns is
// the sample count the reference divides by (the samples taken, or the index
// of the sample at which adaptive sampling stopped).
__device__ __forceinline__ void write_pixel(const OutPtrs& o, size_t p, int ns, const float* __restrict__ a,
                                            uint32_t lane, const PassInfo& pi, const void* inst_table) {
    float sf = (float)ns;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        float rad = __fdiv_rn(ACC(a, A_RAD_S + k, lane), sf);
        float dr = __fdiv_rn(ACC(a, A_DIR_S + k, lane), sf);
        if (o.radiance) o.radiance[NB * p + k] = rad;
        if (o.direct) o.direct[NB * p + k] = dr;
        if (o.indirect) o.indirect[NB * p + k] = rad - dr;
        if (o.albedo) o.albedo[NB * p + k] = __fdiv_rn(acc_sum_get(a, A_ALB + k, lane), sf);
    }
    float cnt = acc_sum_get(a, A_COUNT, lane);
    if (o.normal || o.normal_camera) {
        f3 n = mk3(acc_sum_get(a, A_NRM, lane), acc_sum_get(a, A_NRM + 1, lane), acc_sum_get(a, A_NRM + 2, lane));
        if (cnt > 0.0f) n = normalize(n);
        if (o.normal) {
            o.normal[3 * p] = n.x;
            o.normal[3 * p + 1] = n.y;
            o.normal[3 * p + 2] = n.z;
        }
        if (o.normal_camera) {  // normalize(getViewNormalMatrix() * normal_global) (cpu_path_tracer.ipp:237-239)
            f3 c = mk3(0.f, 0.f, 0.f);
            if (cnt > 0.0f)
                c = normalize(mk3((pi.vn[0] * n.x + pi.vn[1] * n.y) + pi.vn[2] * n.z,
                                  (pi.vn[3] * n.x + pi.vn[4] * n.y) + pi.vn[5] * n.z,
                                  (pi.vn[6] * n.x + pi.vn[7] * n.y) + pi.vn[8] * n.z));
            o.normal_camera[3 * p] = c.x;
            o.normal_camera[3 * p + 1] = c.y;
            o.normal_camera[3 * p + 2] = c.z;
        }
    }
    if (o.depth) o.depth[p] = cnt > 0.0f ? __fdiv_rn(acc_sum_get(a, A_DEPTH, lane), cnt) : kInf;
    if (o.alpha) o.alpha[p] = __fdiv_rn(cnt, sf);
    const uint32_t first = __float_as_uint(ACC(a, A_TRI, lane));
    const uint32_t finst = __float_as_uint(ACC(a, A_INST, lane));
    const bool none = first == 0xffffffffu;
    if (o.tri) o.tri[p] = first;
    if (o.inst) o.inst[p] = none ? first : finst;
    if (o.mesh_id || o.material_id) {
        uint32_t mid = pi.mesh_id, tid = pi.material_id;
        if (inst_table && !none) {
            const InstG& I = static_cast<const InstG*>(inst_table)[finst];
            mid = I.mesh_id;
            tid = I.material_id;
        }
        if (o.mesh_id) o.mesh_id[p] = none ? 0xffffffffu : mid;
        if (o.material_id) o.material_id[p] = none ? 0xffffffffu : tid;
    }
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
    for (int f = 0; f < A_FIELDS; ++f) {
        if (f >= A_ALB && f <= A_COUNT) acc_sum_set(a, f, lane, 0.0f);
        else ACC(a, f, lane) = f == A_TRI ? __uint_as_float(0xffffffffu) : 0.0f;
    }
}

// Kahan-compensated add of v[] into a lane's pair of NB sums and NB
// compensations kept in LDS (column r, stride kFusedBS).
__device__ __forceinline__ void kahan_lds(float* r, const float* v) {
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        float sv = r[k * kFusedBS], cv = r[(NB + k) * kFusedBS];
        kahan_add(sv, cv, v[k]);
        r[k * kFusedBS] = sv;
        r[(NB + k) * kFusedBS] = cv;
    }
}

enum : int { M_NEW = 0, M_CLOSEST = 1, M_SHADOW = 2, M_DONE = 3 };

// Adaptive sampling (CPUPathTracerOptions::adaptive_sampling and its three
// parameters, cpu_path_tracer.hpp:39-42); adaptive = 0 takes every sample.
struct Adaptive {
    int on;
    int batch;       // samples_per_batch
    int miss;        // samples_to_detect_miss
    float tol;       // sampling_tolerance
};

// What a render request needs besides the render loop's own inputs: the
// output passes, the light environment, adaptive sampling and the pass
// values.  It lives in device memory and the loop reaches it by one pointer,
// so the kernel's uniform registers hold only what the traversal uses.
struct JobExt {
    OutPtrs out;
    EnvF env;
    Adaptive adapt;
    PassInfo pass;
};

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
    const InstG* inst;  // instanced scenes: the camera-relative instance table
    const TNode* tlas;  // and its TLAS nodes
    int ntlas;
    const JobExt* ext;  // device memory
};

// Per-lane LDS of the several-light renderer: the vertex's throughput before
// the bounce update, shading normal, albedo and its radiance so far.
enum : int { ML_BETA = 0, ML_N = 8, ML_ALB = 11, ML_VRAD = 19, ML_FIELDS = 27 };
// Per-lane LDS floats of the per-pixel renderers besides the stack: the staged
// light contribution (one light) or the vertex (several lights), then the
// radiance Kahan pair (2 x NB) from fused_rad_lds on.
template <bool MULTI>
constexpr int fused_rad_lds() { return MULTI ? ML_FIELDS : NB; }
template <bool MULTI>
constexpr int fused_lds_floats() { return fused_rad_lds<MULTI>() + 2 * NB; }

// End of sample s of a lane's pixel: fold the path radiance into the Kahan
// sums (rk: the lane's LDS pair) and, with adaptive sampling, apply the
// reference's stopping rules (cpu_path_tracer.ipp:198-227).  Returns true when
// the pixel is finished (s is then the divisor, as dataPayload.sample is in
// the reference).
template <bool FEAT>
__device__ __forceinline__ bool end_sample(float* __restrict__ a, uint32_t lane, const float path[NB],
                                           const Adaptive& ad, uint32_t& s, float* rk) {
    kahan_lds(rk, path);
    if (FEAT && ad.on) {
        float sq = 0.0f;
#pragma unroll
        for (int k = 0; k < NB; ++k) sq = sq + path[k] * path[k];
        const float illum = __builtin_sqrtf(sq);
        const float s1 = ACC(a, A_S1, lane) + illum;
        const float s2 = ACC(a, A_S2, lane) + illum * illum;
        ACC(a, A_S1, lane) = s1;
        ACC(a, A_S2, lane) = s2;
        bool stop = false;
        if (s % (uint32_t)ad.miss == 0 && s != 0 && acc_sum_get(a, A_DEPTH, lane) == 0.0f) stop = true;
        if (!stop && s % (uint32_t)ad.batch == 0 && s != 0) {
            if (s1 == 0.0f) {
                stop = true;
            } else {
                const float n = (float)s + 1.0f;
                const float mu = __fdiv_rn(s1, n);
                const float var = __fdiv_rn(1.0f, n - 1.0f) * (s2 - __fdiv_rn(s1 * s1, n));
                const float I = 1.96f * __builtin_sqrtf(__fdiv_rn(var, n));
                if (I <= ad.tol * mu) stop = true;
            }
        }
        if (stop) return true;
    }
    s++;
    return false;
}

// One wave renders the 8x8 tile `tile` of job J: lane = pixel; accumulators at
// a (this wave's A_FIELDS x 64 block); LDS slices of this wave (s_lds holds
// fused_lds_floats<MULTI>() x 64 floats).  Adds the wave's ray counts to the
// per-lane counters.
//
// FEAT = false compiles out background, adaptive sampling, normal maps and
// spectral albedo maps (the caller picks FEAT = true when any is in use).
// MULTI = false: at most one light and no ambient light; the light's
// contribution is staged in LDS while its shadow ray is traced.
// MULTI = true: any number of lights and ambient light; the vertex's
// throughput, normal, albedo and radiance stay in LDS while the shadow rays of
// the lights are traced one after the other, and each light's (deterministic)
// sample is drawn again when its shadow ray returns unoccluded.  Both give the
// numbers of the reference's loop (cpu_path_tracer.ipp:384-428): radiance =
// 0 (+ ambient at bounce 0) + the lit lights' terms in light order.
//
// STATS = true adds each traversal's visits to tcount[0] (camera and bounce
// rays) or tcount[1] (shadow rays); maxsp keeps the deepest stack.
template <int W, bool INST = false, bool MULTI = false, bool FEAT = false, bool STATS = false>
__device__ __forceinline__ void render_tile(const TileJob& J, uint32_t tile, float* __restrict__ a,
                                            uint32_t* s_ref, float* s_dist, float* s_lds, uint32_t lane,
                                            uint32_t& n_prim, uint32_t& n_shadow, uint32_t& n_bounce,
                                            uint32_t& n_hit, TravCount* tcount = nullptr) {
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
    float* const rk = s_lds + lane + fused_rad_lds<MULTI>() * kFusedBS;  // radiance Kahan pair
    bool first_seen = false;  // the pixel's first hit is recorded
    if (valid) {
        acc_clear(a, lane);
#pragma unroll
        for (int k = 0; k < 2 * NB; ++k) rk[k * kFusedBS] = 0.0f;
    }
    float beta[NB], path[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) beta[k] = path[k] = 0.0f;
    f3 o = mk3(cam.ox, cam.oy, cam.oz), d = mk3(0.f, 0.f, 1.f), nd = mk3(0.f, 0.f, 1.f);
    float tmax = kInf;
    uint32_t s = 0, b = 0;
    int mode = valid ? M_NEW : M_DONE;
    bool has_next = false;
    int li = 0;  // MULTI: the light whose shadow ray is in flight
    const bool one_light = lights.n > 0;
    float* const sm = s_lds + lane;  // this lane's LDS column (stride kFusedBS)
    // MULTI: from light l0 on, the first light that can be sampled gets its
    // shadow ray; false when none is left
    auto next_light = [&](int l0) -> bool {
        for (int l = l0; l < lights.n; ++l) {
            float lpdf, lrad[NB], dist;
            f3 ld;
            sample_light(lights.l[l], o, seed, pix, s, (b << 8) | (SLOT_LIGHT0 + 2u * (uint32_t)l), ld, dist, lpdf,
                         lrad);
            if (lpdf > 0.0f) {
                d = ld;
                tmax = dist;
                li = l;
                mode = M_SHADOW;
                n_shadow++;
                return true;
            }
        }
        return false;
    };
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
            TravCount tc{};
            if (INST)
                h = trace_instances<W, STATS>(J.inst, J.tlas, J.ntlas, o, d, tmax, mode == M_SHADOW, s_ref + lane,
                                              s_dist + lane, kFusedBS, hinst, &tc);
            else
                h = traverse<W, STATS>(nodes, tris, o, d, tmax, mode == M_SHADOW, s_ref + lane, s_dist + lane,
                                       kFusedBS, &tc);
            if (STATS) {
                auto add = [&](TravCount& to) {
                    to.nodes += tc.nodes;
                    to.leaves += tc.leaves;
                    to.tris += tc.tris;
                    to.spills += tc.spills;
                    to.tlas += tc.tlas;
                    to.wsteps += tc.wsteps;
                    to.maxsp = to.maxsp > tc.maxsp ? to.maxsp : tc.maxsp;
                };
                if (mode == M_SHADOW) add(tcount[1]);
                else add(tcount[0]);
            }
        }
        bool vertex_done = false, lit = false;
        if (mode == M_CLOSEST) {
            if (h.id == 0xffffffffu) {
                // miss: background radiance when shown (cpu_path_tracer.ipp:256-270)
                if (FEAT && J.ext->env.show_bg) {
                    float bg[NB];
                    background_radiance(J.ext->env, d, bg);
#pragma unroll
                    for (int k = 0; k < NB; ++k) path[k] = path[k] + beta[k] * (bg[k] * 1.0f);
                }
                if (b == 0) kahan_acc(a, lane, A_DIR_S, A_DIR_C, path, true);
                mode = end_sample<FEAT>(a, lane, path, J.ext->adapt, s, rk) ? M_DONE : M_NEW;
            } else {
                Surface S = INST ? surface_instance<FEAT>(J.inst[hinst], h.id, h.b1, h.b2)
                                 : surface_at<FEAT>(verts, idx, mat, h.id, h.b1, h.b2);
                n_hit++;
                if (b == 0) {
                    if (!first_seen) {
                        ACC(a, A_TRI, lane) = __uint_as_float(h.id);
                        ACC(a, A_INST, lane) = __uint_as_float(hinst);
                        first_seen = true;
                    }
                    accumulate_first_hit_sums(a, lane, h.t, S);
                }
                o = S.org;
                if (MULTI) {
                    // vertex radiance starts at 0 (+ ambient x albedo at bounce 0, cpu_path_tracer.ipp:389-396)
#pragma unroll
                    for (int k = 0; k < NB; ++k) {
                        float vr = 0.0f;
                        if (b == 0 && J.ext->env.has_ambient) vr = vr + beta[k] * (J.ext->env.ambient[k] * S.alb[k]);
                        sm[(ML_VRAD + k) * kFusedBS] = vr;
                        sm[(ML_BETA + k) * kFusedBS] = beta[k];
                        sm[(ML_ALB + k) * kFusedBS] = S.alb[k];
                    }
                    sm[(ML_N + 0) * kFusedBS] = S.N.x;
                    sm[(ML_N + 1) * kFusedBS] = S.N.y;
                    sm[(ML_N + 2) * kFusedBS] = S.N.z;
                    has_next =
                        (int)b < max_bounces && sample_bounce(S, lights, beta, seed, pix, s, b, rr, rr_start, nd);
                    if (!next_light(0)) vertex_done = true;
                } else {
                    f3 ld = mk3(0.f, 0.f, 1.f);
                    float dist = 0.f, c[NB];
                    bool sh = one_light && light_contrib(S, lights.l[0], beta, seed, pix, s,
                                                         (b << 8) | SLOT_LIGHT0, ld, dist, c);
                    if (sh) {
#pragma unroll
                        for (int k = 0; k < NB; ++k) sm[k * kFusedBS] = c[k];
                    }
                    has_next =
                        (int)b < max_bounces && sample_bounce(S, lights, beta, seed, pix, s, b, rr, rr_start, nd);
                    if (sh) {
                        d = ld;
                        tmax = dist;
                        mode = M_SHADOW;
                        n_shadow++;
                    } else {
                        vertex_done = true;
                    }
                }
            }
        } else if (mode == M_SHADOW) {
            lit = h.id == 0xffffffffu;
            if (MULTI) {
                if (lit) {
                    // this light's term, drawn again from the kept surface values
                    float bo[NB], al[NB], c[NB], dist;
                    f3 ld;
#pragma unroll
                    for (int k = 0; k < NB; ++k) {
                        bo[k] = sm[(ML_BETA + k) * kFusedBS];
                        al[k] = sm[(ML_ALB + k) * kFusedBS];
                    }
                    const f3 Nv = mk3(sm[ML_N * kFusedBS], sm[(ML_N + 1) * kFusedBS], sm[(ML_N + 2) * kFusedBS]);
                    light_contrib_at(o, Nv, al, lights.l[li], bo, seed, pix, s,
                                     (b << 8) | (SLOT_LIGHT0 + 2u * (uint32_t)li), ld, dist, c);
#pragma unroll
                    for (int k = 0; k < NB; ++k) sm[(ML_VRAD + k) * kFusedBS] = sm[(ML_VRAD + k) * kFusedBS] + c[k];
                }
                if (!next_light(li + 1)) vertex_done = true;
            } else {
                vertex_done = true;
            }
        }
        if (vertex_done) {
            // pathRadiance += radiance (cpu_path_tracer.ipp:272,384-428); the direct pass is the
            // bounce-0 radiance (:463-465)
            float v[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                if (MULTI) v[k] = sm[(ML_VRAD + k) * kFusedBS];
                else v[k] = 0.0f + (lit ? sm[k * kFusedBS] : 0.0f);
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
                mode = end_sample<FEAT>(a, lane, path, J.ext->adapt, s, rk) ? M_DONE : M_NEW;
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            ACC(a, A_RAD_S + k, lane) = rk[k * kFusedBS];
            ACC(a, A_RAD_C + k, lane) = rk[(NB + k) * kFusedBS];
        }
        write_pixel(J.ext->out, pix, (int)s, a, lane, J.ext->pass, INST ? (const void*)J.inst : nullptr);
    }
}

}  // namespace vg

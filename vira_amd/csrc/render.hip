// render.hip -- the path tracer and the trace_rays hook.
//
// CPUPathTracer::render / unidirectional / simulatePath / processIntersection
// (cpu_path_tracer.ipp:29-511) on the GPU, in two schedules that share every
// formula (shading.hpp) and produce the same numbers:
//
// * fused (the default; any number of lights): one wave per 8x8 pixel tile,
//   one lane per pixel.  Each lane walks its pixel's samples and bounces as a
//   small state machine; every loop iteration traces exactly one ray per lane
//   (camera, shadow or bounce ray, with a per-lane any-hit flag), so lanes
//   whose path ended regenerate the next sample instead of idling.  Path
//   throughput and path radiance stay in registers and the radiance Kahan
//   sums in LDS; the direct pass's Kahan sums and the first-hit passes live in
//   the tile-blocked accumulator block in global memory (the first-hit sums
//   updated by atomic adds that return nothing).
// * wavefront (point and sphere lights, no adaptive sampling, ambient light or
//   shown background): queues of rays, closest-hit and any-hit
//   traversal kernels, shading into per-hit shadow slots, bounce rays
//   compacted with a wave ballot and one atomic per wave.
//
// Both fold samples into Kahan-compensated FP32 sums in sample order, so the
// image does not depend on scheduling.
#include <cmath>
#include <memory>
#include <mutex>
#include <string>

#include "context.hpp"
#include "fused.hpp"
#include "scene_params.hpp"
#include "shading.hpp"
#include "traverse.hpp"

namespace vg {
namespace {

constexpr int kTraceBS = 128;  // traversal block size of the wavefront kernels (2 waves)
constexpr uint32_t kNoPath = 0xffffffffu;

enum : int { P_BETA = 0, P_RAD = 8, P_DIR = 16, P_FIELDS = 24 };

// Compacted append: returns this lane's slot (only meaningful where pred).
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    if (!mask) return 0;
    int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    unsigned long long below = mask & ((1ull << __lane_id()) - 1ull);
    return base + (uint32_t)__popcll(below);
}

// ------------------------------------------------------------ trace_rays hook
template <int W>
__global__ void __launch_bounds__(kTraceBS) k_trace_rays(const void* __restrict__ nodes, const float4* __restrict__ tris,
                                                         const float* __restrict__ rays, size_t n, int any,
                                                         float* t_out, uint32_t* id_out, float* b1_out, float* b2_out) {
    __shared__ uint32_t s_ref[kStackLds * kTraceBS];
    __shared__ float s_dist[kStackLds * kTraceBS];
    size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float4 a = reinterpret_cast<const float4*>(rays)[2 * r];
    const float4 b = reinterpret_cast<const float4*>(rays)[2 * r + 1];
    HitRec h = traverse<W>(nodes, tris, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), b.w, any != 0, s_ref + threadIdx.x,
                        s_dist + threadIdx.x, kTraceBS);
    if (t_out) t_out[r] = any ? (h.id != 0xffffffffu ? 0.0f : kInf) : h.t;
    if (id_out) id_out[r] = h.id;
    if (b1_out) b1_out[r] = h.b1;
    if (b2_out) b2_out[r] = h.b2;
}

// ============================================================ fused schedule
// Per-request launch: block b renders tile b.
template <int W, bool INST, bool MULTI, bool FEAT>
__global__ void __launch_bounds__(kFusedBS, kFusedWavesPerSimd) k_render_fused(TileJob J, float* __restrict__ acc,
                                                         unsigned long long* __restrict__ counters) {
    __shared__ uint32_t s_ref[kStackLds * kFusedBS];
    __shared__ float s_dist[kStackLds * kFusedBS];
    // staged light contribution (one light) or the vertex state (several lights), then the
    // radiance Kahan pair
    __shared__ float s_lds[fused_lds_floats<MULTI>() * kFusedBS];
    const uint32_t lane = threadIdx.x;
    uint32_t n_prim = 0, n_shadow = 0, n_bounce = 0, n_hit = 0;
    render_tile<W, INST, MULTI, FEAT>(J, blockIdx.x, acc + (size_t)blockIdx.x * (A_FIELDS * 64), s_ref, s_dist, s_lds,
                                lane, n_prim, n_shadow, n_bounce, n_hit);
    wave_add(&counters[0], n_prim);
    wave_add(&counters[1], n_shadow);
    wave_add(&counters[2], n_bounce);
    wave_add(&counters[3], n_hit);
}

// Counter words of the visit-counting launch: after the four ray counts, seven
// per ray kind (camera and bounce rays, then shadow rays): nodes, leaves,
// triangles, spilled pushes, deepest stack, TLAS nodes, wave node steps.
enum : int { kVisitBase = 4, kVisitWords = 7 };

// The same render with the traversals' visit counts (opt-in; one light or
// none, no background, adaptive sampling or material maps).
template <int W, bool INST>
__global__ void __launch_bounds__(kFusedBS, kFusedWavesPerSimd) k_render_fused_stats(TileJob J, float* __restrict__ acc,
                                                         unsigned long long* __restrict__ counters) {
    __shared__ uint32_t s_ref[kStackLds * kFusedBS];
    __shared__ float s_dist[kStackLds * kFusedBS];
    __shared__ float s_lds[fused_lds_floats<false>() * kFusedBS];
    const uint32_t lane = threadIdx.x;
    uint32_t n_prim = 0, n_shadow = 0, n_bounce = 0, n_hit = 0;
    TravCount tc[2] = {};
    render_tile<W, INST, false, false, true>(J, blockIdx.x, acc + (size_t)blockIdx.x * (A_FIELDS * 64), s_ref, s_dist,
                                             s_lds, lane, n_prim, n_shadow, n_bounce, n_hit, tc);
    wave_add(&counters[0], n_prim);
    wave_add(&counters[1], n_shadow);
    wave_add(&counters[2], n_bounce);
    wave_add(&counters[3], n_hit);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        unsigned long long* c = counters + kVisitBase + k * kVisitWords;
        wave_add64(&c[0], tc[k].nodes);
        wave_add64(&c[1], tc[k].leaves);
        wave_add64(&c[2], tc[k].tris);
        wave_add64(&c[3], tc[k].spills);
        wave_max(&c[4], tc[k].maxsp);
        wave_add64(&c[5], tc[k].tlas);
        wave_add64(&c[6], tc[k].wsteps);
    }
}

// Camera-relative FP32 placement of every instance and light, computed in
// FP64 from the FP64 model matrices and positions (one thread per item).
__global__ void k_rebase(const double* __restrict__ M, int n, const double* __restrict__ lpos, int nl, double cx,
                         double cy, double cz, InstG* __restrict__ inst, LightF* __restrict__ lights) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rebase_instance(M + 16 * i, cx, cy, cz, inst[i]);
    if (i < nl) {
        lights[i].px = (float)(lpos[3 * i] - cx);
        lights[i].py = (float)(lpos[3 * i + 1] - cy);
        lights[i].pz = (float)(lpos[3 * i + 2] - cz);
    }
}

// Camera-relative TLAS node boxes: a leaf takes its instance's placed box, an
// internal node the union of the leaves of its subtree ([i + 1, skip) in
// depth-first order), so every box holds what its subtree holds, exactly.
__global__ void k_tlas_boxes(const int2* __restrict__ topo, int nn, const InstG* __restrict__ inst,
                             TNode* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int2 t = topo[i];
    TNode N;
    N.inst = t.x;
    N.skip = t.y;
    for (int k = 0; k < 3; ++k) {
        N.lo[k] = __builtin_huge_valf();
        N.hi[k] = -__builtin_huge_valf();
    }
    for (int j = i; j < t.y; ++j) {
        const int leaf = topo[j].x;
        if (leaf < 0) continue;
        const InstG& I = inst[leaf];
        for (int k = 0; k < 3; ++k) {
            N.lo[k] = fminf(N.lo[k], I.wlo[k]);
            N.hi[k] = fmaxf(N.hi[k], I.whi[k]);
        }
    }
    out[i] = N;
}

// ============================================================ wavefront schedule
// Queue entries carry the tile-ordered path index q in origin.w (kNoPath = none).
__global__ void k_camera(CamF cam, uint32_t seed, uint32_t s, uint32_t nq, uint32_t tiles_x, float4* ray_o,
                         float4* ray_d, float* path, unsigned long long* cnt_primary) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint32_t x, y;
    tile_xy(q, tiles_x, x, y);
    bool valid = x < (uint32_t)cam.W && y < (uint32_t)cam.H;
    wave_add(cnt_primary, valid ? 1u : 0u);
    if (!valid) {
        ray_o[q] = make_float4(0.f, 0.f, 0.f, __uint_as_float(kNoPath));
        return;
    }
    uint32_t pix = x + y * (uint32_t)cam.W;
    f3 d = camera_dir(cam, x, y, pix, seed, s);
    ray_o[q] = make_float4(cam.ox, cam.oy, cam.oz, __uint_as_float(q));
    ray_d[q] = make_float4(d.x, d.y, d.z, kInf);
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        path[(P_BETA + k) * nq + q] = 1.0f;
        path[(P_RAD + k) * nq + q] = 0.0f;
        path[(P_DIR + k) * nq + q] = 0.0f;
    }
}

template <int W>
__global__ void __launch_bounds__(kTraceBS) k_trace_queue(const void* __restrict__ nodes,
                                                          const float4* __restrict__ tris,
                                                          const float4* __restrict__ ray_o,
                                                          const float4* __restrict__ ray_d,
                                                          const uint32_t* __restrict__ n_ptr, uint32_t n_fixed,
                                                          uint32_t mult, int any, float4* __restrict__ hit) {
    __shared__ uint32_t s_ref[kStackLds * kTraceBS];
    __shared__ float s_dist[kStackLds * kTraceBS];
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = (n_ptr ? *n_ptr : n_fixed) * mult;  // queue entries x slots per entry
    if (e >= n) return;
    float4 o = ray_o[e];
    if (__float_as_uint(o.w) == kNoPath) return;
    float4 d = ray_d[e];
    HitRec h = traverse<W>(nodes, tris, mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z), d.w, any != 0, s_ref + threadIdx.x,
                        s_dist + threadIdx.x, kTraceBS);
    hit[e] = make_float4(h.t, h.b1, h.b2, __uint_as_float(h.id));
}

__global__ void __launch_bounds__(128) k_shade(
    const float4* __restrict__ ray_o, const float4* __restrict__ hit, const uint32_t* __restrict__ nq_ptr,
    uint32_t nq_fixed, uint32_t nq, uint32_t tiles_x, int W, const GVertex* __restrict__ verts,
    const uint32_t* __restrict__ idx, MatF mat, Lights lights, uint32_t seed, uint32_t s, uint32_t bounce,
    int max_bounces, int rr, int rr_start, float* __restrict__ path, float* __restrict__ acc,
    float4* __restrict__ s_o, float4* __restrict__ s_d, float* __restrict__ s_contrib, uint32_t nslots,
    float4* __restrict__ next_o, float4* __restrict__ next_d, uint32_t* __restrict__ next_count,
    unsigned long long* __restrict__ counters) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = nq_ptr ? *nq_ptr : nq_fixed;
    if (e >= n) return;
    const int nl = lights.n;
    uint32_t q = __float_as_uint(ray_o[e].w);
    float4 h4 = q == kNoPath ? make_float4(kInf, 0.f, 0.f, __uint_as_float(0xffffffffu)) : hit[e];
    uint32_t id = __float_as_uint(h4.w);
    bool alive = q != kNoPath && id != 0xffffffffu;
    for (int l = 0; l < nl; ++l) s_o[(size_t)e * nl + l].w = __uint_as_float(kNoPath);
    bool emit = false;
    f3 org = mk3(0.f, 0.f, 0.f), nd = mk3(0.f, 0.f, 1.f);
    uint32_t nsh = 0;
    if (alive) {
        uint32_t x, y;
        tile_xy(q, tiles_x, x, y);
        uint32_t pix = x + y * (uint32_t)W;
        Surface S = surface_at(verts, idx, mat, id, h4.y, h4.z);
        if (bounce == 0) accumulate_first_hit(acc_tile(acc, q), q & 63u, id, 0u, h4.x, S);
        org = S.org;
        float beta[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) beta[k] = path[(P_BETA + k) * nq + q];
        for (int l = 0; l < nl; ++l) {
            f3 ld;
            float dist, c[NB];
            if (!light_contrib(S, lights.l[l], beta, seed, pix, s, (bounce << 8) | (SLOT_LIGHT0 + 2 * l), ld, dist, c))
                continue;
            size_t slot = (size_t)e * nl + l;
#pragma unroll
            for (int k = 0; k < NB; ++k) s_contrib[(size_t)k * nslots + slot] = c[k];
            s_o[slot] = make_float4(org.x, org.y, org.z, __uint_as_float(q));
            s_d[slot] = make_float4(ld.x, ld.y, ld.z, dist);
            nsh++;
        }
        if ((int)bounce < max_bounces) {
            emit = sample_bounce(S, lights, beta, seed, pix, s, bounce, rr, rr_start, nd);
            if (emit) {
#pragma unroll
                for (int k = 0; k < NB; ++k) path[(P_BETA + k) * nq + q] = beta[k];
            }
        }
    }
    uint32_t slot = wave_append(next_count, emit);
    if (emit) {
        next_o[slot] = make_float4(org.x, org.y, org.z, __uint_as_float(q));
        next_d[slot] = make_float4(nd.x, nd.y, nd.z, kInf);
    }
    wave_add(&counters[1], nsh);
    wave_add(&counters[2], emit ? 1u : 0u);
}

// Ordered sum of the lit contributions of each hit in light order, as the
// `radiance +=` loop of cpu_path_tracer.ipp:399-428; a hit with no lit light
// still adds its (zero) vertex radiance.
__global__ void k_accum(const float4* __restrict__ ray_o, const float4* __restrict__ hit,
                        const uint32_t* __restrict__ nq_ptr, uint32_t nq_fixed, uint32_t nq, int nl,
                        const float4* __restrict__ s_o, const float4* __restrict__ s_hit,
                        const float* __restrict__ s_contrib, uint32_t nslots, uint32_t bounce,
                        float* __restrict__ path) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = nq_ptr ? *nq_ptr : nq_fixed;
    if (e >= n) return;
    uint32_t q = __float_as_uint(ray_o[e].w);
    if (q == kNoPath || __float_as_uint(hit[e].w) == 0xffffffffu) return;
    float vrad[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) vrad[k] = 0.0f;
    for (int l = 0; l < nl; ++l) {
        size_t slot = (size_t)e * nl + l;
        if (__float_as_uint(s_o[slot].w) == kNoPath) continue;
        if (__float_as_uint(s_hit[slot].w) != 0xffffffffu) continue;  // occluded
#pragma unroll
        for (int k = 0; k < NB; ++k) vrad[k] = vrad[k] + s_contrib[(size_t)k * nslots + slot];
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) path[(P_RAD + k) * nq + q] = path[(P_RAD + k) * nq + q] + vrad[k];
    if (bounce == 0) {
#pragma unroll
        for (int k = 0; k < NB; ++k) path[(P_DIR + k) * nq + q] = path[(P_DIR + k) * nq + q] + vrad[k];
    }
}

// Kahan-compensated fold of one sample pass into the pixel sums.
__global__ void k_resolve(uint32_t nq, const float* __restrict__ path, float* __restrict__ acc) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    float v[NB], w[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        v[k] = path[(P_RAD + k) * nq + q];
        w[k] = path[(P_DIR + k) * nq + q];
    }
    float* a = acc_tile(acc, q);
    kahan_acc(a, q & 63u, A_RAD_S, A_RAD_C, v, false);
    kahan_acc(a, q & 63u, A_DIR_S, A_DIR_C, w, false);
}

__global__ void k_acc_init(uint32_t nq, float* acc) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    acc_clear(acc_tile(acc, q), q & 63u);
}

__global__ void k_finalize(uint32_t nq, uint32_t tiles_x, int W, int H, int spp, const float* __restrict__ acc,
                           OutPtrs out, PassInfo pi) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint32_t x, y;
    tile_xy(q, tiles_x, x, y);
    if (x >= (uint32_t)W || y >= (uint32_t)H) return;
    write_pixel(out, x + (size_t)y * W, spp, acc_tile(acc, q), q & 63u, pi, nullptr);
}

}  // namespace

void trace_rays(Ctx& c, MeshG& m, const float* rays, size_t n, int any_hit, float* t, uint32_t* id, float* b1,
                float* b2, double* ms) {
    VG_CHECK(m.blas.built, VIRA_GPU_ERR_STATE, "trace_rays: BLAS not built");
    if (n == 0) {
        if (ms) *ms = 0;
        return;
    }
    EventPair ev;
    VG_HIP(hipEventRecord(ev.a, c.stream));
    if (m.blas.width == 4 && m.blas.compressed)
        k_trace_rays<kWidthQ><<<blocks_for(n, kTraceBS), kTraceBS, 0, c.stream>>>(
            m.blas.qnodes.p, m.blas.tris.as<float4>(), rays, n, any_hit, t, id, b1, b2);
    else if (m.blas.width == 4)
        k_trace_rays<4><<<blocks_for(n, kTraceBS), kTraceBS, 0, c.stream>>>(m.blas.nodes.p, m.blas.tris.as<float4>(),
                                                                            rays, n, any_hit, t, id, b1, b2);
    else
        k_trace_rays<2><<<blocks_for(n, kTraceBS), kTraceBS, 0, c.stream>>>(m.blas.nodes.p, m.blas.tris.as<float4>(),
                                                                            rays, n, any_hit, t, id, b1, b2);
    VG_HIP(hipGetLastError());
    VG_HIP(hipEventRecord(ev.b, c.stream));
    VG_HIP(hipEventSynchronize(ev.b));
    float f = 0.f;
    VG_HIP(hipEventElapsedTime(&f, ev.a, ev.b));
    if (ms) *ms = f;
}

namespace {

// Scratch of one fused (per-pixel) render: the tile accumulator blocks, the
// ray counters, the light table and, for instanced scenes, the instance table
// with its FP64 inputs.  vira_gpu_render takes these from buffers the context
// grows on demand; vira_gpu_render_noalloc from the file-scope array below.
struct FusedScratch {
    float* acc;                    // A_FIELDS x 64 floats per tile
    unsigned long long* counters;  // kCounterWords
    LightF* lights;                // VIRA_GPU_MAX_LIGHTS
    InstG* inst;                   // ninst
    double* inst_m;                // 16 per instance
    double* light_pos;             // 3 per light
    JobExt* ext;                   // one
    TNode* tnodes;                 // TLAS nodes of the instanced scene
    bool host_staging;             // the instance table may be staged through host memory the call allocates
};
constexpr int kCounterWords = 64;

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int n) {
    const int i = threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ---------------------------------------------------------------- allocation-free render
// vira_gpu_render_noalloc's scratch: one array per GPU, living as long as the
// process.  Accumulator blocks of up to kNoAllocMaxTiles 8x8 tiles, then the
// tables.  Host code never names the array: it uses the device address
// hipGetSymbolAddress returns (kept per device by noalloc_init).
constexpr uint32_t kNoAllocMaxTiles = 4096;                                     // e.g. 512 x 512 pixels
constexpr size_t kNoAllocAccBytes = (size_t)kNoAllocMaxTiles * A_FIELDS * 64 * 4;  // 46 MiB
constexpr int kNoAllocMaxInstances = 64;
constexpr size_t kNoAllocCounterOff = kNoAllocAccBytes;
constexpr size_t kNoAllocLightOff = kNoAllocCounterOff + kCounterWords * 8;
constexpr size_t kNoAllocLightPosOff = kNoAllocLightOff + 1024;
constexpr size_t kNoAllocInstOff = kNoAllocLightPosOff + 256;
constexpr size_t kNoAllocInstMOff = kNoAllocInstOff + sizeof(InstG) * kNoAllocMaxInstances;
constexpr size_t kNoAllocExtOff = kNoAllocInstMOff + 16 * 8 * kNoAllocMaxInstances;
constexpr size_t kNoAllocTNodeOff = kNoAllocExtOff + ((sizeof(JobExt) + 255) / 256) * 256;
constexpr size_t kNoAllocBytes = kNoAllocTNodeOff + sizeof(TNode) * (2 * kNoAllocMaxInstances);
static_assert(sizeof(LightF) * VIRA_GPU_MAX_LIGHTS <= 1024, "light table fits its slot");
static_assert(3 * 8 * VIRA_GPU_MAX_LIGHTS <= 256, "light positions fit their slot");
static_assert(kNoAllocInstOff % 16 == 0 && kNoAllocInstMOff % 16 == 0 && kNoAllocExtOff % 16 == 0,
              "aligned tables");

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

// Validation shared by both render entry points; returns the light count.
int render_checks(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_render_opts& o) {
    VG_CHECK(c.scene_mesh >= 0, VIRA_GPU_ERR_STATE, "render: no TLAS built");
    MeshG& m = c.mesh(c.scene_mesh);
    VG_CHECK(m.blas.built, VIRA_GPU_ERR_STATE, "render: BLAS not built");
    const int ninst = c.instanced ? (int)c.tlas.size() : 0;
    for (int i = 0; i < ninst; ++i) {
        const MeshG& mi = c.mesh(c.tlas[i].mesh_id);
        VG_CHECK(mi.blas.built && mi.blas.width == m.blas.width, VIRA_GPU_ERR_STATE,
                 "render: every instance needs a built BLAS of one width");
    }
    VG_CHECK(cam.width > 0 && cam.height > 0, VIRA_GPU_ERR_INVALID, "render: empty image");
    VG_CHECK(o.samples > 0 && o.bounces >= 0 && o.bounces < 64, VIRA_GPU_ERR_INVALID, "render: bad samples/bounces");
    VG_CHECK(o.integrator >= 0 && o.integrator <= 2, VIRA_GPU_ERR_INVALID, "render: unknown integrator");
    const size_t npix = (size_t)cam.width * cam.height;
    VG_CHECK(npix < (1ull << 30), VIRA_GPU_ERR_INVALID, "render: image too large");
    return (int)c.lights.size();
}

// The fused (per-pixel) render into scratch S: one wave per 8x8 tile.  The
// caller has checked the request (render_checks) and that the fused schedule
// applies.  Timing through ev (two events owned by the caller).
void run_fused(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_render_opts& o, const vira_gpu_outputs& out,
               vira_gpu_render_stats* st, const FusedScratch& S, hipEvent_t ev0, hipEvent_t ev1) {
    MeshG& m = c.mesh(c.scene_mesh);
    const bool inst = c.instanced;
    const int ninst = inst ? (int)c.tlas.size() : 0;
    const int nl = (int)c.lights.size();
    const uint32_t tiles_x = (uint32_t)((cam.width + 7) / 8);
    const uint32_t tiles = tiles_x * (uint32_t)((cam.height + 7) / 8);
    CamF cf = cam_params(cam);
    {
        LightF hl[VIRA_GPU_MAX_LIGHTS];
        for (int l = 0; l < nl; ++l) hl[l] = light_params(c.lights[l]);
        if (nl > 0)
            VG_HIP(hipMemcpyAsync(S.lights, hl, sizeof(LightF) * nl, hipMemcpyHostToDevice, c.stream));
    }
    const Lights lf{nl, S.lights};
    const MatF mf = mat_params(m);
    // the quantised nodes when every BLAS of the scene has them
    bool wq = m.blas.width == 4 && m.blas.compressed;
    for (int i = 0; i < ninst && wq; ++i) wq = c.mesh(c.tlas[i].mesh_id).blas.compressed;
    if (inst) {
        // instance table (pointers, materials) and FP64 inputs; k_rebase fills the FP32 placement.
        // The allocation-free call stages blocks of kStage instances on the stack; others take the
        // whole table in one host buffer.
        constexpr int kStage = 16;
        InstG hi_stack[kStage];
        double hm_stack[16 * kStage];
        std::vector<InstG> hi_vec;
        std::vector<double> hm_vec;
        const int stage = S.host_staging ? std::max(ninst, 1) : kStage;
        if (S.host_staging) {
            hi_vec.resize(stage);
            hm_vec.resize(16 * (size_t)stage);
        }
        InstG* hi = S.host_staging ? hi_vec.data() : hi_stack;
        double* hm = S.host_staging ? hm_vec.data() : hm_stack;
        for (int i0 = 0; i0 < ninst; i0 += stage) {
            const int n = std::min(stage, ninst - i0);
            for (int i = 0; i < n; ++i) {
                const MeshG& mi = c.mesh(c.tlas[i0 + i].mesh_id);
                hi[i] = InstG{};
                hi[i].nodes = wq ? mi.blas.qnodes.p : mi.blas.nodes.p;
                hi[i].tris = mi.blas.tris.as<float4>();
                hi[i].verts = mi.verts.as<GVertex>();
                hi[i].idx = mi.idx.as<uint32_t>();
                hi[i].mat = mat_params(mi);
                hi[i].mesh_id = (uint32_t)c.tlas[i0 + i].mesh_id;
                hi[i].material_id = mi.mat.material_id;
                for (int k = 0; k < 3; ++k) {
                    hi[i].blo[k] = mi.blas.lo[k];
                    hi[i].bhi[k] = mi.blas.hi[k];
                }
                for (int k = 0; k < 16; ++k) hm[16 * i + k] = c.tlas[i0 + i].transform[k];
            }
            VG_HIP(hipMemcpyAsync(S.inst + i0, hi, sizeof(InstG) * n, hipMemcpyHostToDevice, c.stream));
            VG_HIP(hipMemcpyAsync(S.inst_m + 16 * (size_t)i0, hm, 16 * 8 * n, hipMemcpyHostToDevice, c.stream));
            // pageable sources: wait before the staging arrays are refilled
            if (i0 + stage < ninst) VG_HIP(hipStreamSynchronize(c.stream));
        }
        double hl[3 * VIRA_GPU_MAX_LIGHTS];
        for (int l = 0; l < nl; ++l)
            for (int k = 0; k < 3; ++k) hl[3 * l + k] = c.lights[l].position[k];
        if (nl > 0) VG_HIP(hipMemcpyAsync(S.light_pos, hl, 3 * 8 * nl, hipMemcpyHostToDevice, c.stream));
        k_rebase<<<blocks_for(std::max(ninst, nl), 64), 64, 0, c.stream>>>(
            S.inst_m, ninst, S.light_pos, nl, cam.position[0], cam.position[1], cam.position[2], S.inst, S.lights);
        VG_HIP(hipGetLastError());
        k_tlas_boxes<<<blocks_for(c.tlas_nodes, 64), 64, 0, c.stream>>>(c.tlas_topo.as<int2>(), c.tlas_nodes, S.inst,
                                                                        S.tnodes);
        VG_HIP(hipGetLastError());
        cf.ox = cf.oy = cf.oz = 0.0f;  // the camera is the origin of the render frame
    }
    {
        JobExt X{};
        X.out = out_params(out);
        X.env = env_params(c, o);
        X.adapt = adaptive_params(o);
        X.pass = pass_params(cam, c.scene_mesh, m);
        VG_HIP(hipMemcpyAsync(S.ext, &X, sizeof(X), hipMemcpyHostToDevice, c.stream));
    }
    k_zero_u64<<<1, kCounterWords, 0, c.stream>>>(S.counters, kCounterWords);
    VG_HIP(hipGetLastError());
    VG_HIP(hipEventRecord(ev0, c.stream));
    const bool w4 = m.blas.width == 4;
    TileJob J{};
    J.nodes = wq ? m.blas.qnodes.p : m.blas.nodes.p;
    J.tris = m.blas.tris.as<float4>();
    J.verts = m.verts.as<GVertex>();
    J.idx = m.idx.as<uint32_t>();
    J.mat = mf;
    J.lights = lf;
    J.cam = cf;
    J.seed = o.seed;
    J.spp = o.samples;
    J.max_bounces = o.bounces;
    J.rr = o.russian_roulette;
    J.rr_start = o.rr_start;
    J.tiles_x = tiles_x;
    J.inst = inst ? S.inst : nullptr;
    J.tlas = inst ? S.tnodes : nullptr;
    J.ntlas = inst ? c.tlas_nodes : 0;
    J.ext = S.ext;
    const bool multi = fused_multi(c);
    const bool feat = fused_feat(c, o);
    using K = void (*)(TileJob, float*, unsigned long long*);
#define VG_FUSED_PAIR(W_, I_, M_) {k_render_fused<W_, I_, M_, false>, k_render_fused<W_, I_, M_, true>}
    const K table[3][2][2][2] = {
        {{VG_FUSED_PAIR(2, false, false), VG_FUSED_PAIR(2, false, true)},
         {VG_FUSED_PAIR(2, true, false), VG_FUSED_PAIR(2, true, true)}},
        {{VG_FUSED_PAIR(4, false, false), VG_FUSED_PAIR(4, false, true)},
         {VG_FUSED_PAIR(4, true, false), VG_FUSED_PAIR(4, true, true)}},
        {{VG_FUSED_PAIR(kWidthQ, false, false), VG_FUSED_PAIR(kWidthQ, false, true)},
         {VG_FUSED_PAIR(kWidthQ, true, false), VG_FUSED_PAIR(kWidthQ, true, true)}}};
#undef VG_FUSED_PAIR
    const int wi = w4 ? (wq ? 2 : 1) : 0;
    K kern = table[wi][inst ? 1 : 0][multi ? 1 : 0][feat ? 1 : 0];
    if (o.count_visits) {
        VG_CHECK(!multi && !feat, VIRA_GPU_ERR_UNSUPPORTED,
                 "render: visit counting takes one light or none, no ambient light, background, adaptive "
                 "sampling or material maps");
        const K stats_table[3][2] = {{k_render_fused_stats<2, false>, k_render_fused_stats<2, true>},
                                     {k_render_fused_stats<4, false>, k_render_fused_stats<4, true>},
                                     {k_render_fused_stats<kWidthQ, false>, k_render_fused_stats<kWidthQ, true>}};
        kern = stats_table[wi][inst ? 1 : 0];
    }
    kern<<<tiles, kFusedBS, 0, c.stream>>>(J, S.acc, S.counters);
    VG_HIP(hipGetLastError());
    VG_HIP(hipEventRecord(ev1, c.stream));
    unsigned long long hc[kVisitBase + 2 * kVisitWords];
    VG_HIP(hipMemcpyAsync(hc, S.counters, sizeof(hc), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    if (st) {
        *st = vira_gpu_render_stats{};
        st->primary_rays = hc[0];
        st->shadow_rays = hc[1];
        st->bounce_rays = hc[2];
        st->closest_hits = hc[3];
        if (o.count_visits) {
            st->visits_counted = 1;
            vira_gpu_visit_counts* vc[2] = {&st->closest_visits, &st->shadow_visits};
            for (int k = 0; k < 2; ++k) {
                const unsigned long long* h = hc + kVisitBase + k * kVisitWords;
                vc[k]->nodes = h[0];
                vc[k]->leaves = h[1];
                vc[k]->triangles = h[2];
                vc[k]->stack_spills = h[3];
                vc[k]->max_stack = (uint32_t)h[4];
                vc[k]->tlas_nodes = h[5];
                vc[k]->wave_node_steps = h[6];
            }
        }
        float f = 0.f;
        VG_HIP(hipEventElapsedTime(&f, ev0, ev1));
        st->total_ms = f;
        st->fused_ms = f;
        st->fused_launches = 1;
        st->integrator = 1;
    }
}

}  // namespace

void noalloc_init(int device) {
    VG_CHECK(device >= 0 && device < kMaxDevices, VIRA_GPU_ERR_INVALID, "noalloc_init: device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    VG_CHECK(e == hipSuccess && ptr != nullptr, VIRA_GPU_ERR_HIP,
             std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc_ptr[device] = ptr;
}

size_t noalloc_max_tiles() { return kNoAllocMaxTiles; }

void render(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_render_opts& o, const vira_gpu_outputs& out,
            vira_gpu_render_stats* st) {
    const int nl = render_checks(c, cam, o);
    MeshG& m = c.mesh(c.scene_mesh);
    const bool inst = c.instanced;
    const int ninst = inst ? (int)c.tlas.size() : 0;
    const uint32_t tiles_x = (uint32_t)((cam.width + 7) / 8);
    const uint32_t tiles = tiles_x * (uint32_t)((cam.height + 7) / 8);
    const uint32_t nq = tiles * 64u;
    const bool fused = o.integrator != 2;
    VG_CHECK(!inst || fused, VIRA_GPU_ERR_UNSUPPORTED, "render: instanced scenes use the per-pixel integrator");
    VG_CHECK(fused || !o.count_visits, VIRA_GPU_ERR_UNSUPPORTED,
             "render: visit counting runs in the per-pixel integrator only");
    if (!fused) {
        VG_CHECK(!o.adaptive_sampling, VIRA_GPU_ERR_UNSUPPORTED,
                 "render: adaptive sampling runs in the per-pixel integrator only");
        VG_CHECK(!c.env_has_ambient, VIRA_GPU_ERR_UNSUPPORTED,
                 "render: ambient light runs in the per-pixel integrator only");
        VG_CHECK(!(o.show_background && c.env_bg), VIRA_GPU_ERR_UNSUPPORTED,
                 "render: a shown background runs in the per-pixel integrator only");
    }

    c.r_acc.alloc((size_t)A_FIELDS * nq * 4);
    c.r_counters.alloc(kCounterWords * 8);
    c.r_lights.alloc(sizeof(LightF) * VIRA_GPU_MAX_LIGHTS);
    if (fused) {
        if (inst) {
            c.r_inst.alloc(sizeof(InstG) * ninst);
            c.r_inst_m.alloc((size_t)16 * 8 * ninst);
            c.r_tnodes.alloc(sizeof(TNode) * (size_t)c.tlas_nodes);
            c.r_light_pos.alloc((size_t)3 * 8 * VIRA_GPU_MAX_LIGHTS);
        }
        c.r_ext.alloc(sizeof(JobExt));
        const FusedScratch S{c.r_acc.as<float>(), c.r_counters.as<unsigned long long>(), c.r_lights.as<LightF>(),
                             c.r_inst.as<InstG>(), c.r_inst_m.as<double>(), c.r_light_pos.as<double>(),
                             c.r_ext.as<JobExt>(), c.r_tnodes.as<TNode>(), true};
        EventPair ev;
        run_fused(c, cam, o, out, st, S, ev.a, ev.b);
        return;
    }

    // ---- wavefront schedule
    CamF cf = cam_params(cam);
    const Lights lf = upload_lights(c, c.r_lights);
    const MatF mf = mat_params(m);
    const OutPtrs op = out_params(out);
    float* acc = c.r_acc.as<float>();
    unsigned long long* cnt = c.r_counters.as<unsigned long long>();
    uint32_t* qcount = reinterpret_cast<uint32_t*>(c.r_counters.as<char>() + 64);  // uint32 queue lengths per bounce
    const int maxq = 64;
    k_zero_u64<<<1, kCounterWords, 0, c.stream>>>(cnt, kCounterWords);
    VG_HIP(hipGetLastError());
    const void* nodes = m.blas.nodes.p;
    const bool w4 = m.blas.width == 4;
    const float4* tris = m.blas.tris.as<float4>();

    // one event pair per launch, owned here and released on every path
    std::vector<std::unique_ptr<EventPair>> evs;
    std::vector<int> kinds;
    auto ev_begin = [&](int kind) {
        evs.push_back(std::make_unique<EventPair>());
        kinds.push_back(kind);
        VG_HIP(hipEventRecord(evs.back()->a, c.stream));
    };
    auto ev_end = [&]() { VG_HIP(hipEventRecord(evs.back()->b, c.stream)); };
    EventPair total;
    VG_HIP(hipEventRecord(total.a, c.stream));
    {
        const uint32_t nslots = nq * (uint32_t)std::max(1, nl);
        c.r_ray_o.alloc((size_t)nq * 16 * 2);
        c.r_ray_d.alloc((size_t)nq * 16 * 2);
        c.r_hit.alloc((size_t)nq * 16);
        c.r_path_rad.alloc((size_t)P_FIELDS * nq * 4);
        c.r_srays_o.alloc((size_t)nslots * 16);
        c.r_srays_d.alloc((size_t)nslots * 16);
        c.r_scontrib.alloc((size_t)nslots * NB * 4);
        c.r_svis.alloc((size_t)nslots * 16);
        float4* ro[2] = {c.r_ray_o.as<float4>(), c.r_ray_o.as<float4>() + nq};
        float4* rd[2] = {c.r_ray_d.as<float4>(), c.r_ray_d.as<float4>() + nq};
        float* path = c.r_path_rad.as<float>();
        k_acc_init<<<blocks_for(nq, 256), 256, 0, c.stream>>>(nq, acc);
        VG_HIP(hipGetLastError());
        for (int s = 0; s < o.samples; ++s) {
            VG_HIP(hipMemsetAsync(qcount, 0, maxq * 4, c.stream));
            k_camera<<<blocks_for(nq, 256), 256, 0, c.stream>>>(cf, o.seed, (uint32_t)s, nq, tiles_x, ro[0], rd[0],
                                                                path, &cnt[0]);
            VG_HIP(hipGetLastError());
            for (int b = 0; b <= o.bounces; ++b) {
                const int cur = b & 1, nxt = cur ^ 1;
                const uint32_t* nqp = b == 0 ? nullptr : &qcount[b];
                ev_begin(0);
                auto tq = w4 ? k_trace_queue<4> : k_trace_queue<2>;
                tq<<<blocks_for(nq, kTraceBS), kTraceBS, 0, c.stream>>>(nodes, tris, ro[cur], rd[cur], nqp, nq, 1u, 0,
                                                                        c.r_hit.as<float4>());
                VG_HIP(hipGetLastError());
                ev_end();
                ev_begin(2);
                k_shade<<<blocks_for(nq, 128), 128, 0, c.stream>>>(
                    ro[cur], c.r_hit.as<float4>(), nqp, nq, nq, tiles_x, cam.width, m.verts.as<GVertex>(),
                    m.idx.as<uint32_t>(), mf, lf, o.seed, (uint32_t)s, (uint32_t)b, o.bounces, o.russian_roulette,
                    o.rr_start, path, acc, c.r_srays_o.as<float4>(), c.r_srays_d.as<float4>(),
                    c.r_scontrib.as<float>(), nslots, ro[nxt], rd[nxt], &qcount[b + 1], cnt);
                VG_HIP(hipGetLastError());
                ev_end();
                if (nl > 0) {
                    // the slot array holds nl entries per queue entry of this bounce (unused ones are marked)
                    ev_begin(1);
                    tq<<<blocks_for(nslots, kTraceBS), kTraceBS, 0, c.stream>>>(
                        nodes, tris, c.r_srays_o.as<float4>(), c.r_srays_d.as<float4>(), nqp, nq,
                        (uint32_t)nl, 1, c.r_svis.as<float4>());
                    VG_HIP(hipGetLastError());
                    ev_end();
                }
                ev_begin(2);
                k_accum<<<blocks_for(nq, 256), 256, 0, c.stream>>>(ro[cur], c.r_hit.as<float4>(), nqp, nq, nq, nl,
                                                                   c.r_srays_o.as<float4>(), c.r_svis.as<float4>(),
                                                                   c.r_scontrib.as<float>(), nslots, (uint32_t)b, path);
                VG_HIP(hipGetLastError());
                ev_end();
            }
            k_resolve<<<blocks_for(nq, 256), 256, 0, c.stream>>>(nq, path, acc);
            VG_HIP(hipGetLastError());
        }
        k_finalize<<<blocks_for(nq, 256), 256, 0, c.stream>>>(nq, tiles_x, cam.width, cam.height, o.samples, acc, op,
                                                              pass_params(cam, c.scene_mesh, m));
        VG_HIP(hipGetLastError());
    }
    VG_HIP(hipEventRecord(total.b, c.stream));
    unsigned long long hc[4];
    VG_HIP(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    if (st) {
        *st = vira_gpu_render_stats{};
        st->primary_rays = hc[0];
        st->shadow_rays = hc[1];
        st->bounce_rays = hc[2];
        st->closest_hits = hc[3];
        float f = 0.f;
        VG_HIP(hipEventElapsedTime(&f, total.a, total.b));
        st->total_ms = f;
        st->integrator = 2;
        for (size_t i = 0; i < evs.size(); ++i) {
            float g = 0.f;
            VG_HIP(hipEventElapsedTime(&g, evs[i]->a, evs[i]->b));
            if (kinds[i] == 0) { st->trace_closest_ms += g; st->trace_closest_launches++; }
            else if (kinds[i] == 1) { st->trace_shadow_ms += g; st->trace_shadow_launches++; }
            else st->shade_ms += g;
        }
    }
}

void render_noalloc(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_render_opts& o, const vira_gpu_outputs& out,
                    vira_gpu_render_stats* st) {
    render_checks(c, cam, o);
    VG_CHECK(o.integrator != 2, VIRA_GPU_ERR_UNSUPPORTED,
             "render_noalloc: only the per-pixel (fused) integrator runs without allocation");
    const uint32_t tiles = (uint32_t)((cam.width + 7) / 8) * (uint32_t)((cam.height + 7) / 8);
    VG_CHECK(tiles <= kNoAllocMaxTiles, VIRA_GPU_ERR_INVALID,
             "render_noalloc: " + std::to_string(tiles) + " 8x8 tiles exceed the scratch maximum of " +
                 std::to_string(kNoAllocMaxTiles) + " (e.g. 512 x 512 pixels)");
    const int ninst = c.instanced ? (int)c.tlas.size() : 0;
    VG_CHECK(ninst <= kNoAllocMaxInstances, VIRA_GPU_ERR_INVALID, "render_noalloc: too many instances");
    VG_CHECK(c.device >= 0 && c.device < kMaxDevices, VIRA_GPU_ERR_STATE, "render_noalloc: device out of range");
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        base = static_cast<unsigned char*>(g_noalloc_ptr[c.device]);
    }
    VG_CHECK(base != nullptr, VIRA_GPU_ERR_STATE, "render_noalloc: scratch not set up on this device");
    const FusedScratch S{reinterpret_cast<float*>(base), reinterpret_cast<unsigned long long*>(base + kNoAllocCounterOff),
                         reinterpret_cast<LightF*>(base + kNoAllocLightOff), reinterpret_cast<InstG*>(base + kNoAllocInstOff),
                         reinterpret_cast<double*>(base + kNoAllocInstMOff),
                         reinterpret_cast<double*>(base + kNoAllocLightPosOff),
                         reinterpret_cast<JobExt*>(base + kNoAllocExtOff),
                         reinterpret_cast<TNode*>(base + kNoAllocTNodeOff), false};
    run_fused(c, cam, o, out, st, S, c.na_ev[0], c.na_ev[1]);
}

}  // namespace vg

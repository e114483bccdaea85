// render.hip -- wavefront path tracer (rows a4-a10) and the trace_rays hook.
//
// CPUPathTracer::render / unidirectional / simulatePath / processIntersection
// (cpu_path_tracer.ipp:29-511) reorganised as queues of rays:
//   camera rays (8x8 pixel tiles, one wave per tile)
//   -> closest-hit traversal -> shading (direct-light shadow rays into a
//   per-hit slot array, bounce rays compacted into the next queue with a wave
//   ballot + one atomic per wave) -> any-hit traversal -> ordered
//   accumulation of the lit contributions -> next bounce.
// One path per pixel per sample pass; passes run in sample order and are
// folded into per-pixel Kahan-compensated FP32 sums, so the image does not
// depend on scheduling.  Every shading formula follows the reference's
// operation order (see the CPU restatement used as the oracle in the tests).
#include <cmath>

#include "context.hpp"
#include "traverse.hpp"

namespace vg {
namespace {

constexpr int kTraceBS = 128;  // traversal block size (2 waves)
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr uint32_t kNoPixel = 0xffffffffu;
constexpr float kInf = __builtin_huge_valf();

struct LightF {
    int kind;
    float px, py, pz;
    float radius;
    float value[NB];
};
struct Lights {
    int n;
    LightF l[VIRA_GPU_MAX_LIGHTS];
};
struct CamF {
    int W, H;
    float kinv[6];
    float rot[9];
    float ox, oy, oz;
    float zdir;
};
struct MatF {
    float albedo[NB];
    const float* tex;
    int tw, th;
    int smooth;
};

// Per-pixel accumulator fields (SoA, field-major: acc[f * npix + pixel]).
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
// Per-path fields (one path per pixel within a sample pass).
enum : int { P_BETA = 0, P_RAD = 8, P_DIR = 16, P_FIELDS = 24 };

__device__ __forceinline__ void wave_add(unsigned long long* counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    if (mask && __lane_id() == (unsigned)(__ffsll((long long)mask) - 1))
        atomicAdd(counter, (unsigned long long)__popcll(mask));
}

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
template <bool ANY>
__global__ void __launch_bounds__(kTraceBS) k_trace_rays(const Node2* __restrict__ nodes, const float4* __restrict__ tris,
                                                         const float* __restrict__ rays, size_t n, float* t_out,
                                                         uint32_t* id_out, float* b1_out, float* b2_out) {
    __shared__ uint32_t s_ref[kStackLds * kTraceBS];
    __shared__ float s_dist[kStackLds * kTraceBS];
    size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float4 a = reinterpret_cast<const float4*>(rays)[2 * r];
    const float4 b = reinterpret_cast<const float4*>(rays)[2 * r + 1];
    HitRec h = traverse<ANY>(nodes, tris, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), b.w, s_ref + threadIdx.x,
                             s_dist + threadIdx.x, kTraceBS);
    if (ANY) {
        if (t_out) t_out[r] = h.id != 0xffffffffu ? 0.0f : kInf;
    } else if (t_out) {
        t_out[r] = h.t;
    }
    if (id_out) id_out[r] = h.id;
    if (b1_out) b1_out[r] = h.b1;
    if (b2_out) b2_out[r] = h.b2;
}

// ------------------------------------------------------------ camera rays (a4)
__global__ void k_camera(CamF cam, uint32_t seed, uint32_t s, uint32_t nq, float4* ray_o, float4* ray_d,
                         float* path, size_t npix, unsigned long long* cnt_primary) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint32_t tiles_x = (uint32_t)(cam.W + 7) >> 3;
    uint32_t tile = q >> 6, lane = q & 63u;
    uint32_t x = (tile % tiles_x) * 8u + (lane & 7u), y = (tile / tiles_x) * 8u + (lane >> 3);
    bool valid = x < (uint32_t)cam.W && y < (uint32_t)cam.H;
    wave_add(cnt_primary, valid);
    if (!valid) {
        ray_o[q] = make_float4(0.f, 0.f, 0.f, __uint_as_float(kNoPixel));
        return;
    }
    uint32_t pix = x + y * (uint32_t)cam.W;
    float X = (float)x, Y = (float)y;
    if (s > 0) {
        X = X + rng_u01(seed, pix, s, SLOT_JITTER_X);
        Y = Y + rng_u01(seed, pix, s, SLOT_JITTER_Y);
    }
    // pixelToDirectionHelper (camera.ipp:1738-1749) + localDirectionToGlobal + normalize
    float hx = (cam.kinv[0] * X + cam.kinv[2] * Y) + cam.kinv[4];
    float hy = (cam.kinv[1] * X + cam.kinv[3] * Y) + cam.kinv[5];
    f3 dl = mk3(hx, cam.zdir * hy, cam.zdir);
    f3 dg = mk3((cam.rot[0] * dl.x + cam.rot[1] * dl.y) + cam.rot[2] * dl.z,
                (cam.rot[3] * dl.x + cam.rot[4] * dl.y) + cam.rot[5] * dl.z,
                (cam.rot[6] * dl.x + cam.rot[7] * dl.y) + cam.rot[8] * dl.z);
    f3 d = normalize(dg);
    ray_o[q] = make_float4(cam.ox, cam.oy, cam.oz, __uint_as_float(pix));
    ray_d[q] = make_float4(d.x, d.y, d.z, kInf);
    for (int k = 0; k < NB; ++k) {
        path[(P_BETA + k) * npix + pix] = 1.0f;
        path[(P_RAD + k) * npix + pix] = 0.0f;
        path[(P_DIR + k) * npix + pix] = 0.0f;
    }
}

// ------------------------------------------------------------ closest-hit traversal (a5)
__global__ void __launch_bounds__(kTraceBS) k_trace_closest(const Node2* __restrict__ nodes,
                                                            const float4* __restrict__ tris,
                                                            const float4* __restrict__ ray_o,
                                                            const float4* __restrict__ ray_d,
                                                            const uint32_t* __restrict__ nq_ptr, uint32_t nq_fixed,
                                                            float4* __restrict__ hit) {
    __shared__ uint32_t s_ref[kStackLds * kTraceBS];
    __shared__ float s_dist[kStackLds * kTraceBS];
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nq = nq_ptr ? *nq_ptr : nq_fixed;
    if (q >= nq) return;
    float4 o = ray_o[q];
    if (__float_as_uint(o.w) == kNoPixel) return;
    float4 d = ray_d[q];
    HitRec h = traverse<false>(nodes, tris, mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z), d.w, s_ref + threadIdx.x,
                               s_dist + threadIdx.x, kTraceBS);
    hit[q] = make_float4(h.t, h.b1, h.b2, __uint_as_float(h.id));
}

// ------------------------------------------------------------ any-hit traversal (a6)
__global__ void __launch_bounds__(kTraceBS) k_trace_shadow(const Node2* __restrict__ nodes,
                                                           const float4* __restrict__ tris,
                                                           const float4* __restrict__ s_o,
                                                           const float4* __restrict__ s_d,
                                                           const uint32_t* __restrict__ nq_ptr, uint32_t nq_fixed,
                                                           int nl, uint32_t* __restrict__ vis) {
    __shared__ uint32_t s_ref[kStackLds * kTraceBS];
    __shared__ float s_dist[kStackLds * kTraceBS];
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nq = nq_ptr ? *nq_ptr : nq_fixed;
    if (e >= nq * (uint32_t)nl) return;
    float4 o = s_o[e];
    if (__float_as_uint(o.w) == kNoPixel) return;
    float4 d = s_d[e];
    HitRec h = traverse<true>(nodes, tris, mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z), d.w, s_ref + threadIdx.x,
                              s_dist + threadIdx.x, kTraceBS);
    vis[e] = h.id == 0xffffffffu ? 1u : 0u;
}

// ------------------------------------------------------------ shading (a6, a7)
// offsetIntersection (rendering/ray.ipp:107-139)
__device__ __forceinline__ f3 offset_point(f3 p, f3 n) {
    int32_t oi[3] = {(int32_t)(256.0f * n.x), (int32_t)(256.0f * n.y), (int32_t)(256.0f * n.z)};
    float pc[3] = {p.x, p.y, p.z}, nc[3] = {n.x, n.y, n.z}, r[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float pi = __int_as_float(__float_as_int(pc[k]) + ((pc[k] < 0.0f) ? -oi[k] : oi[k]));
        r[k] = (fabsf(pc[k]) < (1.0f / 32.0f)) ? pc[k] + (1.0f / 65536.0f) * nc[k] : pi;
    }
    return mk3(r[0], r[1], r[2]);
}

__device__ __forceinline__ float power_heuristic(float f, float g) { return (f * f) / (f * f + g * g); }

__device__ __forceinline__ float tex_bilinear(const MatF& m, float u, float v) {
    float rx = (float)m.tw, ry = (float)m.th;
    float px = fmodf(u * rx, rx), py = fmodf(v * ry, ry);
    if (px < 0.f) px += rx;
    if (py < 0.f) py += ry;
    int x0 = (int)floorf(px), y0 = (int)floorf(py);
    x0 = min(x0, m.tw - 1);
    y0 = min(y0, m.th - 1);
    int x1 = min(x0 + 1, m.tw - 1), y1 = min(y0 + 1, m.th - 1);
    float p00 = m.tex[x0 + (size_t)y0 * m.tw], p10 = m.tex[x1 + (size_t)y0 * m.tw];
    float p01 = m.tex[x0 + (size_t)y1 * m.tw], p11 = m.tex[x1 + (size_t)y1 * m.tw];
    float dx = px - (float)x0, dy = py - (float)y0;
    return ((((1.0f - dx) * (1.0f - dy)) * p00 + (dx * (1.0f - dy)) * p10) + ((1.0f - dx) * dy) * p01) +
           (dx * dy) * p11;
}

// Light::sample (point_light.ipp:18-38, sphere_light.ipp:31-125)
__device__ __forceinline__ void sample_light(const LightF& L, f3 p, uint32_t seed, uint32_t pix, uint32_t s,
                                             uint32_t dim0, f3& dir, float& distance, float& pdf, float rad[NB]) {
    f3 c = mk3(L.px, L.py, L.pz);
    f3 toL = sub(c, p);
    distance = length(toL);
    float dInv = __fdiv_rn(1.0f, distance);
    if (L.kind == 0) {
        float dSq = distance * distance;
        pdf = 1.0f;
        dir = mul(dInv, toL);
#pragma unroll
        for (int k = 0; k < NB; ++k) rad[k] = __fdiv_rn(L.value[k], dSq);
        return;
    }
    float radius = L.radius;
    if (distance <= radius) {
        pdf = __fdiv_rn(1.0f, (4.0f * kPi * radius * radius));
        dir = distance < 1e-3f ? mk3(1.f, 0.f, 0.f) : mul(dInv, toL);
        distance = radius;
#pragma unroll
        for (int k = 0; k < NB; ++k) rad[k] = L.value[k];
        return;
    }
    float r1 = rng_u01(seed, pix, s, dim0), r2 = rng_u01(seed, pix, s, dim0 + 1);
    f3 W = mul(dInv, toL);
    f3 N = (W.x < W.y) ? mk3(1.f, 0.f, 0.f) : mk3(0.f, 1.f, 0.f);
    f3 Vv = normalize(cross(W, N));
    f3 U = normalize(cross(W, Vv));
    float phi = 2.0f * kPi * r1;
    float zeta = r2;
    float cc = __fdiv_rn(radius, distance);
    float theta = acosf((1.0f - zeta) + zeta * __builtin_sqrtf(1.0f - cc * cc));
    float sinT = sinf(theta), sinP = sinf(phi), cosT = cosf(theta), cosP = cosf(phi);
    float a = cosP * sinT, bb = cosT, cz = sinP * sinT;
    dir = mk3((a * U.x + bb * W.x) + cz * Vv.x, (a * U.y + bb * W.y) + cz * Vv.y, (a * U.z + bb * W.z) + cz * Vv.z);
    float r2s = radius * radius;
    float Omega = 2.0f * kPi * (1.0f - __fdiv_rn(__builtin_sqrtf(distance * distance - r2s), distance));
    pdf = __fdiv_rn(1.0f, Omega);
#pragma unroll
    for (int k = 0; k < NB; ++k) rad[k] = L.value[k];
}

// Light::getPDF (point_light.ipp:40-57, sphere_light.ipp:127-153)
__device__ __forceinline__ float light_pdf(const LightF& L, f3 p, f3 dir) {
    f3 c = mk3(L.px, L.py, L.pz);
    if (L.kind == 0) {
        f3 toL = normalize(sub(c, p));
        f3 q = normalize(dir);
        float d = dot(toL, q);
        return (d > 1.0f - 1e-6f) ? 1.0f : 0.0f;
    }
    f3 toC = sub(c, p);
    float dc = length(toC);
    float radius = L.radius;
    if (dc <= radius) return __fdiv_rn(1.0f, (4.0f * kPi * radius * radius));
    f3 tn = mul(toC, __fdiv_rn(1.0f, dc));
    float cosMax = __builtin_sqrtf(1.0f - __fdiv_rn(radius * radius, dc * dc));
    float ca = dot(dir, tn);
    if (ca < cosMax) return 0.0f;
    float Omega = 2.0f * kPi * (1.0f - __fdiv_rn(__builtin_sqrtf(dc * dc - radius * radius), dc));
    return __fdiv_rn(1.0f, Omega);
}

__device__ __forceinline__ f3 vp(const GVertex& v) { return mk3(v.px, v.py, v.pz); }
__device__ __forceinline__ f3 vn(const GVertex& v) { return mk3(v.nx, v.ny, v.nz); }

// processIntersection (cpu_path_tracer.ipp:279-469) for one queue entry.
__global__ void __launch_bounds__(128) k_shade(
    const float4* __restrict__ ray_o, const float4* __restrict__ hit, const uint32_t* __restrict__ nq_ptr,
    uint32_t nq_fixed, const GVertex* __restrict__ verts, const uint32_t* __restrict__ idx, MatF mat, Lights lights,
    uint32_t seed, uint32_t s, uint32_t bounce, int max_bounces, int rr, int rr_start, float* __restrict__ path,
    float* __restrict__ acc, size_t npix, float4* __restrict__ s_o, float4* __restrict__ s_d,
    float* __restrict__ s_contrib, uint32_t nslots, float4* __restrict__ next_o, float4* __restrict__ next_d,
    uint32_t* __restrict__ next_count, unsigned long long* __restrict__ counters) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nq = nq_ptr ? *nq_ptr : nq_fixed;
    if (q >= nq) return;
    const int nl = lights.n;
    float4 o4 = ray_o[q];
    uint32_t pix = __float_as_uint(o4.w);
    float4 h4 = pix == kNoPixel ? make_float4(kInf, 0.f, 0.f, __uint_as_float(0xffffffffu)) : hit[q];
    uint32_t id = __float_as_uint(h4.w);
    bool alive = pix != kNoPixel && id != 0xffffffffu;
    // shadow slots of this entry start invalid
    for (int l = 0; l < nl; ++l) s_o[(size_t)q * nl + l].w = __uint_as_float(kNoPixel);
    bool emit_bounce = false;
    f3 org = mk3(0.f, 0.f, 0.f), nd = mk3(0.f, 0.f, 0.f);
    bool any_shadow = false;
    if (alive) {
        const uint32_t i0 = idx[3 * (size_t)id], i1 = idx[3 * (size_t)id + 1], i2 = idx[3 * (size_t)id + 2];
        const GVertex V0 = verts[i0], V1 = verts[i1], V2 = verts[i2];
        const f3 p0 = vp(V0), p1 = vp(V1), p2 = vp(V2);
        const float w0 = h4.y, w1 = h4.z, w2 = (1.0f - h4.y) - h4.z;
        const float uvu = (w0 * V1.u + w1 * V2.u) + w2 * V0.u;
        const float uvv = (w0 * V1.v + w1 * V2.v) + w2 * V0.v;
        const f3 fn = normalize(cross(sub(p1, p0), sub(p2, p0)));
        const f3 n0 = vn(V0), n1 = vn(V1), n2 = vn(V2);
        f3 Nl = fn;
        if (mat.smooth) Nl = normalize(add(add(mul(w0, n1), mul(w1, n2)), mul(w2, n0)));
        const f3 N = Nl;
        const f3 arb = (fabsf(N.x) < 0.9f) ? mk3(1.f, 0.f, 0.f) : mk3(0.f, 1.f, 0.f);
        const f3 tg = normalize(cross(arb, N));
        const f3 bt = cross(N, tg);
        float texv = 1.0f;
        if (mat.tex) texv = tex_bilinear(mat, uvu, uvv);
        float alb[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            float av = (w0 * V1.alb[k] + w1 * V2.alb[k]) + w2 * V0.alb[k];
            float mv = mat.tex ? texv * mat.albedo[k] : mat.albedo[k];
            alb[k] = av * mv;
        }
        f3 P = add(add(mul(w0, p1), mul(w1, p2)), mul(w2, p0));
        {  // computeShadingPoint (cpu_path_tracer.ipp:481-511)
            f3 q0 = sub(P, mul(dot(sub(P, p1), n1), n1));
            f3 q1 = sub(P, mul(dot(sub(P, p2), n2), n2));
            f3 q2 = sub(P, mul(dot(sub(P, p0), n0), n0));
            f3 sp = add(add(mul(q0, w0), mul(q1, w1)), mul(q2, w2));
            f3 diff = sub(sp, P);
            if (dot(diff, Nl) > 0.0f) P = sp;
        }
        if (bounce == 0) {
            if (__float_as_uint(acc[A_TRI * npix + pix]) == 0xffffffffu) acc[A_TRI * npix + pix] = __uint_as_float(id);
#pragma unroll
            for (int k = 0; k < NB; ++k) acc[(A_ALB + k) * npix + pix] = acc[(A_ALB + k) * npix + pix] + alb[k];
            acc[(A_NRM + 0) * npix + pix] = acc[(A_NRM + 0) * npix + pix] + N.x;
            acc[(A_NRM + 1) * npix + pix] = acc[(A_NRM + 1) * npix + pix] + N.y;
            acc[(A_NRM + 2) * npix + pix] = acc[(A_NRM + 2) * npix + pix] + N.z;
            acc[A_DEPTH * npix + pix] = acc[A_DEPTH * npix + pix] + h4.x;
            acc[A_COUNT * npix + pix] = acc[A_COUNT * npix + pix] + 1.0f;
        }
        org = offset_point(P, fn);
        float beta[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) beta[k] = path[(P_BETA + k) * npix + pix];
        // direct lighting: one shadow ray per light, contribution staged in the slot
        for (int l = 0; l < nl; ++l) {
            f3 ld;
            float dist, lpdf, lrad[NB];
            sample_light(lights.l[l], org, seed, pix, s, (bounce << 8) | (SLOT_LIGHT0 + 2 * l), ld, dist, lpdf, lrad);
            if (!(lpdf > 0.0f)) continue;
            any_shadow = true;
            float ndl = dot(N, ld);
            float mx = fmaxf(ndl, 0.0f);
            float cth = fmaxf(0.0f, dot(ld, N));
            float mpdf = (ndl <= 0.0f) ? 0.0f : ndl * kInvPi;
            float wgt = power_heuristic(lpdf, mpdf);
            size_t e = (size_t)q * nl + l;
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                float bsdf = (alb[k] * kInvPi) * mx;
                s_contrib[(size_t)k * nslots + e] = __fdiv_rn((((beta[k] * bsdf) * lrad[k]) * cth) * wgt, lpdf);
            }
            s_o[e] = make_float4(org.x, org.y, org.z, __uint_as_float(pix));
            s_d[e] = make_float4(ld.x, ld.y, ld.z, dist);
        }
        // BSDF sampling of the next bounce (material.ipp:8-33, sampling.ipp:53-84)
        if ((int)bounce < max_bounces) {
            uint32_t db = bounce << 8;
            float u1 = rng_u01(seed, pix, s, db | SLOT_BSDF_U);
            float u2 = rng_u01(seed, pix, s, db | SLOT_BSDF_V);
            float ox = 2.0f * u1 - 1.0f, oy = 2.0f * u2 - 1.0f;
            float dx = 0.f, dy = 0.f;
            if (!(fabsf(ox) < 1e-6f && fabsf(oy) < 1e-6f)) {
                float r, th;
                if (fabsf(ox) > fabsf(oy)) {
                    r = ox;
                    th = (kPi / 4.0f) * __fdiv_rn(oy, ox);
                } else {
                    r = oy;
                    th = (kPi / 2.0f) - (kPi / 4.0f) * __fdiv_rn(ox, oy);
                }
                dx = r * cosf(th);
                dy = r * sinf(th);
            }
            float lz = __builtin_sqrtf(fmaxf(0.0f, (1.0f - dx * dx) - dy * dy));
            float mpdf = lz * kInvPi;
            if (mpdf > 0.0f) {
                nd = mk3((tg.x * dx + bt.x * dy) + N.x * lz, (tg.y * dx + bt.y * dy) + N.y * lz,
                         (tg.z * dx + bt.z * dy) + N.z * lz);
                float ndd = dot(N, nd);
                float mx = fmaxf(ndd, 0.0f);
                float cth = fmaxf(0.0f, dot(nd, N));
                float lp = 0.0f;
                for (int l = 0; l < nl; ++l) lp = lp + light_pdf(lights.l[l], org, nd);
                if (nl > 0) lp = __fdiv_rn(lp, (float)nl);
                float wgt = power_heuristic(mpdf, lp);
                float qmax = 0.0f;
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    float bsdf = (alb[k] * kInvPi) * mx;
                    beta[k] = beta[k] * __fdiv_rn((bsdf * cth) * wgt, mpdf);
                    qmax = fmaxf(qmax, beta[k]);
                }
                emit_bounce = true;
                if (rr && (int)bounce + 1 >= rr_start) {
                    float qq = fminf(1.0f, qmax);
                    float ur = rng_u01(seed, pix, s, db | SLOT_RR);
                    if (ur >= qq) {
                        emit_bounce = false;
                    } else {
#pragma unroll
                        for (int k = 0; k < NB; ++k) beta[k] = __fdiv_rn(beta[k], qq);
                    }
                }
                if (emit_bounce) {
#pragma unroll
                    for (int k = 0; k < NB; ++k) path[(P_BETA + k) * npix + pix] = beta[k];
                }
            }
        }
    }
    uint32_t slot = wave_append(next_count, emit_bounce);
    if (emit_bounce) {
        next_o[slot] = make_float4(org.x, org.y, org.z, __uint_as_float(pix));
        next_d[slot] = make_float4(nd.x, nd.y, nd.z, kInf);
    }
    wave_add(&counters[2], emit_bounce);
    (void)any_shadow;
}

// Ordered sum of the lit contributions of each hit (light order), as the
// `radiance +=` loop of cpu_path_tracer.ipp:399-428.
__global__ void k_accum(const float4* __restrict__ ray_o, const uint32_t* __restrict__ nq_ptr, uint32_t nq_fixed,
                        int nl, const float4* __restrict__ s_o, const uint32_t* __restrict__ vis,
                        const float* __restrict__ s_contrib, uint32_t nslots, uint32_t bounce, float* __restrict__ path,
                        size_t npix, unsigned long long* counters) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nq = nq_ptr ? *nq_ptr : nq_fixed;
    bool inq = q < nq;
    uint32_t nsh = 0;
    if (inq) {
        uint32_t pix = __float_as_uint(ray_o[q].w);
        if (pix != kNoPixel) {
            float vrad[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) vrad[k] = 0.0f;
            bool any = false;
            for (int l = 0; l < nl; ++l) {
                size_t e = (size_t)q * nl + l;
                if (__float_as_uint(s_o[e].w) == kNoPixel) continue;
                nsh++;
                if (!vis[e]) continue;
                any = true;
#pragma unroll
                for (int k = 0; k < NB; ++k) vrad[k] = vrad[k] + s_contrib[(size_t)k * nslots + e];
            }
            // a hit adds its (possibly zero) vertex radiance, as pathRadiance += radiance
            (void)any;
#pragma unroll
            for (int k = 0; k < NB; ++k) path[(P_RAD + k) * npix + pix] = path[(P_RAD + k) * npix + pix] + vrad[k];
            if (bounce == 0) {
#pragma unroll
                for (int k = 0; k < NB; ++k) path[(P_DIR + k) * npix + pix] = path[(P_DIR + k) * npix + pix] + vrad[k];
            }
        }
    }
    if (inq && nsh) atomicAdd(&counters[1], (unsigned long long)nsh);
}

// Kahan-compensated fold of one sample pass into the pixel sums.
__global__ void k_resolve(size_t npix, const float* __restrict__ path, float* __restrict__ acc) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        {
            float x = path[(P_RAD + k) * npix + p];
            float s = acc[(A_RAD_S + k) * npix + p], c = acc[(A_RAD_C + k) * npix + p];
            float y = x - c;
            float t = s + y;
            c = (t - s) - y;
            acc[(A_RAD_S + k) * npix + p] = t;
            acc[(A_RAD_C + k) * npix + p] = c;
        }
        {
            float x = path[(P_DIR + k) * npix + p];
            float s = acc[(A_DIR_S + k) * npix + p], c = acc[(A_DIR_C + k) * npix + p];
            float y = x - c;
            float t = s + y;
            c = (t - s) - y;
            acc[(A_DIR_S + k) * npix + p] = t;
            acc[(A_DIR_C + k) * npix + p] = c;
        }
    }
}

__global__ void k_acc_init(size_t npix, float* acc) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    for (int f = 0; f < A_FIELDS; ++f) acc[(size_t)f * npix + p] = f == A_TRI ? __uint_as_float(0xffffffffu) : 0.0f;
}

// Averages and passes (cpu_path_tracer.ipp:229-239, passes.ipp:81-127).
__global__ void k_finalize(size_t npix, int spp, const float* __restrict__ acc, float* radiance, float* direct,
                           float* indirect, float* albedo, float* normal, float* depth, float* alpha, uint32_t* tri) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    float sf = (float)spp;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        float rad = __fdiv_rn(acc[(A_RAD_S + k) * npix + p], sf);
        float dr = __fdiv_rn(acc[(A_DIR_S + k) * npix + p], sf);
        if (radiance) radiance[NB * p + k] = rad;
        if (direct) direct[NB * p + k] = dr;
        if (indirect) indirect[NB * p + k] = rad - dr;
        if (albedo) albedo[NB * p + k] = __fdiv_rn(acc[(A_ALB + k) * npix + p], sf);
    }
    float cnt = acc[A_COUNT * npix + p];
    if (normal) {
        f3 n = mk3(acc[A_NRM * npix + p], acc[(A_NRM + 1) * npix + p], acc[(A_NRM + 2) * npix + p]);
        if (cnt > 0.0f) n = normalize(n);
        normal[3 * p] = n.x;
        normal[3 * p + 1] = n.y;
        normal[3 * p + 2] = n.z;
    }
    if (depth) depth[p] = cnt > 0.0f ? __fdiv_rn(acc[A_DEPTH * npix + p], cnt) : kInf;
    if (alpha) alpha[p] = __fdiv_rn(cnt, sf);
    if (tri) tri[p] = __float_as_uint(acc[A_TRI * npix + p]);
}

struct EvPair {
    hipEvent_t a, b;
    int kind;  // 0 closest, 1 shadow, 2 shade
};

}  // namespace

void trace_rays(Ctx& c, MeshG& m, const float* rays, size_t n, int any_hit, float* t, uint32_t* id, float* b1,
                float* b2, double* ms) {
    VG_CHECK(m.blas.built, VIRA_GPU_ERR_STATE, "trace_rays: BLAS not built");
    if (n == 0) {
        if (ms) *ms = 0;
        return;
    }
    hipEvent_t e0, e1;
    VG_HIP(hipEventCreate(&e0));
    VG_HIP(hipEventCreate(&e1));
    VG_HIP(hipEventRecord(e0, c.stream));
    if (any_hit)
        k_trace_rays<true><<<blocks_for(n, kTraceBS), kTraceBS, 0, c.stream>>>(
            m.blas.nodes.as<Node2>(), m.blas.tris.as<float4>(), rays, n, t, id, b1, b2);
    else
        k_trace_rays<false><<<blocks_for(n, kTraceBS), kTraceBS, 0, c.stream>>>(
            m.blas.nodes.as<Node2>(), m.blas.tris.as<float4>(), rays, n, t, id, b1, b2);
    VG_HIP(hipGetLastError());
    VG_HIP(hipEventRecord(e1, c.stream));
    VG_HIP(hipEventSynchronize(e1));
    float f = 0.f;
    VG_HIP(hipEventElapsedTime(&f, e0, e1));
    if (ms) *ms = f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

void render(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_render_opts& o, const vira_gpu_outputs& out,
            vira_gpu_render_stats* st) {
    VG_CHECK(c.scene_mesh >= 0, VIRA_GPU_ERR_STATE, "render: no TLAS built");
    MeshG& m = c.mesh(c.scene_mesh);
    VG_CHECK(m.blas.built, VIRA_GPU_ERR_STATE, "render: BLAS not built");
    VG_CHECK(cam.width > 0 && cam.height > 0, VIRA_GPU_ERR_INVALID, "render: empty image");
    VG_CHECK(o.samples > 0 && o.bounces >= 0 && o.bounces < 64, VIRA_GPU_ERR_INVALID, "render: bad samples/bounces");
    const size_t npix = (size_t)cam.width * cam.height;
    VG_CHECK(npix < (1ull << 31), VIRA_GPU_ERR_INVALID, "render: image too large");
    const uint32_t tiles = (uint32_t)(((cam.width + 7) / 8) * ((cam.height + 7) / 8));
    const uint32_t nq = tiles * 64u;
    const int nl = (int)c.lights.size();
    const uint32_t nslots = nq * (uint32_t)std::max(1, nl);

    CamF cf;
    cf.W = cam.width;
    cf.H = cam.height;
    for (int k = 0; k < 6; ++k) cf.kinv[k] = cam.kinv[k];
    for (int k = 0; k < 9; ++k) cf.rot[k] = cam.rot[k];
    cf.ox = (float)cam.position[0];
    cf.oy = (float)cam.position[1];
    cf.oz = (float)cam.position[2];
    cf.zdir = cam.z_dir;
    Lights lf{};
    lf.n = nl;
    for (int l = 0; l < nl; ++l) {
        const vira_gpu_light& L = c.lights[l];
        lf.l[l].kind = L.kind;
        lf.l[l].px = (float)L.position[0];
        lf.l[l].py = (float)L.position[1];
        lf.l[l].pz = (float)L.position[2];
        lf.l[l].radius = L.radius;
        for (int k = 0; k < NB; ++k) lf.l[l].value[k] = L.value[k];
    }
    MatF mf;
    for (int k = 0; k < NB; ++k) mf.albedo[k] = m.mat.albedo[k];
    mf.tex = m.mat.texture_dev;
    mf.tw = m.mat.tex_width;
    mf.th = m.mat.tex_height;
    mf.smooth = m.mat.smooth_shading;

    c.r_ray_o.alloc((size_t)nq * 16 * 2);
    c.r_ray_d.alloc((size_t)nq * 16 * 2);
    c.r_hit.alloc((size_t)nq * 16);
    c.r_path_rad.alloc((size_t)P_FIELDS * npix * 4);
    c.r_acc.alloc((size_t)A_FIELDS * npix * 4);
    c.r_srays_o.alloc((size_t)nslots * 16);
    c.r_srays_d.alloc((size_t)nslots * 16);
    c.r_scontrib.alloc((size_t)nslots * NB * 4);
    c.r_svis.alloc((size_t)nslots * 4);
    c.r_counters.alloc(64 * 8);
    float4* ro[2] = {c.r_ray_o.as<float4>(), c.r_ray_o.as<float4>() + nq};
    float4* rd[2] = {c.r_ray_d.as<float4>(), c.r_ray_d.as<float4>() + nq};
    float* path = c.r_path_rad.as<float>();
    float* acc = c.r_acc.as<float>();
    unsigned long long* cnt = c.r_counters.as<unsigned long long>();
    // counters: [0] primary [1] shadow [2] bounce ; uint32 queue counts at byte 64 + 4*b
    uint32_t* qcount = reinterpret_cast<uint32_t*>(c.r_counters.as<char>() + 64);
    const int maxq = 64;
    VG_HIP(hipMemsetAsync(c.r_counters.p, 0, 64 * 8, c.stream));

    std::vector<EvPair> evs;
    auto ev_begin = [&](int kind) {
        EvPair p;
        VG_HIP(hipEventCreate(&p.a));
        VG_HIP(hipEventCreate(&p.b));
        p.kind = kind;
        VG_HIP(hipEventRecord(p.a, c.stream));
        evs.push_back(p);
    };
    auto ev_end = [&]() { VG_HIP(hipEventRecord(evs.back().b, c.stream)); };
    hipEvent_t t0, t1;
    VG_HIP(hipEventCreate(&t0));
    VG_HIP(hipEventCreate(&t1));
    VG_HIP(hipEventRecord(t0, c.stream));

    k_acc_init<<<blocks_for(npix, 256), 256, 0, c.stream>>>(npix, acc);
    VG_HIP(hipGetLastError());
    const Node2* nodes = m.blas.nodes.as<Node2>();
    const float4* tris = m.blas.tris.as<float4>();
    for (int s = 0; s < o.samples; ++s) {
        VG_HIP(hipMemsetAsync(qcount, 0, maxq * 4, c.stream));
        k_camera<<<blocks_for(nq, 256), 256, 0, c.stream>>>(cf, o.seed, (uint32_t)s, nq, ro[0], rd[0], path, npix, &cnt[0]);
        VG_HIP(hipGetLastError());
        for (int b = 0; b <= o.bounces; ++b) {
            const int cur = b & 1, nxt = cur ^ 1;
            const uint32_t* nqp = b == 0 ? nullptr : &qcount[b];
            ev_begin(0);
            k_trace_closest<<<blocks_for(nq, kTraceBS), kTraceBS, 0, c.stream>>>(nodes, tris, ro[cur], rd[cur], nqp, nq,
                                                                                 c.r_hit.as<float4>());
            VG_HIP(hipGetLastError());
            ev_end();
            ev_begin(2);
            k_shade<<<blocks_for(nq, 128), 128, 0, c.stream>>>(
                ro[cur], c.r_hit.as<float4>(), nqp, nq, m.verts.as<GVertex>(), m.idx.as<uint32_t>(), mf, lf, o.seed,
                (uint32_t)s, (uint32_t)b, o.bounces, o.russian_roulette, o.rr_start, path, acc, npix,
                c.r_srays_o.as<float4>(), c.r_srays_d.as<float4>(), c.r_scontrib.as<float>(), nslots, ro[nxt],
                rd[nxt], &qcount[b + 1], cnt);
            VG_HIP(hipGetLastError());
            ev_end();
            if (nl > 0) {
                ev_begin(1);
                k_trace_shadow<<<blocks_for(nslots, kTraceBS), kTraceBS, 0, c.stream>>>(
                    nodes, tris, c.r_srays_o.as<float4>(), c.r_srays_d.as<float4>(), nqp, nq, nl, c.r_svis.as<uint32_t>());
                VG_HIP(hipGetLastError());
                ev_end();
                ev_begin(2);
                k_accum<<<blocks_for(nq, 256), 256, 0, c.stream>>>(ro[cur], nqp, nq, nl, c.r_srays_o.as<float4>(),
                                                                   c.r_svis.as<uint32_t>(), c.r_scontrib.as<float>(),
                                                                   nslots, (uint32_t)b, path, npix, cnt);
                VG_HIP(hipGetLastError());
                ev_end();
            }
        }
        k_resolve<<<blocks_for(npix, 256), 256, 0, c.stream>>>(npix, path, acc);
        VG_HIP(hipGetLastError());
    }
    k_finalize<<<blocks_for(npix, 256), 256, 0, c.stream>>>(npix, o.samples, acc, out.radiance_dev, out.direct_dev,
                                                             out.indirect_dev, out.albedo_dev, out.normal_dev,
                                                             out.depth_dev, out.alpha_dev, out.tri_id_dev);
    VG_HIP(hipGetLastError());
    VG_HIP(hipEventRecord(t1, c.stream));
    unsigned long long hc[3];
    VG_HIP(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    if (st) {
        *st = vira_gpu_render_stats{};
        st->primary_rays = hc[0];
        st->shadow_rays = hc[1];
        st->bounce_rays = hc[2];
        float f = 0.f;
        VG_HIP(hipEventElapsedTime(&f, t0, t1));
        st->total_ms = f;
        for (const EvPair& p : evs) {
            float g = 0.f;
            VG_HIP(hipEventElapsedTime(&g, p.a, p.b));
            if (p.kind == 0) { st->trace_closest_ms += g; st->trace_closest_launches++; }
            else if (p.kind == 1) { st->trace_shadow_ms += g; st->trace_shadow_launches++; }
            else st->shade_ms += g;
        }
    }
    for (const EvPair& p : evs) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
}

}  // namespace vg

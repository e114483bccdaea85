// shading.hpp -- surface interaction, light sampling and BSDF sampling.
//
// CPUPathTracer::processIntersection (cpu_path_tracer.ipp:279-469) split into
// three steps shared by both integrators (fused per-pixel and wavefront), so
// the two compute the same numbers:
//   surface_at()     interpolation, shading frame, albedo, shading point,
//                    integer-ULP offset of the new origin;
//   light_contrib()  Light::sample + the MIS-weighted Lambertian term of one
//                    light (the shadow ray decides whether it is added);
//   sample_bounce()  cosine-weighted direction, throughput update and the
//                    Russian-roulette decision.
// Operation order follows the reference expression by expression.
#pragma once

#include "device_math.hpp"
#include "fp_trig.hpp"
#include "../../include/vira_gpu.h"

namespace vg {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kInf = __builtin_huge_valf();

struct LightF {
    int kind;
    float px, py, pz;
    float radius;
    float value[NB];
};
// Lights live in device memory (a dynamically indexed by-value kernel argument
// would be copied into registers).
struct Lights {
    int n;
    const LightF* __restrict__ l;
};
struct CamF {
    int W, H;
    float kinv[6];
    float rot[9];
    float ox, oy, oz;
    float zdir;
};
struct MatF {
    const float4* fnrm; // the mesh's unit face normals (FP64-derived, FP32), per triangle
    float albedo[NB];
    const float* tex;   // albedo map: tch floats per texel
    int tw, th;
    int smooth;
    int tch;            // 1: grey texture x albedo[k]; 8: spectral albedo map (Image<TSpectral>)
    const float* nmap;  // tangent-space normal map, 3 floats per texel (Material::getNormal), or null
    int nw, nh;
};
// The scene's light environment: ambient light (Scene::getAmbient/hasAmbient)
// and the background image shown on misses (Scene::getBackgroundEmission).
struct EnvF {
    float ambient[NB];
    int has_ambient;
    int show_bg;        // CPUPathTracerOptions::show_background
    const float* bg;    // W*H*8 spectral image, or null (zero background)
    int bw, bh;
};

// Pixel tiles of 8x8: queue/accumulator index q = tile * 64 + (y & 7) * 8 + (x & 7).
__device__ __forceinline__ void tile_xy(uint32_t q, uint32_t tiles_x, uint32_t& x, uint32_t& y) {
    uint32_t tile = q >> 6, lane = q & 63u;
    x = (tile % tiles_x) * 8u + (lane & 7u);
    y = (tile / tiles_x) * 8u + (lane >> 3);
}
__device__ __forceinline__ uint32_t tile_index(uint32_t x, uint32_t y, uint32_t tiles_x) {
    return (((y >> 3) * tiles_x + (x >> 3)) << 6) + ((y & 7u) << 3) + (x & 7u);
}

// Camera::pixelToRay for sample s of pixel (x, y): sample 0 through the pixel
// corner, later samples jittered (cpu_path_tracer.ipp:178-192,
// camera.ipp:1126-1163,1738-1749).
__device__ __forceinline__ f3 camera_dir(const CamF& cam, uint32_t x, uint32_t y, uint32_t pix, uint32_t seed,
                                         uint32_t s) {
    float X = (float)x, Y = (float)y;
    if (s > 0) {
        X = X + rng_u01(seed, pix, s, SLOT_JITTER_X);
        Y = Y + rng_u01(seed, pix, s, SLOT_JITTER_Y);
    }
    float hx = (cam.kinv[0] * X + cam.kinv[2] * Y) + cam.kinv[4];
    float hy = (cam.kinv[1] * X + cam.kinv[3] * Y) + cam.kinv[5];
    f3 dl = mk3(hx, cam.zdir * hy, cam.zdir);
    f3 dg = mk3((cam.rot[0] * dl.x + cam.rot[1] * dl.y) + cam.rot[2] * dl.z,
                (cam.rot[3] * dl.x + cam.rot[4] * dl.y) + cam.rot[5] * dl.z,
                (cam.rot[6] * dl.x + cam.rot[7] * dl.y) + cam.rot[8] * dl.z);
    return normalize(dg);
}

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

// PowerHeuristic (cpu_path_tracer.ipp:472-478), one sample each
__device__ __forceinline__ float power_heuristic(float f, float g) { return (f * f) / (f * f + g * g); }

// Image::sampleUVs + interpolatePixel (image.ipp:589-670): uv * resolution
// modulo the resolution (negative wrapped up), the floor texel and its +1
// neighbours clamped to the last texel, bilinear weights.
struct Taps {
    size_t o00, o10, o01, o11;  // texel indices
    float w00, w10, w01, w11;
};
__device__ __forceinline__ Taps bilinear_taps(int w, int h, float u, float v) {
    const float rx = (float)w, ry = (float)h;
    float px = fmodf(u * rx, rx), py = fmodf(v * ry, ry);
    if (px < 0.f) px += rx;
    if (py < 0.f) py += ry;
    if (!(px >= 0.f)) px = 0.f;  // non-finite uv: keep the fetch in range
    if (!(py >= 0.f)) py = 0.f;
    int x0 = (int)floorf(px), y0 = (int)floorf(py);
    x0 = min(x0, w - 1);
    y0 = min(y0, h - 1);
    const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
    const float dx = px - (float)x0, dy = py - (float)y0;
    Taps T;
    T.o00 = x0 + (size_t)y0 * w;
    T.o10 = x1 + (size_t)y0 * w;
    T.o01 = x0 + (size_t)y1 * w;
    T.o11 = x1 + (size_t)y1 * w;
    T.w00 = (1.0f - dx) * (1.0f - dy);
    T.w10 = dx * (1.0f - dy);
    T.w01 = (1.0f - dx) * dy;
    T.w11 = dx * dy;
    return T;
}
// Channel c of a C-channel image at the taps.
__device__ __forceinline__ float taps_eval(const Taps& T, const float* __restrict__ img, int C, int c) {
    return ((T.w00 * img[T.o00 * C + c] + T.w10 * img[T.o10 * C + c]) + T.w01 * img[T.o01 * C + c]) +
           T.w11 * img[T.o11 * C + c];
}

// Background radiance along direction d (cpu_path_tracer.ipp:257-270):
// uv = (atan2(y, x) / 2 pi, atan2(sqrt(x^2 + y^2), z) / pi), sampleUVs.
__device__ __forceinline__ void background_radiance(const EnvF& E, f3 d, float bg[NB]) {
    if (!E.bg) {
#pragma unroll
        for (int k = 0; k < NB; ++k) bg[k] = 0.0f;
        return;
    }
    const float u = __fdiv_rn(vgm::atan2_f32(d.y, d.x), 2.0f * kPi);
    const float v = __fdiv_rn(vgm::atan2_f32(__builtin_sqrtf(d.x * d.x + d.y * d.y), d.z), kPi);
    const Taps T = bilinear_taps(E.bw, E.bh, u, v);
#pragma unroll
    for (int k = 0; k < NB; ++k) bg[k] = taps_eval(T, E.bg, NB, k);
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
    // cos(theta) = (1 - zeta) + zeta cos(theta_max), with 1 - cos(theta_max)
    // = 1 - sqrt(1 - (r/d)^2) written as x / (1 + sqrt(1 - x)), x = (r/d)^2,
    // which keeps its precision in FP32 for a sun at 1 AU (the reference's
    // 1 - sqrt(d^2 - r^2)/d loses about 0.3 % there)
    float cc = __fdiv_rn(radius, distance);
    float x = cc * cc;
    float q = __fdiv_rn(x, 1.0f + __builtin_sqrtf(1.0f - x));  // 1 - cos(theta_max)
    float t = zeta * q;                                          // 1 - cos(theta)
    float cosT = 1.0f - t, sinT = __builtin_sqrtf(t * (2.0f - t));
    float sinP, cosP;
    vgm::sincos_f32(phi, sinP, cosP);
    float a = cosP * sinT, bb = cosT, cz = sinP * sinT;
    dir = mk3((a * U.x + bb * W.x) + cz * Vv.x, (a * U.y + bb * W.y) + cz * Vv.y, (a * U.z + bb * W.z) + cz * Vv.z);
    float Omega = (2.0f * kPi) * q;  // 2 pi (1 - cos(theta_max))
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
    float cc = __fdiv_rn(radius, dc);
    float x = cc * cc;
    float q = __fdiv_rn(x, 1.0f + __builtin_sqrtf(1.0f - x));  // 1 - cos(theta_max), as in sample_light
    float cosMax = 1.0f - q;
    float ca = dot(dir, tn);
    if (ca < cosMax) return 0.0f;
    float Omega = (2.0f * kPi) * q;
    return __fdiv_rn(1.0f, Omega);
}

struct Surface {
    f3 org;         // offset shading point (origin of shadow and bounce rays)
    f3 N;           // shading normal (after the normal map)
    f3 Nf, tg, bt;  // tangent frame [tg bt Nf] of the normal before the map (BSDF sampling uses it)
    float alb[NB];  // vertex albedo x material albedo map
};

// Mesh-local part of a hit (cpu_path_tracer.ipp:284-349): shading normal,
// face normal, albedo and the shading point, all in the mesh's frame.
struct LocalHit {
    f3 P;    // computeShadingPoint of the barycentric point
    f3 Nl;   // shading normal (face normal, or interpolated vertex normals when smooth)
    f3 fn;   // unit face normal
    float u, v;  // interpolated uv
    float alb[NB];
};

// FEAT = false takes the grey-texture branch only (no spectral albedo map).
template <bool FEAT = true>
__device__ __forceinline__ void local_hit(const GVertex* __restrict__ verts, const uint32_t* __restrict__ idx,
                                          const MatF& mat, uint32_t id, float b1, float b2, LocalHit& L) {
    const uint32_t i0 = idx[3 * (size_t)id], i1 = idx[3 * (size_t)id + 1], i2 = idx[3 * (size_t)id + 2];
    const GVertex V0 = verts[i0], V1 = verts[i1], V2 = verts[i2];
    const f3 p0 = mk3(V0.px, V0.py, V0.pz), p1 = mk3(V1.px, V1.py, V1.pz), p2 = mk3(V2.px, V2.py, V2.pz);
    // Embree barycentrics: w = (u, v, 1-u-v) paired with (v1, v2, v0) (embree_tlas.ipp:160-166)
    const float w0 = b1, w1 = b2, w2 = (1.0f - b1) - b2;
    // normalize(cross(p1 - p0, p2 - p0)) of the mesh's FP64 positions (triangle.ipp:36-44)
    const float4 fn4 = mat.fnrm[id];
    L.fn = mk3(fn4.x, fn4.y, fn4.z);
    const f3 n0 = mk3(V0.nx, V0.ny, V0.nz), n1 = mk3(V1.nx, V1.ny, V1.nz), n2 = mk3(V2.nx, V2.ny, V2.nz);
    L.Nl = L.fn;
    if (mat.smooth) L.Nl = normalize(add(add(mul(w0, n1), mul(w1, n2)), mul(w2, n0)));
    L.u = (w0 * V1.u + w1 * V2.u) + w2 * V0.u;
    L.v = (w0 * V1.v + w1 * V2.v) + w2 * V0.v;
    // Material::getAlbedo (material.ipp:114-116): a spectral map per band, or a grey map x albedo
    float mv[NB];
    if (mat.tex) {
        const Taps T = bilinear_taps(mat.tw, mat.th, L.u, L.v);
        if (FEAT && mat.tch == NB) {
#pragma unroll
            for (int k = 0; k < NB; ++k) mv[k] = taps_eval(T, mat.tex, NB, k);
        } else {
            const float texv = taps_eval(T, mat.tex, 1, 0);
#pragma unroll
            for (int k = 0; k < NB; ++k) mv[k] = texv * mat.albedo[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < NB; ++k) mv[k] = mat.albedo[k];
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        float av = (w0 * V1.alb[k] + w1 * V2.alb[k]) + w2 * V0.alb[k];
        L.alb[k] = av * mv[k];
    }
    f3 P = add(add(mul(w0, p1), mul(w1, p2)), mul(w2, p0));
    {  // computeShadingPoint (cpu_path_tracer.ipp:481-511)
        f3 q0 = sub(P, mul(dot(sub(P, p1), n1), n1));
        f3 q1 = sub(P, mul(dot(sub(P, p2), n2), n2));
        f3 q2 = sub(P, mul(dot(sub(P, p0), n0), n0));
        f3 sp = add(add(mul(q0, w0), mul(q1, w1)), mul(q2, w2));
        f3 diff = sub(sp, P);
        if (dot(diff, L.Nl) > 0.0f) P = sp;
    }
    L.P = P;
}

// Tangent frame, normal map, albedo and offset origin from global-frame
// quantities (cpu_path_tracer.ipp:307-330,386): N is the global shading normal,
// fng the global face normal the offset follows, P the global shading point,
// (u, v) the hit's uv for the normal map (Material::getNormal,
// material.ipp:129-143: the tangent-space normal through [tg bt N], normalised).
// FEAT = false leaves the normal map out.
template <bool FEAT = true>
__device__ __forceinline__ Surface finish_surface(f3 P, f3 N, f3 fng, const float alb[NB], const MatF& mat, float u,
                                                  float v) {
    Surface S;
    S.Nf = N;
    const f3 arb = (fabsf(N.x) < 0.9f) ? mk3(1.f, 0.f, 0.f) : mk3(0.f, 1.f, 0.f);
    S.tg = normalize(cross(arb, N));
    S.bt = cross(N, S.tg);
    S.N = N;
    if (FEAT && mat.nmap) {
        const Taps T = bilinear_taps(mat.nw, mat.nh, u, v);
        const float nx = taps_eval(T, mat.nmap, 3, 0), ny = taps_eval(T, mat.nmap, 3, 1),
                    nz = taps_eval(T, mat.nmap, 3, 2);
        S.N = normalize(mk3((S.tg.x * nx + S.bt.x * ny) + N.x * nz, (S.tg.y * nx + S.bt.y * ny) + N.y * nz,
                            (S.tg.z * nx + S.bt.z * ny) + N.z * nz));
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) S.alb[k] = alb[k];
    S.org = offset_point(P, fng);
    return S;
}

// A hit on a mesh whose frame is the render frame (identity instance).
template <bool FEAT = true>
__device__ __forceinline__ Surface surface_at(const GVertex* __restrict__ verts, const uint32_t* __restrict__ idx,
                                              const MatF& mat, uint32_t id, float b1, float b2) {
    LocalHit L;
    local_hit<FEAT>(verts, idx, mat, id, b1, b2, L);
    return finish_surface<FEAT>(L.P, L.Nl, L.fn, L.alb, mat, L.u, L.v);
}

// One light's direct-lighting term (cpu_path_tracer.ipp:399-428): returns
// false when the light cannot be sampled (pdf == 0); otherwise the shadow ray
// (S.org, dir, tmax = dist) and the contribution added if it is unoccluded.
// The same from the surface's origin, shading normal and albedo alone (the
// several-light renderer keeps only these while it traces the shadow rays).
__device__ __forceinline__ bool light_contrib_at(f3 org, f3 N, const float alb[NB], const LightF& L,
                                                 const float beta[NB], uint32_t seed, uint32_t pix, uint32_t s,
                                                 uint32_t dim0, f3& dir, float& dist, float contrib[NB]) {
    float lpdf, lrad[NB];
    sample_light(L, org, seed, pix, s, dim0, dir, dist, lpdf, lrad);
    if (!(lpdf > 0.0f)) return false;
    float ndl = dot(N, dir);
    float mx = fmaxf(ndl, 0.0f);
    float cth = fmaxf(0.0f, dot(dir, N));
    float mpdf = (ndl <= 0.0f) ? 0.0f : ndl * kInvPi;
    float wgt = power_heuristic(lpdf, mpdf);
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        float bsdf = (alb[k] * kInvPi) * mx;
        contrib[k] = __fdiv_rn((((beta[k] * bsdf) * lrad[k]) * cth) * wgt, lpdf);
    }
    return true;
}
__device__ __forceinline__ bool light_contrib(const Surface& S, const LightF& L, const float beta[NB], uint32_t seed,
                                              uint32_t pix, uint32_t s, uint32_t dim0, f3& dir, float& dist,
                                              float contrib[NB]) {
    return light_contrib_at(S.org, S.N, S.alb, L, beta, seed, pix, s, dim0, dir, dist, contrib);
}

// BSDF sampling of the next bounce (cpu_path_tracer.ipp:431-461,
// material.ipp:8-33, sampling.ipp:53-84) and Russian roulette.  Updates beta
// and returns whether the path continues along nd.
__device__ __forceinline__ bool sample_bounce(const Surface& S, const Lights& lights, float beta[NB], uint32_t seed,
                                              uint32_t pix, uint32_t s, uint32_t bounce, int rr, int rr_start,
                                              f3& nd) {
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
        float sn, cs;
        vgm::sincos_f32(th, sn, cs);
        dx = r * cs;
        dy = r * sn;
    }
    float lz = __builtin_sqrtf(fmaxf(0.0f, (1.0f - dx * dx) - dy * dy));
    float mpdf = lz * kInvPi;
    if (!(mpdf > 0.0f)) return false;
    // the frame of the normal before the normal map (material_sampling.hpp:15-21);
    // the weights below use the mapped normal (material.ipp:8-67)
    nd = mk3((S.tg.x * dx + S.bt.x * dy) + S.Nf.x * lz, (S.tg.y * dx + S.bt.y * dy) + S.Nf.y * lz,
             (S.tg.z * dx + S.bt.z * dy) + S.Nf.z * lz);
    float ndd = dot(S.N, nd);
    float mx = fmaxf(ndd, 0.0f);
    float cth = fmaxf(0.0f, dot(nd, S.N));
    float lp = 0.0f;
    for (int l = 0; l < lights.n; ++l) lp = lp + light_pdf(lights.l[l], S.org, nd);
    if (lights.n > 0) lp = __fdiv_rn(lp, (float)lights.n);
    float wgt = power_heuristic(mpdf, lp);
    float qmax = 0.0f;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        float bsdf = (S.alb[k] * kInvPi) * mx;
        beta[k] = beta[k] * __fdiv_rn((bsdf * cth) * wgt, mpdf);
        qmax = fmaxf(qmax, beta[k]);
    }
    if (rr && (int)bounce + 1 >= rr_start) {
        float qq = fminf(1.0f, qmax);
        float ur = rng_u01(seed, pix, s, db | SLOT_RR);
        if (ur >= qq) return false;
#pragma unroll
        for (int k = 0; k < NB; ++k) beta[k] = __fdiv_rn(beta[k], qq);
    }
    return true;
}

// Kahan-compensated add of x into (s, c).
__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

}  // namespace vg

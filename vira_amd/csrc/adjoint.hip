// adjoint.hip -- fixed-visibility direct illumination: a forward pass that can
// keep its hit records, and the hand-written adjoint giving dL/dheights (DEM
// grid nodes) and dL/dtexels (grey albedo texture) from dL/dimage.
//
// Forward model per (pixel, sample): the bounce-0 direct-light term of
// processIntersection (cpu_path_tracer.ipp:284-428) -- identical arithmetic to
// the path tracer with bounces = 0, so the image equals vira_gpu_render's.
// The adjoint holds the records fixed (triangle, barycentrics, per-light
// visibility) and differentiates, for point lights,
//   L_k = albedo_k * I_k / |L - P|^2 * c^2 / (pi (1 + (c/pi)^2)),  c = N . (L - P)/|L - P|,
// through the shading point P (barycentric point moved by computeShadingPoint,
// cpu_path_tracer.ipp:481-511), the shading normal N (face normal, or the
// interpolated vertex normals when smooth), the vertex normals (normalised
// sums of unit face normals, mesh.ipp:199-233), the node positions and finally
// the heights along d position / d height.  The shadow-ray ULP offset
// (ray.ipp:107-139) is piecewise constant and contributes no gradient.
//
// Three passes:
//   k_di_adjoint    one lane per pixel (8x8 tiles): reads or re-traces each
//                   sample's record, scatters dL/dposition and dL/dnormal of the
//                   hit triangle's corners and dL/dtexel with float atomics;
//   k_adj_ngrad     per vertex: dL/dnormal -> dL/d(sum of face normals);
//   k_adj_heights   per vertex: gathers the face-normal gradients of incident
//                   triangles (CSR adjacency) onto its position and projects
//                   on d position / d height.
#include <cstring>

#include "context.hpp"
#include "scene_params.hpp"
#include "shading.hpp"
#include "traverse.hpp"

namespace vg {
namespace {

constexpr int kDiBS = 64;  // one wave = one 8x8 pixel tile
constexpr uint32_t kMiss = 0xffffffffu;

// One sample of one pixel: camera ray, closest hit, one shadow ray per light.
// Returns the record (id, b1 bits, b2 bits, visibility bits) and, in v, the
// sample's radiance (0 + lit contributions in light order).
template <int W>
__device__ __forceinline__ uint4 di_sample(const void* __restrict__ nodes, const float4* __restrict__ tris,
                                           const GVertex* __restrict__ verts, const uint32_t* __restrict__ idx,
                                           const MatF& mat, const Lights& lights, const CamF& cam, uint32_t seed,
                                           uint32_t x, uint32_t y, uint32_t pix, uint32_t s, uint32_t* lds_ref,
                                           float* lds_dist, float v[NB], uint32_t& nshadow) {
    uint4 r = make_uint4(kMiss, 0u, 0u, 0u);
#pragma unroll
    for (int k = 0; k < NB; ++k) v[k] = 0.0f;
    const f3 o = mk3(cam.ox, cam.oy, cam.oz);
    const f3 d = camera_dir(cam, x, y, pix, seed, s);
    const HitRec h = traverse<W>(nodes, tris, o, d, kInf, false, lds_ref, lds_dist, kDiBS);
    if (h.id == kMiss) return r;
    r.x = h.id;
    r.y = __float_as_uint(h.b1);
    r.z = __float_as_uint(h.b2);
    const Surface S = surface_at(verts, idx, mat, h.id, h.b1, h.b2);
    float beta[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) beta[k] = 1.0f;
    for (int l = 0; l < lights.n; ++l) {
        f3 ld;
        float dist, c[NB];
        if (!light_contrib(S, lights.l[l], beta, seed, pix, s, SLOT_LIGHT0 + 2u * (uint32_t)l, ld, dist, c)) continue;
        nshadow++;
        const HitRec sh = traverse<W>(nodes, tris, S.org, ld, dist, true, lds_ref, lds_dist, kDiBS);
        if (sh.id != kMiss) continue;
        r.w |= 1u << l;
#pragma unroll
        for (int k = 0; k < NB; ++k) v[k] = v[k] + c[k];
    }
    return r;
}

template <int W>
__global__ void __launch_bounds__(kDiBS, 4) k_di_forward(const void* __restrict__ nodes, const float4* __restrict__ tris,
                                                      const GVertex* __restrict__ verts,
                                                      const uint32_t* __restrict__ idx, MatF mat, Lights lights,
                                                      CamF cam, uint32_t seed, int spp, uint32_t tiles_x, uint32_t nq,
                                                      float* __restrict__ radiance, uint4* __restrict__ rec,
                                                      unsigned long long* __restrict__ counters) {
    __shared__ uint32_t s_ref[kStackLds * kDiBS];
    __shared__ float s_dist[kStackLds * kDiBS];
    const uint32_t lane = threadIdx.x;
    const uint32_t q = blockIdx.x * kDiBS + lane;
    uint32_t x, y;
    tile_xy(q, tiles_x, x, y);
    const bool valid = x < (uint32_t)cam.W && y < (uint32_t)cam.H;
    const uint32_t pix = x + y * (uint32_t)cam.W;
    uint32_t nshadow = 0;
    if (valid) {
        float sum[NB], comp[NB], v[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) sum[k] = comp[k] = 0.0f;
        for (int s = 0; s < spp; ++s) {
            const uint4 r = di_sample<W>(nodes, tris, verts, idx, mat, lights, cam, seed, x, y, pix, (uint32_t)s,
                                         s_ref + lane, s_dist + lane, v, nshadow);
#pragma unroll
            for (int k = 0; k < NB; ++k) kahan_add(sum[k], comp[k], v[k]);
            if (rec) rec[(size_t)s * nq + q] = r;
        }
        const float sf = (float)spp;
#pragma unroll
        for (int k = 0; k < NB; ++k) radiance[(size_t)NB * pix + k] = __fdiv_rn(sum[k], sf);
    }
    wave_add(&counters[0], valid ? (uint32_t)spp : 0u);
    wave_add(&counters[1], nshadow);
}

// Bilinear taps of Image::sampleUVs + interpolatePixel (image.ipp:589-670),
// as tex_bilinear evaluates them: texel indices and weights.
__device__ __forceinline__ void tex_taps(const MatF& m, float u, float v, uint32_t ti[4], float tw[4]) {
    float rx = (float)m.tw, ry = (float)m.th;
    float px = fmodf(u * rx, rx), py = fmodf(v * ry, ry);
    if (px < 0.f) px += rx;
    if (py < 0.f) py += ry;
    if (!(px >= 0.f)) px = 0.f;
    if (!(py >= 0.f)) py = 0.f;
    int x0 = min((int)floorf(px), m.tw - 1), y0 = min((int)floorf(py), m.th - 1);
    int x1 = min(x0 + 1, m.tw - 1), y1 = min(y0 + 1, m.th - 1);
    float dx = px - (float)x0, dy = py - (float)y0;
    ti[0] = (uint32_t)(x0 + y0 * m.tw);
    ti[1] = (uint32_t)(x1 + y0 * m.tw);
    ti[2] = (uint32_t)(x0 + y1 * m.tw);
    ti[3] = (uint32_t)(x1 + y1 * m.tw);
    tw[0] = (1.0f - dx) * (1.0f - dy);
    tw[1] = dx * (1.0f - dy);
    tw[2] = (1.0f - dx) * dy;
    tw[3] = dx * dy;
}

__device__ __forceinline__ void atomic_add3(float4* p, f3 g) {
    float* a = reinterpret_cast<float*>(p);
    unsafeAtomicAdd(a, g.x);
    unsafeAtomicAdd(a + 1, g.y);
    unsafeAtomicAdd(a + 2, g.z);
}

__device__ __forceinline__ bool nonzero(f3 g) { return g.x != 0.0f || g.y != 0.0f || g.z != 0.0f; }

// Gradient of one record's contribution, G = dL/d(contribution) per bin.
__device__ __forceinline__ void di_backward(const GVertex* __restrict__ verts, const uint32_t* __restrict__ idx,
                                            const MatF& mat, const Lights& lights, uint4 r, const float G[NB],
                                            bool want_geo, float4* __restrict__ gpos, float4* __restrict__ gnrm,
                                            float* __restrict__ gtex) {
    const uint32_t id = r.x;
    const float b1 = __uint_as_float(r.y), b2 = __uint_as_float(r.z);
    const uint32_t vi[3] = {idx[3 * (size_t)id], idx[3 * (size_t)id + 1], idx[3 * (size_t)id + 2]};
    f3 p[3], n[3];
    float au[3], avv[3], va[3][NB];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const GVertex V = verts[vi[a]];
        p[a] = mk3(V.px, V.py, V.pz);
        n[a] = mk3(V.nx, V.ny, V.nz);
        au[a] = V.u;
        avv[a] = V.v;
#pragma unroll
        for (int k = 0; k < NB; ++k) va[a][k] = V.alb[k];
    }
    const float om[3] = {(1.0f - b1) - b2, b1, b2};  // weights of corners 0, 1, 2 (embree_tlas.ipp:160-166)
    const f3 ea = sub(p[1], p[0]), eb = sub(p[2], p[0]);
    const f3 e = cross(ea, eb);
    const float le = length(e);
    if (!(le > 0.0f) || !isfinite(le)) return;
    const f3 fn = mul(__fdiv_rn(1.0f, le), e);
    const f3 m = add(add(mul(om[1], n[1]), mul(om[2], n[2])), mul(om[0], n[0]));
    const float lm = length(m);
    if (mat.smooth && !(lm > 0.0f)) return;
    const f3 N = mat.smooth ? mul(__fdiv_rn(1.0f, lm), m) : fn;
    const f3 P0 = add(add(mul(om[1], p[1]), mul(om[2], p[2])), mul(om[0], p[0]));
    float del[3];
    f3 sp = mk3(0.f, 0.f, 0.f);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        del[a] = dot(sub(P0, p[a]), n[a]);
        sp = add(sp, mul(om[a], sub(P0, mul(del[a], n[a]))));
    }
    const bool use = dot(sub(sp, P0), N) > 0.0f;
    const f3 P = use ? sp : P0;
    float texv = 1.0f;
    uint32_t ti[4] = {0, 0, 0, 0};
    float tw[4] = {0.f, 0.f, 0.f, 0.f};
    if (mat.tex) {
        const float u = (om[1] * au[1] + om[2] * au[2]) + om[0] * au[0];
        const float v = (om[1] * avv[1] + om[2] * avv[2]) + om[0] * avv[0];
        tex_taps(mat, u, v, ti, tw);
        texv = ((tw[0] * mat.tex[ti[0]] + tw[1] * mat.tex[ti[1]]) + tw[2] * mat.tex[ti[2]]) + tw[3] * mat.tex[ti[3]];
    }
    float av[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) av[k] = (om[1] * va[1][k] + om[2] * va[2][k]) + om[0] * va[0][k];
    f3 gN = mk3(0.f, 0.f, 0.f), gP = mk3(0.f, 0.f, 0.f);
    float gtexv = 0.0f;
    constexpr float kPi2 = kPi * kPi;
    for (int l = 0; l < lights.n; ++l) {
        if (!((r.w >> l) & 1u)) continue;
        const LightF& L = lights.l[l];
        const f3 toL = sub(mk3(L.px, L.py, L.pz), P);
        const float dist = length(toL);
        const float inv_d = __fdiv_rn(1.0f, dist);
        const f3 dir = mul(inv_d, toL);
        const float c = dot(N, dir);
        if (!(c > 0.0f)) continue;
        const float den = kPi2 + c * c;
        const float Fc = __fdiv_rn(kPi * c * c, den);
        const float dFc = __fdiv_rn(2.0f * kPi * kPi2 * c, den * den);
        const float inv_d2 = inv_d * inv_d;
        float gc = 0.0f, gd = 0.0f;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const float mv = mat.tex ? texv * mat.albedo[k] : mat.albedo[k];
            const float rad = L.value[k] * inv_d2;
            const float a = G[k] * (av[k] * mv) * rad;
            gc += a * dFc;
            gd += a * Fc;
            if (mat.tex) gtexv += G[k] * av[k] * mat.albedo[k] * rad * Fc;
        }
        gd = gd * (-2.0f * inv_d);
        gN = add(gN, mul(gc, dir));
        const f3 gdir = mul(gc, N);
        gP = sub(sub(gP, mul(inv_d, sub(gdir, mul(dot(gdir, dir), dir)))), mul(gd, dir));
    }
    if (gtex && mat.tex && gtexv != 0.0f) {
#pragma unroll
        for (int t = 0; t < 4; ++t) unsafeAtomicAdd(&gtex[ti[t]], gtexv * tw[t]);
    }
    if (!want_geo) return;
    f3 gp[3], gn[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) gp[a] = gn[a] = mk3(0.f, 0.f, 0.f);
    f3 gP0 = gP;
    if (use) {
        gP0 = mul((om[0] + om[1]) + om[2], gP);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float ng = dot(n[a], gP);
            gP0 = sub(gP0, mul(om[a] * ng, n[a]));
            gp[a] = add(gp[a], mul(om[a] * ng, n[a]));
            gn[a] = sub(gn[a], mul(om[a], add(mul(ng, sub(P0, p[a])), mul(del[a], gP))));
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) gp[a] = add(gp[a], mul(om[a], gP0));
    if (mat.smooth) {
        const f3 gm = mul(__fdiv_rn(1.0f, lm), sub(gN, mul(dot(gN, N), N)));
#pragma unroll
        for (int a = 0; a < 3; ++a) gn[a] = add(gn[a], mul(om[a], gm));
    } else {
        // N = fn = e/|e|, e = cross(p1 - p0, p2 - p0)
        const f3 ge = mul(__fdiv_rn(1.0f, le), sub(gN, mul(dot(gN, fn), fn)));
        const f3 ga = cross(eb, ge), gb = cross(ge, ea);
        gp[1] = add(gp[1], ga);
        gp[2] = add(gp[2], gb);
        gp[0] = sub(gp[0], add(ga, gb));
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (nonzero(gp[a])) atomic_add3(&gpos[vi[a]], gp[a]);
        if (nonzero(gn[a])) atomic_add3(&gnrm[vi[a]], gn[a]);
    }
}

template <int W, bool RETRACE>
__global__ void __launch_bounds__(kDiBS, 4) k_di_adjoint(const void* __restrict__ nodes, const float4* __restrict__ tris,
                                                      const GVertex* __restrict__ verts,
                                                      const uint32_t* __restrict__ idx, MatF mat, Lights lights,
                                                      CamF cam, uint32_t seed, int spp, uint32_t tiles_x, uint32_t nq,
                                                      const float* __restrict__ gimg, const uint4* __restrict__ rec,
                                                      int want_geo, float4* __restrict__ gpos,
                                                      float4* __restrict__ gnrm, float* __restrict__ gtex,
                                                      unsigned long long* __restrict__ counters) {
    __shared__ uint32_t s_ref[kStackLds * kDiBS];
    __shared__ float s_dist[kStackLds * kDiBS];
    const uint32_t lane = threadIdx.x;
    const uint32_t q = blockIdx.x * kDiBS + lane;
    uint32_t x, y;
    tile_xy(q, tiles_x, x, y);
    const bool valid = x < (uint32_t)cam.W && y < (uint32_t)cam.H;
    const uint32_t pix = x + y * (uint32_t)cam.W;
    uint32_t nshadow = 0;
    if (valid) {
        float G[NB];
        bool any = false;
        const float sf = (float)spp;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            G[k] = __fdiv_rn(gimg[(size_t)NB * pix + k], sf);
            any = any || G[k] != 0.0f;
        }
        for (int s = 0; any && s < spp; ++s) {
            uint4 r;
            if (!RETRACE) {
                r = rec[(size_t)s * nq + q];
            } else {
                float v[NB];
                r = di_sample<W>(nodes, tris, verts, idx, mat, lights, cam, seed, x, y, pix, (uint32_t)s,
                                 s_ref + lane, s_dist + lane, v, nshadow);
            }
            if (r.x == kMiss || r.w == 0u) continue;
            di_backward(verts, idx, mat, lights, r, G, want_geo != 0, gpos, gnrm, gtex);
        }
    }
    if (RETRACE) {
        wave_add(&counters[0], valid ? (uint32_t)spp : 0u);
        wave_add(&counters[1], nshadow);
    }
}

// dL/dn_v -> dL/ds_v for n_v = s_v / |s_v|, s_v = sum of unit face normals.
__global__ void k_adj_ngrad(size_t nv, const uint32_t* __restrict__ offs, const uint32_t* __restrict__ count,
                            const uint32_t* __restrict__ adj, const float4* __restrict__ fnrm,
                            const float4* __restrict__ gnrm, float4* __restrict__ gs) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const float4 g4 = gnrm[v];
    const f3 g = mk3(g4.x, g4.y, g4.z);
    f3 out = mk3(0.f, 0.f, 0.f);
    if (nonzero(g)) {
        f3 s = mk3(0.f, 0.f, 0.f);
        const uint32_t b = offs[v], n = count[v];
        for (uint32_t a = 0; a < n; ++a) {
            const float4 fn = fnrm[adj[b + a]];
            if (fn.w != 0.0f) s = add(s, mk3(fn.x, fn.y, fn.z));
        }
        const float ls = length(s);
        if (ls > 0.0f) {
            const float il = __fdiv_rn(1.0f, ls);
            const f3 nn = mul(il, s);
            out = mul(il, sub(g, mul(dot(g, nn), nn)));
        }
    }
    gs[v] = make_float4(out.x, out.y, out.z, 0.0f);
}

// d position / d height of grid node v: +z on the flat grid, the radial unit
// vector of the gnomonic drape on the sphere (FP64, as the positions).
__device__ __forceinline__ f3 dem_up(const vira_gpu_projection& p, int W, int H, size_t v) {
    if (p.kind == 0) return mk3(0.f, 0.f, 1.f);
    const int i = (int)(v % (size_t)W), j = (int)(v / (size_t)W);
    const double cx = (double)(W - 1) * 0.5, cy = (double)(H - 1) * 0.5;
    const double a = __ddiv_rn(p.spacing, p.radius);
    const double dx = 1.0, dy = ((double)i - cx) * a, dz = (cy - (double)j) * a;
    const double inv = __ddiv_rn(1.0, __dsqrt_rn((dx * dx + dy * dy) + dz * dz));
    return mk3((float)(dx * inv), (float)(dy * inv), (float)(dz * inv));
}

// dL/dh_v = (dL/dp_v + sum over incident faces of their normal gradients
// pulled back onto corner v) . d p_v / d h_v.
__global__ void k_adj_heights(size_t nv, int W, int H, vira_gpu_projection proj, const uint32_t* __restrict__ offs,
                              const uint32_t* __restrict__ count, const uint32_t* __restrict__ adj,
                              const uint32_t* __restrict__ idx, const GVertex* __restrict__ verts,
                              const float4* __restrict__ gpos, const float4* __restrict__ gs,
                              float* __restrict__ dh) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const float4 g4 = gpos[v];
    f3 acc = mk3(g4.x, g4.y, g4.z);
    const uint32_t b = offs[v], n = count[v];
    for (uint32_t a = 0; a < n; ++a) {
        const uint32_t t = adj[b + a];
        const uint32_t i0 = idx[3 * (size_t)t], i1 = idx[3 * (size_t)t + 1], i2 = idx[3 * (size_t)t + 2];
        const float4 s0 = gs[i0], s1 = gs[i1], s2 = gs[i2];
        const f3 gfn = mk3((s0.x + s1.x) + s2.x, (s0.y + s1.y) + s2.y, (s0.z + s1.z) + s2.z);
        if (!nonzero(gfn)) continue;
        const f3 p0 = vpos(verts, i0), p1 = vpos(verts, i1), p2 = vpos(verts, i2);
        if (isinf(p0.x) || isinf(p1.x) || isinf(p2.x)) continue;
        const f3 ea = sub(p1, p0), eb = sub(p2, p0);
        const f3 e = cross(ea, eb);
        const float le = length(e);
        if (!(le > 0.0f)) continue;
        const float il = __fdiv_rn(1.0f, le);
        const f3 fn = mul(il, e);
        const f3 ge = mul(il, sub(gfn, mul(dot(gfn, fn), fn)));
        const f3 ga = cross(eb, ge), gb = cross(ge, ea);
        if (i1 == (uint32_t)v) acc = add(acc, ga);
        if (i2 == (uint32_t)v) acc = add(acc, gb);
        if (i0 == (uint32_t)v) acc = sub(acc, add(ga, gb));
    }
    dh[v] = dot(acc, dem_up(proj, W, H, v));
}

__global__ void k_zero(size_t n, float* __restrict__ p) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

void zero_async(Ctx& c, float* p, size_t n) {
    if (n == 0) return;
    k_zero<<<blocks_for(n, 256), 256, 0, c.stream>>>(n, p);
    VG_HIP(hipGetLastError());
}

bool same_camera(const vira_gpu_camera& a, const vira_gpu_camera& b) {
    if (a.width != b.width || a.height != b.height || a.z_dir != b.z_dir) return false;
    for (int k = 0; k < 6; ++k)
        if (a.kinv[k] != b.kinv[k]) return false;
    for (int k = 0; k < 9; ++k)
        if (a.rot[k] != b.rot[k]) return false;
    for (int k = 0; k < 3; ++k)
        if (a.position[k] != b.position[k]) return false;
    return true;
}

MeshG& di_scene_mesh(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_di_opts& o, const char* who) {
    VG_CHECK(c.scene_mesh >= 0, VIRA_GPU_ERR_STATE, std::string(who) + ": no TLAS built");
    VG_CHECK(!c.instanced, VIRA_GPU_ERR_UNSUPPORTED,
             std::string(who) + ": direct illumination takes a TLAS of one identity instance");
    MeshG& m = c.mesh(c.scene_mesh);
    VG_CHECK(m.blas.built, VIRA_GPU_ERR_STATE, std::string(who) + ": BLAS not built");
    VG_CHECK(cam.width > 0 && cam.height > 0, VIRA_GPU_ERR_INVALID, std::string(who) + ": empty image");
    VG_CHECK((size_t)cam.width * cam.height < (1ull << 30), VIRA_GPU_ERR_INVALID,
             std::string(who) + ": image too large");
    VG_CHECK(o.samples >= 1 && o.samples <= 65536, VIRA_GPU_ERR_INVALID, std::string(who) + ": samples 1..65536");
    VG_CHECK(!m.mat.normal_map_dev && m.mat.tex_channels <= 1, VIRA_GPU_ERR_UNSUPPORTED,
             std::string(who) + ": the differentiable model takes no normal map and a grey albedo texture");
    VG_CHECK(!c.env_has_ambient, VIRA_GPU_ERR_UNSUPPORTED, std::string(who) + ": no ambient light in the model");
    return m;
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    explicit Timer(hipStream_t s) {
        VG_HIP(hipEventCreate(&a));
        VG_HIP(hipEventCreate(&b));
        VG_HIP(hipEventRecord(a, s));
    }
    void stop(hipStream_t s) { VG_HIP(hipEventRecord(b, s)); }
    double ms() {
        float f = 0.f;
        VG_HIP(hipEventElapsedTime(&f, a, b));
        return f;
    }
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace

void di_render(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_di_opts& o, float* radiance,
               vira_gpu_di_stats* st) {
    VG_CHECK(radiance, VIRA_GPU_ERR_INVALID, "di_render: null radiance");
    MeshG& m = di_scene_mesh(c, cam, o, "di_render");
    const TileGrid g(cam);
    const CamF cf = cam_params(cam);
    const Lights lf = upload_lights(c, c.r_lights);
    const MatF mf = mat_params(m);
    const uint64_t rec_bytes = (uint64_t)o.samples * g.nq * 16ull;
    const bool keep = rec_bytes <= o.record_budget;
    c.di_rec_valid = false;
    if (keep) c.di_rec.alloc(rec_bytes);
    c.r_counters.alloc(64 * 8);
    unsigned long long* cnt = c.r_counters.as<unsigned long long>();
    VG_HIP(hipMemsetAsync(cnt, 0, 64 * 8, c.stream));
    Timer tm(c.stream);
    auto kern = m.blas.width == 4 ? k_di_forward<4> : k_di_forward<2>;
    kern<<<g.tiles, kDiBS, 0, c.stream>>>(m.blas.nodes.p, m.blas.tris.as<float4>(), m.verts.as<GVertex>(),
                                         m.idx.as<uint32_t>(), mf, lf, cf, o.seed, o.samples, g.tiles_x, g.nq,
                                         radiance, keep ? c.di_rec.as<uint4>() : nullptr, cnt);
    VG_HIP(hipGetLastError());
    tm.stop(c.stream);
    unsigned long long hc[2];
    VG_HIP(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    c.di_rec_valid = keep;
    c.di_gen = c.scene_gen;
    c.di_cam = cam;
    c.di_spp = o.samples;
    c.di_seed = o.seed;
    if (st) {
        *st = vira_gpu_di_stats{};
        st->primary_rays = hc[0];
        st->shadow_rays = hc[1];
        st->ms = tm.ms();
        st->record_bytes = keep ? rec_bytes : 0;
    }
}

void di_adjoint(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_di_opts& o, const float* grad_image,
                float* grad_heights, float* grad_texture, vira_gpu_di_stats* st) {
    VG_CHECK(grad_image, VIRA_GPU_ERR_INVALID, "adjoint: null dL/dimage");
    MeshG& m = di_scene_mesh(c, cam, o, "adjoint");
    VG_CHECK(!grad_heights || m.is_dem, VIRA_GPU_ERR_INVALID,
             "adjoint: height gradients need a scene mesh made by dem_mesh");
    VG_CHECK(!grad_texture || m.mat.texture_dev, VIRA_GPU_ERR_INVALID,
             "adjoint: texel gradients need a material texture");
    for (const vira_gpu_light& L : c.lights)
        VG_CHECK(L.kind == 0, VIRA_GPU_ERR_UNSUPPORTED, "adjoint: point lights only");
    const TileGrid g(cam);
    const CamF cf = cam_params(cam);
    const Lights lf = upload_lights(c, c.r_lights);
    const MatF mf = mat_params(m);
    const bool use_rec = c.di_rec_valid && c.di_gen == c.scene_gen && same_camera(c.di_cam, cam) &&
                         c.di_spp == o.samples && c.di_seed == o.seed;
    const bool geo = grad_heights != nullptr;
    const size_t nv = m.nv;
    if (geo) {
        mesh_adjacency(c, m);
        c.di_gpos.alloc(std::max<size_t>(1, nv) * 16);
        c.di_gnrm.alloc(std::max<size_t>(1, nv) * 16);
        c.di_gs.alloc(std::max<size_t>(1, nv) * 16);
    }
    c.r_counters.alloc(64 * 8);
    unsigned long long* cnt = c.r_counters.as<unsigned long long>();
    VG_HIP(hipMemsetAsync(cnt, 0, 64 * 8, c.stream));
    Timer tm(c.stream);
    if (geo) {
        zero_async(c, c.di_gpos.as<float>(), 4 * nv);
        zero_async(c, c.di_gnrm.as<float>(), 4 * nv);
    }
    if (grad_texture) zero_async(c, grad_texture, (size_t)m.mat.tex_width * m.mat.tex_height);
    Timer tt(c.stream);
    auto kern = use_rec ? (m.blas.width == 4 ? k_di_adjoint<4, false> : k_di_adjoint<2, false>)
                        : (m.blas.width == 4 ? k_di_adjoint<4, true> : k_di_adjoint<2, true>);
    kern<<<g.tiles, kDiBS, 0, c.stream>>>(m.blas.nodes.p, m.blas.tris.as<float4>(), m.verts.as<GVertex>(),
                                         m.idx.as<uint32_t>(), mf, lf, cf, o.seed, o.samples, g.tiles_x, g.nq,
                                         grad_image, use_rec ? c.di_rec.as<uint4>() : nullptr, geo ? 1 : 0,
                                         geo ? c.di_gpos.as<float4>() : nullptr,
                                         geo ? c.di_gnrm.as<float4>() : nullptr, grad_texture, cnt);
    VG_HIP(hipGetLastError());
    tt.stop(c.stream);
    if (geo && nv) {
        k_adj_ngrad<<<blocks_for(nv, 256), 256, 0, c.stream>>>(nv, m.adj_offs.as<uint32_t>(),
                                                                m.adj_count.as<uint32_t>(), m.adj.as<uint32_t>(),
                                                                m.fnrm.as<float4>(), c.di_gnrm.as<float4>(),
                                                                c.di_gs.as<float4>());
        VG_HIP(hipGetLastError());
        k_adj_heights<<<blocks_for(nv, 256), 256, 0, c.stream>>>(
            nv, m.dem_w, m.dem_h, m.dem_proj, m.adj_offs.as<uint32_t>(), m.adj_count.as<uint32_t>(),
            m.adj.as<uint32_t>(), m.idx.as<uint32_t>(), m.verts.as<GVertex>(), c.di_gpos.as<float4>(),
            c.di_gs.as<float4>(), grad_heights);
        VG_HIP(hipGetLastError());
    }
    tm.stop(c.stream);
    unsigned long long hc[2];
    VG_HIP(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    if (st) {
        *st = vira_gpu_di_stats{};
        st->primary_rays = hc[0];
        st->shadow_rays = hc[1];
        st->ms = tm.ms();
        st->trace_ms = use_rec ? 0.0 : tt.ms();
        st->used_records = use_rec ? 1 : 0;
    }
}

}  // namespace vg

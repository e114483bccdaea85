// oracle_capi.cpp -- extern "C" surface of the CPU oracle for ctypes.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include <atomic>
#include <chrono>
#include <thread>

#include "oracle_core.hpp"
#include "oracle_diff.hpp"
#include "../vira_amd/csrc/fp_trig.hpp"

using namespace vo;

namespace {
struct OMesh {
    Mesh m;
    BVH b;
    double build_ms = 0;
};
}  // namespace

extern "C" {

void orc_fbm_heights(int W, int H, uint32_t seed, double amplitude, int octaves, double base_cells,
                     float* out, int nthreads) {
    auto work = [&](int j0, int j1) {
        for (int j = j0; j < j1; ++j)
            for (int i = 0; i < W; ++i)
                out[(size_t)j * W + i] = (float)fbm_height(i, j, seed, amplitude, octaves, base_cells);
    };
    int nt = std::max(1, std::min(nthreads, H));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(work, (int)((int64_t)H * t / nt), (int)((int64_t)H * (t + 1) / nt));
    for (auto& x : th) x.join();
}

void orc_vertex_albedo(size_t nv, uint32_t seed, float* out) {
    for (size_t v = 0; v < nv; ++v) out[v] = vertex_albedo((uint32_t)v, seed);
}

int64_t orc_dem_indices(const float* heights, int W, int H, uint32_t* idx_out) {
    std::vector<uint32_t> idx;
    dem_indices(heights, W, H, idx);
    if (idx_out) std::copy(idx.begin(), idx.end(), idx_out);
    return (int64_t)(idx.size() / 3);
}

// Positions of the DEM grid, FP64 and rebased FP32.  Non-finite heights give
// +inf positions (dem.ipp:186-188).
void orc_dem_positions(const float* heights, int W, int H, const Projection* proj, float* pos,
                       double* pos64) {
    for (int j = 0; j < H; ++j)
        for (int i = 0; i < W; ++i) {
            size_t v = (size_t)j * W + i;
            float h = heights[v];
            if (!std::isfinite(h)) {
                for (int k = 0; k < 3; ++k) {
                    pos[3 * v + k] = INFINITY;
                    if (pos64) pos64[3 * v + k] = INFINITY;
                }
                continue;
            }
            V3<double> p = dem_world(*proj, W, H, i, j, (double)h);
            double r[3] = {p.x - proj->offset[0], p.y - proj->offset[1], p.z - proj->offset[2]};
            for (int k = 0; k < 3; ++k) {
                pos[3 * v + k] = (float)r[k];
                if (pos64) pos64[3 * v + k] = r[k];
            }
        }
}

void orc_vertex_normals(const float* pos, const double* pos64, size_t nv, const uint32_t* idx, size_t nt,
                        float* nrm) {
    vertex_normals(pos, pos64, nv, idx, nt, nrm);
}

void* orc_mesh_create(const float* pos, const double* pos64, size_t nv, const uint32_t* idx, size_t nt,
                      const float* nrm, const float* alb, const float* uv, int nthreads) {
    auto* o = new OMesh;
    o->m.nv = nv;
    o->m.nt = nt;
    o->m.pos.assign(pos, pos + 3 * nv);
    if (pos64) o->m.pos64.assign(pos64, pos64 + 3 * nv);
    o->m.idx.assign(idx, idx + 3 * nt);
    if (nrm) o->m.nrm.assign(nrm, nrm + 3 * nv);
    else o->m.nrm.assign(3 * nv, 0.0f);
    if (alb) o->m.alb.assign(alb, alb + (size_t)NB * nv);
    else o->m.alb.assign((size_t)NB * nv, 1.0f);
    if (uv) o->m.uv.assign(uv, uv + 2 * nv);
    o->m.fn.resize(4 * nt);
    face_normals(o->m.pos.data(), pos64, o->m.idx.data(), nt, o->m.fn.data());
    auto t0 = std::chrono::steady_clock::now();
    o->b = bvh_build(o->m, nthreads);
    o->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return o;
}

void orc_mesh_destroy(void* h) { delete static_cast<OMesh*>(h); }

double orc_mesh_build_ms(void* h) { return static_cast<OMesh*>(h)->build_ms; }
size_t orc_bvh_num_nodes(void* h) { return static_cast<OMesh*>(h)->b.nodes.size(); }

// BVH invariants: every triangle in exactly one leaf, children inside parents,
// leaf boxes contain their triangles, node count <= 2N-1.  Returns 0 if OK,
// otherwise a code naming the first violated invariant.
int orc_bvh_check(void* h, uint32_t* max_depth, uint32_t* max_leaf) {
    auto* o = static_cast<OMesh*>(h);
    const BVH& b = o->b;
    const Mesh& m = o->m;
    if (m.nt == 0) return b.nodes.empty() ? 0 : 1;
    if (b.nodes.size() > 2 * m.nt - 1) return 2;
    std::vector<uint32_t> seen(m.nt, 0);
    uint32_t md = 0, ml = 0;
    struct E { uint32_t n, d; };
    std::vector<E> st{{0, 1}};
    while (!st.empty()) {
        E e = st.back();
        st.pop_back();
        const BNode& n = b.nodes[e.n];
        md = std::max(md, e.d);
        if (n.triCount > 0) {
            ml = std::max(ml, n.triCount);
            for (uint32_t i = 0; i < n.triCount; ++i) {
                uint32_t t = b.triIdx[n.leftFirst + i];
                if (t >= m.nt) return 3;
                seen[t]++;
                for (int v = 0; v < 3; ++v)
                    for (int k = 0; k < 3; ++k) {
                        double p = m.pos[3 * m.idx[3 * t + v] + k];
                        if (p < n.lo[k] || p > n.hi[k]) return 4;
                    }
            }
            continue;
        }
        for (uint32_t c = n.leftFirst; c < n.leftFirst + 2; ++c) {
            if (c >= b.nodes.size()) return 5;
            for (int k = 0; k < 3; ++k)
                if (b.nodes[c].lo[k] < n.lo[k] || b.nodes[c].hi[k] > n.hi[k]) return 6;
            st.push_back({c, e.d + 1});
        }
    }
    for (size_t t = 0; t < m.nt; ++t)
        if (seen[t] != 1) return 7;
    if (max_depth) *max_depth = md;
    if (max_leaf) *max_leaf = ml;
    return 0;
}

// rays: n x 6 floats (ox oy oz dx dy dz); tmax: n floats (any-hit only, may be null = inf).
// counts[0] = internal nodes visited, counts[1] = triangle tests (summed).
void orc_trace(void* h, const float* rays, size_t n, int anyhit, const float* tmax, int brute,
               float* t_out, uint32_t* id_out, float* b1_out, float* b2_out, uint64_t* counts,
               int nthreads) {
    auto* o = static_cast<OMesh*>(h);
    std::atomic<uint64_t> ci{0}, ct{0};
    auto work = [&](size_t a, size_t e) {
        TravCount tc;
        for (size_t r = a; r < e; ++r) {
            RayT<float> ray{{rays[6 * r], rays[6 * r + 1], rays[6 * r + 2]}, {rays[6 * r + 3], rays[6 * r + 4], rays[6 * r + 5]}};
            float tm = tmax ? tmax[r] : INFINITY;
            Hit<float> hh = brute ? brute_trace<float>(o->m, ray, anyhit != 0, tm)
                                  : bvh_trace<float>(o->m, o->b, ray, anyhit != 0, tm, &tc);
            if (t_out) t_out[r] = hh.t;
            if (id_out) id_out[r] = hh.id;
            if (b1_out) b1_out[r] = hh.b1;
            if (b2_out) b2_out[r] = hh.b2;
        }
        ci += tc.internal;
        ct += tc.tris;
    };
    int nt = std::max(1, nthreads);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(work, n * t / nt, n * (t + 1) / nt);
    for (auto& x : th) x.join();
    if (counts) { counts[0] = ci; counts[1] = ct; }
}

// FP64 twin of orc_trace over the FP64 positions (if given at mesh creation).
void orc_trace64(void* h, const double* rays, size_t n, int anyhit, const double* tmax, double* t_out,
                 uint32_t* id_out, double* b1_out, double* b2_out, int nthreads) {
    auto* o = static_cast<OMesh*>(h);
    auto work = [&](size_t a, size_t e) {
        for (size_t r = a; r < e; ++r) {
            RayT<double> ray{{rays[6 * r], rays[6 * r + 1], rays[6 * r + 2]}, {rays[6 * r + 3], rays[6 * r + 4], rays[6 * r + 5]}};
            double tm = tmax ? tmax[r] : INFINITY;
            Hit<double> hh = bvh_trace<double>(o->m, o->b, ray, anyhit != 0, tm, nullptr);
            if (t_out) t_out[r] = hh.t;
            if (id_out) id_out[r] = hh.id;
            if (b1_out) b1_out[r] = hh.b1;
            if (b2_out) b2_out[r] = hh.b2;
        }
    };
    int nt = std::max(1, nthreads);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(work, n * t / nt, n * (t + 1) / nt);
    for (auto& x : th) x.join();
}

// Single-triangle watertight test (FP32), for known-answer tests.
int orc_intersect_tri(const float* ray6, const float* tri9, float* t, float* b1, float* b2) {
    RayT<float> ray{{ray6[0], ray6[1], ray6[2]}, {ray6[3], ray6[4], ray6[5]}};
    WTRay<float> w = wt_setup(ray.d);
    V3<float> v0{tri9[0], tri9[1], tri9[2]}, v1{tri9[3], tri9[4], tri9[5]}, v2{tri9[6], tri9[7], tri9[8]};
    return wt_intersect(ray, w, v0, v1, v2, *t, *b1, *b2) ? 1 : 0;
}

// Node records: lohi 6 doubles per node, leftFirst, triCount.
void orc_bvh_nodes(void* h, double* lohi, uint32_t* left, uint32_t* count) {
    auto* o = static_cast<OMesh*>(h);
    for (size_t n = 0; n < o->b.nodes.size(); ++n) {
        const BNode& b = o->b.nodes[n];
        for (int k = 0; k < 3; ++k) { lohi[6 * n + k] = b.lo[k]; lohi[6 * n + 3 + k] = b.hi[k]; }
        left[n] = b.leftFirst;
        count[n] = b.triCount;
    }
}

void orc_bvh_tri_index(void* h, uint32_t* out) {
    auto* o = static_cast<OMesh*>(h);
    std::copy(o->b.triIdx.begin(), o->b.triIdx.end(), out);
}

void orc_offset_point(const float* p, const float* n, float* out) {
    V3<float> r = offset_point_f32({p[0], p[1], p[2]}, {n[0], n[1], n[2]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

uint32_t orc_lowbias32(uint32_t x) { return lowbias32(x); }
float orc_rng_u01(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
    return rng_u01(seed, pixel, sample, dim);
}

// precision: 32 = FP32 twin, 64 = FP64 reference.
void orc_mesh_set_id(void* h, uint32_t id) { static_cast<OMesh*>(h)->m.id = id; }

void orc_render(void* h, int precision, const Camera* cam, const Light* lights, int nl, const Material* mat,
                const Options* opt, float* radiance, float* direct, float* indirect, float* albedo,
                float* normal, float* depth, float* alpha, uint32_t* tri_id, uint64_t* counts,
                uint32_t* records, int nthreads, const int* pixel_list, size_t npix, float* normal_camera,
                uint32_t* mesh_id, uint32_t* material_id) {
    auto* o = static_cast<OMesh*>(h);
    Outputs out;
    out.radiance = radiance; out.direct = direct; out.indirect = indirect; out.albedo = albedo;
    out.normal = normal; out.depth = depth; out.alpha = alpha; out.tri_id = tri_id; out.counts = counts;
    out.records = records;
    out.normal_camera = normal_camera; out.mesh_id = mesh_id; out.material_id = material_id;
    if (precision == 64) render_f64(o->m, o->b, *cam, lights, nl, *mat, *opt, out, nthreads, pixel_list, npix);
    else render_f32(o->m, o->b, *cam, lights, nl, *mat, *opt, out, nthreads, pixel_list, npix);
}

// Two-level scene of n instances (mesh handles, FP64 column-major model
// matrices, materials); outputs as orc_render plus the first-hit instance.
void orc_render_scene(void* const* meshes, const double* transforms, const Material* mats, int n, int precision,
                      const Camera* cam, const Light* lights, int nl, const Options* opt, float* radiance,
                      float* direct, float* indirect, float* albedo, float* normal, float* depth, float* alpha,
                      uint32_t* tri_id, uint32_t* instance_id, uint64_t* counts, int nthreads, const int* pixel_list,
                      size_t npix, float* normal_camera, uint32_t* mesh_id, uint32_t* material_id) {
    std::vector<InstanceO> inst(n);
    for (int i = 0; i < n; ++i) {
        auto* o = static_cast<OMesh*>(meshes[i]);
        inst[i].m = &o->m;
        inst[i].b = &o->b;
        inst[i].mat = mats[i];
        for (int k = 0; k < 16; ++k) inst[i].M[k] = transforms[16 * i + k];
    }
    Outputs out;
    out.radiance = radiance; out.direct = direct; out.indirect = indirect; out.albedo = albedo;
    out.normal = normal; out.depth = depth; out.alpha = alpha; out.tri_id = tri_id; out.counts = counts;
    out.instance_id = instance_id;
    out.normal_camera = normal_camera; out.mesh_id = mesh_id; out.material_id = material_id;
    render_scene(precision, inst.data(), n, *cam, lights, nl, *opt, out, nthreads, pixel_list, npix);
}

void orc_pixel_solid_angles(const float* kinv, float zdir, int W, int H, float* out) {
    pixel_solid_angles(kinv, zdir, W, H, out);
}

void orc_psf_kernel(const PSFO* p, int K, float* out) { psf_kernel(*p, K, out); }
void orc_denoise(const DenoiseO* o, int W, int H, float* direct, float* indirect, const float* albedo,
                 const float* depth, const float* normal, int literal) {
    denoise_eatwt(*o, W, H, direct, indirect, albedo, depth, normal, literal != 0);
}
void orc_render_stars(const float* rot, int W, int H, const float* eta, float area, int count, const int* sizes,
                      const float* const* kernels, float minimum_power, const float* intrinsic, int cull_behind,
                      const float* dirs, const float* irr, size_t n, float* img) {
    StarPSFO p{};
    p.count = count;
    for (int i = 0; i < count; ++i) {
        p.size[i] = sizes[i];
        p.kernel[i] = kernels[i];
    }
    p.minimum_power = minimum_power;
    for (int i = 0; i < 6; ++i) p.intrinsic[i] = intrinsic[i];
    p.cull_behind = cull_behind;
    render_stars(rot, W, H, eta, area, p, dirs, irr, n, img);
}
void orc_convolve(const float* in, int W, int H, const float* kern, int K, float* out) {
    convolve_spatial(in, W, H, kern, K, out);
}

// ---- fixed-visibility direct illumination, FP64 forward / adjoint / finite differences
void* orc_dem_diff_create(int W, int H, const Projection* proj, const float* heights, const float* grey) {
    return new DemDiff(dem_diff_make(W, H, *proj, heights, grey));
}
void orc_dem_diff_destroy(void* d) { delete static_cast<DemDiff*>(d); }

static DiScene di_scene(void* d, const uint32_t* rec, size_t npix, int spp, const Light* lights, int nl,
                        const Material* mat) {
    DiScene s;
    s.dd = static_cast<DemDiff*>(d);
    s.rec = reinterpret_cast<const DiRecord*>(rec);
    s.npix = npix;
    s.spp = spp;
    s.lights = lights;
    s.nl = nl;
    s.mat = mat;
    return s;
}

void orc_di_forward64(void* d, const double* h, const double* tex, const uint32_t* rec, size_t npix, int spp,
                      const Light* lights, int nl, const Material* mat, double* img, int nthreads) {
    di_forward64(di_scene(d, rec, npix, spp, lights, nl, mat), h, tex, img, nthreads);
}

void orc_di_adjoint64(void* d, const double* h, const double* tex, const uint32_t* rec, size_t npix, int spp,
                      const Light* lights, int nl, const Material* mat, const double* g, double* dh,
                      double* dtex) {
    di_adjoint64(di_scene(d, rec, npix, spp, lights, nl, mat), h, tex, g, dh, dtex);
}

void orc_di_fd64(void* d, const double* h, const double* tex, const uint32_t* rec, size_t npix, int spp,
                 const Light* lights, int nl, const Material* mat, const double* target, const int64_t* nodes,
                 size_t n, double step, double* out, int nthreads) {
    di_fd_heights64(di_scene(d, rec, npix, spp, lights, nl, mat), h, tex, target, nodes, n, step, out, nthreads);
}

void orc_light_sample(int precision, const Light* L, const double* p, uint32_t seed, uint32_t pix, uint32_t dim0,
                      size_t n, double* dirs, double* dist, double* pdf, double* rad) {
    hook_light_sample(precision, *L, p, seed, pix, dim0, n, dirs, dist, pdf, rad);
}
void orc_light_pdf(int precision, const Light* L, const double* p, const double* dirs, size_t n, double* out) {
    hook_light_pdf(precision, *L, p, dirs, n, out);
}
void orc_bounce_dirs(int precision, const double* N, uint32_t seed, uint32_t pix, uint32_t bounce, size_t n,
                     double* dirs, double* pdf) {
    hook_bounce_dirs(precision, N, seed, pix, bounce, n, dirs, pdf);
}
void orc_tex_bilinear(int precision, const Material* m, const double* uv, size_t n, double* out) {
    hook_tex_bilinear(precision, *m, uv, n, out);
}
float orc_kahan_sum_f32(const float* x, size_t n) { return hook_kahan_sum_f32(x, n); }

void orc_atan2_f32(const float* y, const float* x, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) out[i] = vgm::atan2_f32(y[i], x[i]);
}

// Test instrument: plant one known mistake in the oracle (0 = none).
void orc_plant(int k) { vo::g_plant = k; }

// sin and cos of the kernels' FP32 routine (fp_trig.hpp), for the tests that pin it.
void orc_sincos_f32(const float* x, size_t n, float* s, float* c) {
    for (size_t i = 0; i < n; ++i) vgm::sincos_f32(x[i], s[i], c[i]);
}

}  // extern "C"

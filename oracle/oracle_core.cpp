// oracle_core.cpp -- BVH build and integrator of the CPU oracle.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include "oracle_core.hpp"

#include "../vira_amd/csrc/fp_trig.hpp"

#include <atomic>
#include <thread>

namespace vo {

int g_plant = PLANT_NONE;

// ============================================================== BVH build
// ViraBLAS::subdivide / findBestSplitPlane (vira_blas.ipp:132-254).  Subtrees
// below the top levels are built on worker threads into private node arrays
// and spliced afterwards; the tree is identical to the sequential build.
namespace {

struct Builder {
    const Mesh& m;
    std::vector<double> tlo, thi, cen;  // per-triangle bounds and centroid (FP64)
    std::vector<uint32_t>& triIdx;
    Builder(const Mesh& mesh, std::vector<uint32_t>& ti) : m(mesh), triIdx(ti) {
        tlo.resize(3 * m.nt); thi.resize(3 * m.nt); cen.resize(3 * m.nt);
        for (uint32_t t = 0; t < (uint32_t)m.nt; ++t) bvh_tri_bounds(m, t, &tlo[3 * t], &thi[3 * t], &cen[3 * t]);
    }

    void bounds(BNode& n) const {  // updateNodeBounds (vira_blas.ipp:117-129)
        for (int k = 0; k < 3; ++k) { n.lo[k] = INFINITY; n.hi[k] = -INFINITY; }
        for (uint32_t i = 0; i < n.triCount; ++i) {
            uint32_t t = triIdx[n.leftFirst + i];
            for (int k = 0; k < 3; ++k) {
                n.lo[k] = std::min(n.lo[k], tlo[3 * t + k]);
                n.hi[k] = std::max(n.hi[k], thi[3 * t + k]);
            }
        }
    }

    static double area(const double lo[3], const double hi[3]) {  // aabb.ipp:34-38
        double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
        return ex * ey + ey * ez + ez * ex;
    }

    // findBestSplitPlane (vira_blas.ipp:185-254)
    double best_split(const BNode& n, int& axis, double& splitPos) const {
        double bestCost = std::numeric_limits<double>::max();
        int a = 0;
        double ext[3] = {n.hi[0] - n.lo[0], n.hi[1] - n.lo[1], n.hi[2] - n.lo[2]};
        if (ext[1] > ext[0]) a = 1;
        if (ext[2] > ext[a]) a = 2;
        double bmin = INFINITY, bmax = -INFINITY;
        for (uint32_t i = 0; i < n.triCount; ++i) {
            uint32_t t = triIdx[n.leftFirst + i];
            bmin = std::min(bmin, cen[3 * t + a]);
            bmax = std::max(bmax, cen[3 * t + a]);
        }
        constexpr int BINS = 8;
        struct Bin { double lo[3], hi[3]; uint32_t cnt; } bin[BINS];
        for (auto& b : bin) { for (int k = 0; k < 3; ++k) { b.lo[k] = INFINITY; b.hi[k] = -INFINITY; } b.cnt = 0; }
        double scale = BINS / (bmax - bmin);
        for (uint32_t i = 0; i < n.triCount; ++i) {
            uint32_t t = triIdx[n.leftFirst + i];
            double f = (cen[3 * t + a] - bmin) * scale;
            size_t bi = (f >= (double)BINS || !(f == f)) ? BINS - 1 : std::min<size_t>(BINS - 1, (size_t)f);
            bin[bi].cnt++;
            for (int k = 0; k < 3; ++k) {
                bin[bi].lo[k] = std::min(bin[bi].lo[k], tlo[3 * t + k]);
                bin[bi].hi[k] = std::max(bin[bi].hi[k], thi[3 * t + k]);
            }
        }
        double leftArea[BINS - 1], rightArea[BINS - 1];
        size_t leftCount[BINS - 1], rightCount[BINS - 1];
        double llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        double rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        size_t ls = 0, rs = 0;
        for (int i = 0; i < BINS - 1; ++i) {
            ls += bin[i].cnt;
            leftCount[i] = ls;
            if (!std::isinf(bin[i].lo[0]))
                for (int k = 0; k < 3; ++k) { llo[k] = std::min(llo[k], bin[i].lo[k]); lhi[k] = std::max(lhi[k], bin[i].hi[k]); }
            leftArea[i] = area(llo, lhi);
            rs += bin[BINS - 1 - i].cnt;
            rightCount[BINS - 2 - i] = rs;
            const Bin& rb = bin[BINS - 1 - i];
            if (!std::isinf(rb.lo[0]))
                for (int k = 0; k < 3; ++k) { rlo[k] = std::min(rlo[k], rb.lo[k]); rhi[k] = std::max(rhi[k], rb.hi[k]); }
            rightArea[BINS - 2 - i] = area(rlo, rhi);
        }
        scale = (bmax - bmin) / BINS;
        for (int i = 0; i < BINS - 1; ++i) {
            double cost = (double)leftCount[i] * leftArea[i] + (double)rightCount[i] * rightArea[i];
            if (cost < bestCost) {
                axis = a;
                splitPos = bmin + scale * (double)(i + 1);
                bestCost = cost;
            }
        }
        return bestCost;
    }

    // Builds the subtree rooted at local node 0 of `out` (already holding the
    // root with leftFirst/triCount/bounds set).  Child pairs are adjacent.
    void subdivide(std::vector<BNode>& out, uint32_t ni, int depth, int par_depth) {
        BNode n = out[ni];
        if (n.triCount <= 2) return;
        int axis = 0;
        double splitPos = 0;
        double cost = best_split(n, axis, splitPos);
        double ex = n.hi[0] - n.lo[0], ey = n.hi[1] - n.lo[1], ez = n.hi[2] - n.lo[2];
        double nodeCost = (double)n.triCount * (ex * ey + ey * ez + ez * ex);  // vira_blas.ipp:287-292
        if (cost >= nodeCost) return;
        // in-place partition (vira_blas.ipp:149-159)
        int64_t i = n.leftFirst, j = i + n.triCount - 1;
        while (i <= j) {
            if (cen[3 * triIdx[i] + axis] < splitPos) i++;
            else std::swap(triIdx[i], triIdx[j--]);
        }
        uint32_t leftCount = (uint32_t)(i - n.leftFirst);
        if (leftCount == 0 || leftCount == n.triCount) return;
        uint32_t l = (uint32_t)out.size();
        BNode L{}, R{};
        L.leftFirst = n.leftFirst; L.triCount = leftCount;
        R.leftFirst = (uint32_t)i; R.triCount = n.triCount - leftCount;
        bounds(L); bounds(R);
        out.push_back(L); out.push_back(R);
        out[ni].leftFirst = l;
        out[ni].triCount = 0;
        if (depth < par_depth) {
            // build both children as independent subtrees, the right one on a thread
            std::vector<BNode> lsub{out[l]}, rsub{out[l + 1]};
            std::thread th([&] { subdivide(rsub, 0, depth + 1, par_depth); });
            subdivide(lsub, 0, depth + 1, par_depth);
            th.join();
            splice(out, l, lsub);
            splice(out, l + 1, rsub);
        } else {
            subdivide(out, l, depth + 1, par_depth);
            subdivide(out, l + 1, depth + 1, par_depth);
        }
    }

    // Copies a subtree built in a private array (root at index 0) into `out`,
    // with its root replacing out[at].
    static void splice(std::vector<BNode>& out, uint32_t at, const std::vector<BNode>& sub) {
        uint32_t base = (uint32_t)out.size();
        // sub[0] -> out[at]; sub[k>0] -> out[base + k - 1]
        auto map = [&](uint32_t k) { return k == 0 ? at : base + k - 1; };
        out.resize(base + sub.size() - 1);
        for (uint32_t k = 0; k < sub.size(); ++k) {
            BNode c = sub[k];
            if (c.triCount == 0) c.leftFirst = map(c.leftFirst);
            out[map(k)] = c;
        }
    }
};

}  // namespace

BVH bvh_build(const Mesh& m, int nthreads) {
    BVH b;
    b.triIdx.resize(m.nt);
    for (uint32_t i = 0; i < (uint32_t)m.nt; ++i) b.triIdx[i] = i;
    if (m.nt == 0) return b;
    Builder bd(m, b.triIdx);
    b.nodes.reserve(2 * m.nt);
    BNode root{};
    root.leftFirst = 0;
    root.triCount = (uint32_t)m.nt;
    bd.bounds(root);
    b.nodes.push_back(root);
    int par_depth = 0;
    while ((1 << par_depth) < nthreads && par_depth < 6) par_depth++;
    if (m.nt < 100000) par_depth = 0;
    bd.subdivide(b.nodes, 0, 0, par_depth);
    b.flo.resize(3 * b.nodes.size());
    b.fhi.resize(3 * b.nodes.size());
    for (size_t n = 0; n < b.nodes.size(); ++n)
        for (int k = 0; k < 3; ++k) {
            b.flo[3 * n + k] = (float)b.nodes[n].lo[k];  // exact: bounds of FP32 positions
            b.fhi[3 * n + k] = (float)b.nodes[n].hi[k];
        }
    return b;
}

// ============================================================== integrator
namespace {

template <class R> struct Consts;
template <> struct Consts<float> {
    static constexpr float pi = 3.14159265358979323846f;
    static constexpr float inv_pi = 0.318309886183790671538f;
};
template <> struct Consts<double> {
    static constexpr double pi = 3.14159265358979323846;
    static constexpr double inv_pi = 0.318309886183790671538;
};

// offsetIntersection (rendering/ray.ipp:49-139): integer-ULP push along N.
inline V3<float> offset_point(V3<float> p, V3<float> n) {
    auto f2i = [](float v) { int32_t o; std::memcpy(&o, &v, 4); return o; };
    auto i2f = [](int32_t v) { float o; std::memcpy(&o, &v, 4); return o; };
    int32_t oi[3] = {(int32_t)(256.0f * n.x), (int32_t)(256.0f * n.y), (int32_t)(256.0f * n.z)};
    float pc[3] = {p.x, p.y, p.z}, nc[3] = {n.x, n.y, n.z}, r[3];
    for (int k = 0; k < 3; ++k) {
        float pi = i2f(f2i(pc[k]) + ((pc[k] < 0) ? -oi[k] : oi[k]));
        r[k] = (std::fabs(pc[k]) < (1.0f / 32.0f)) ? pc[k] + (1.0f / 65536.0f) * nc[k] : pi;
    }
    return {r[0], r[1], r[2]};
}
inline V3<double> offset_point(V3<double> p, V3<double> n) {
    auto f2i = [](double v) { int64_t o; std::memcpy(&o, &v, 8); return o; };
    auto i2f = [](int64_t v) { double o; std::memcpy(&o, &v, 8); return o; };
    int64_t oi[3] = {(int64_t)(65536.0 * n.x), (int64_t)(65536.0 * n.y), (int64_t)(65536.0 * n.z)};
    double pc[3] = {p.x, p.y, p.z}, nc[3] = {n.x, n.y, n.z}, r[3];
    for (int k = 0; k < 3; ++k) {
        double pi = i2f(f2i(pc[k]) + ((pc[k] < 0) ? -oi[k] : oi[k]));
        r[k] = (std::fabs(pc[k]) < (1.0 / 64.0)) ? pc[k] + (1.0 / 4294967296.0) * nc[k] : pi;
    }
    return {r[0], r[1], r[2]};
}

}  // namespace

V3<float> offset_point_f32(V3<float> p, V3<float> n) { return offset_point(p, n); }

namespace {

// sin and cos of the sampling angles: the FP32 twin uses the kernels' own
// routine (fp_trig.hpp, Cody-Waite reduction + fitted polynomials, pinned
// against FP64 sin/cos by the tests); the FP64 reference answer uses libm.
inline void sincos_r(float x, float& s, float& c) { vgm::sincos_f32(x, s, c); }
inline void sincos_r(double x, double& s, double& c) { s = std::sin(x); c = std::cos(x); }

template <class R>
inline R power_heuristic(R fPdf, R gPdf) {  // cpu_path_tracer.ipp:472-478
    R f = fPdf, g = gPdf;
    if (g_plant == PLANT_MIS_BALANCE) return f / (f + g);
    return (f * f) / (f * f + g * g);
}

// Image::sampleUVs + interpolatePixel (image.ipp:589-670), grey texture.
template <class R>
inline R tex_bilinear(const Material& mat, R u, R v) {
    if (g_plant == PLANT_TEX_TRANSPOSED) std::swap(u, v);
    R rx = (R)mat.tex_w, ry = (R)mat.tex_h;
    R px = std::fmod(u * rx, rx), py = std::fmod(v * ry, ry);
    if (px < 0) px += rx;
    if (py < 0) py += ry;
    if (!(px >= 0)) px = 0;  // non-finite uv: keep the fetch in range
    if (!(py >= 0)) py = 0;
    int x0 = (int)std::floor(px), y0 = (int)std::floor(py);
    if (g_plant == PLANT_TEX_OFF_BY_ONE) x0 = x0 + 1;
    x0 = std::min(x0, mat.tex_w - 1);
    y0 = std::min(y0, mat.tex_h - 1);
    int x1 = std::min(x0 + 1, mat.tex_w - 1), y1 = std::min(y0 + 1, mat.tex_h - 1);
    R p00 = mat.tex[x0 + (size_t)y0 * mat.tex_w], p10 = mat.tex[x1 + (size_t)y0 * mat.tex_w];
    R p01 = mat.tex[x0 + (size_t)y1 * mat.tex_w], p11 = mat.tex[x1 + (size_t)y1 * mat.tex_w];
    R dx = px - (R)x0, dy = py - (R)y0;
    return ((((R)1 - dx) * ((R)1 - dy)) * p00 + (dx * ((R)1 - dy)) * p10) + (((R)1 - dx) * dy) * p01 +
           (dx * dy) * p11;
}

// Image::sampleUVs + interpolatePixel (image.ipp:589-670) of a C-channel image
// (row-major, C floats per texel): the same texel choice and weights as
// tex_bilinear, per channel ((w00 p00 + w10 p10) + w01 p01) + w11 p11.
template <class R>
inline void image_bilinear(const float* img, int w, int h, int C, R u, R v, R* out) {
    R rx = (R)w, ry = (R)h;
    R px = std::fmod(u * rx, rx), py = std::fmod(v * ry, ry);
    if (px < 0) px += rx;
    if (py < 0) py += ry;
    if (!(px >= 0)) px = 0;
    if (!(py >= 0)) py = 0;
    int x0 = (int)std::floor(px), y0 = (int)std::floor(py);
    x0 = std::min(x0, w - 1);
    y0 = std::min(y0, h - 1);
    int x1 = std::min(x0 + 1, w - 1), y1 = std::min(y0 + 1, h - 1);
    R dx = px - (R)x0, dy = py - (R)y0;
    R w00 = ((R)1 - dx) * ((R)1 - dy), w10 = dx * ((R)1 - dy), w01 = ((R)1 - dx) * dy, w11 = dx * dy;
    const float* p00 = img + ((size_t)x0 + (size_t)y0 * w) * C;
    const float* p10 = img + ((size_t)x1 + (size_t)y0 * w) * C;
    const float* p01 = img + ((size_t)x0 + (size_t)y1 * w) * C;
    const float* p11 = img + ((size_t)x1 + (size_t)y1 * w) * C;
    for (int c = 0; c < C; ++c) out[c] = ((w00 * (R)p00[c] + w10 * (R)p10[c]) + w01 * (R)p01[c]) + w11 * (R)p11[c];
}

inline float atan2_r(float y, float x) { return vgm::atan2_f32(y, x); }
inline double atan2_r(double y, double x) { return std::atan2(y, x); }

// Background radiance seen along direction d (cpu_path_tracer.ipp:257-270):
// uv = (atan2(y, x) / 2 pi, atan2(sqrt(x^2 + y^2), z) / pi), then sampleUVs.
template <class R>
inline void background_radiance(const Options& opt, const V3<R>& d, R bg[NB]) {
    if (!opt.background) {
        for (int k = 0; k < NB; ++k) bg[k] = 0;
        return;
    }
    const R u = (g_plant == PLANT_BACKGROUND_SWAP ? atan2_r(d.x, d.y) : atan2_r(d.y, d.x)) / ((R)2 * Consts<R>::pi);
    const R v = atan2_r(std::sqrt(d.x * d.x + d.y * d.y), d.z) / Consts<R>::pi;
    image_bilinear<R>(opt.background, opt.bg_w, opt.bg_h, NB, u, v, bg);
}

// Light::sample (point_light.ipp:18-38, sphere_light.ipp:31-125).
template <class R>
inline bool sample_light(const Light& L, const V3<R>& p, uint32_t seed, uint32_t pix, uint32_t s,
                         uint32_t dim0, V3<R>& dir, R& distance, R& pdf, R rad[NB]) {
    V3<R> c{(R)L.pos[0], (R)L.pos[1], (R)L.pos[2]};
    V3<R> toL = c - p;
    distance = length(toL);
    R dInv = (R)1 / distance;
    if (L.kind == 0) {
        R dSq = distance * distance;
        pdf = 1;
        dir = dInv * toL;
        for (int k = 0; k < NB; ++k) rad[k] = (R)L.value[k] / dSq;
        return true;
    }
    R radius = (R)L.radius;
    if (distance <= radius) {  // inside the sphere: sphere_light.ipp:40-57
        pdf = (R)1 / ((R)4 * Consts<R>::pi * radius * radius);
        if (distance < (R)1e-3) dir = {1, 0, 0};
        else dir = dInv * toL;
        distance = radius;
        for (int k = 0; k < NB; ++k) rad[k] = (R)L.value[k];
        return true;
    }
    R r1 = (R)rng_u01(seed, pix, s, dim0), r2 = (R)rng_u01(seed, pix, s, dim0 + 1);
    V3<R> W = dInv * toL;
    V3<R> N = (W.x < W.y) ? V3<R>{1, 0, 0} : V3<R>{0, 1, 0};
    V3<R> Vv = normalize(cross(W, N));
    V3<R> U = normalize(cross(W, Vv));
    R phi = (R)2 * Consts<R>::pi * r1;
    R zeta = r2;
    // cos(theta) = (1 - zeta) + zeta cos(theta_max); 1 - cos(theta_max) as
    // x / (1 + sqrt(1 - x)), x = (r/d)^2 -- algebraically the reference's
    // 1 - sqrt(d^2 - r^2)/d, without its FP32 cancellation for a sun at 1 AU
    R cc = radius / distance;
    R x = cc * cc;
    R q = x / ((R)1 + std::sqrt((R)1 - x));
    R t = g_plant == PLANT_CONE_SQUARE ? (zeta * zeta) * q : zeta * q;
    R cosT = (R)1 - t, sinT = std::sqrt(t * ((R)2 - t));
    R sinP, cosP;
    sincos_r(phi, sinP, cosP);
    R a = cosP * sinT, bb = cosT, cz = sinP * sinT;
    dir = {(a * U.x + bb * W.x) + cz * Vv.x, (a * U.y + bb * W.y) + cz * Vv.y, (a * U.z + bb * W.z) + cz * Vv.z};
    R Omega = (g_plant == PLANT_CONE_OMEGA ? Consts<R>::pi : (R)2 * Consts<R>::pi) * q;
    pdf = (R)1 / Omega;
    for (int k = 0; k < NB; ++k) rad[k] = (R)L.value[k];
    return true;
}

// Light::getPDF (point_light.ipp:40-57, sphere_light.ipp:127-153).
template <class R>
inline R light_pdf(const Light& L, const V3<R>& p, const V3<R>& dir) {
    V3<R> c{(R)L.pos[0], (R)L.pos[1], (R)L.pos[2]};
    if (L.kind == 0) {
        V3<R> toL = normalize(c - p);
        V3<R> q = normalize(dir);
        R d = dot(toL, q);
        return (d > (R)1 - (R)1e-6) ? (R)1 : (R)0;
    }
    V3<R> toC = c - p;
    R dc = length(toC);
    R radius = (R)L.radius;
    if (dc <= radius) return (R)1 / ((R)4 * Consts<R>::pi * radius * radius);
    V3<R> tn = toC * ((R)1 / dc);
    R cc = radius / dc;
    R x = cc * cc;
    R q = x / ((R)1 + std::sqrt((R)1 - x));
    R cosMax = (R)1 - q;
    R ca = dot(dir, tn);
    if (ca < cosMax) return 0;
    R Omega = (g_plant == PLANT_CONE_OMEGA ? Consts<R>::pi : (R)2 * Consts<R>::pi) * q;
    return (R)1 / Omega;
}

// Shading frame of processIntersection (cpu_path_tracer.ipp:313-324).
template <class R>
inline void tangent_frame(const V3<R>& N, V3<R>& tg, V3<R>& bt) {
    V3<R> arb = (std::fabs(N.x) < (R)0.9) ? V3<R>{1, 0, 0} : V3<R>{0, 1, 0};
    tg = normalize(cross(arb, N));
    bt = cross(N, tg);
}

// Material::sampleDirection for the Lambertian's COSINE_WEIGHTED strategy
// (material.ipp:8-33, material_sampling.hpp:15-21, sampling.ipp:53-84):
// Shirley-Chiu concentric map of (u1, u2) to the disk, lifted to the
// hemisphere, into the frame [tg, bt, N]; pdf = cos(theta) / pi.  Returns
// false where the pdf is 0 (the path then ends).
template <class R>
inline bool cosine_sample(uint32_t seed, uint32_t pix, uint32_t s, uint32_t bounce, const V3<R>& N, const V3<R>& tg,
                          const V3<R>& bt, V3<R>& nd, R& mpdf) {
    const uint32_t db = bounce << 8;
    R u1 = (R)rng_u01(seed, pix, s, db | SLOT_BSDF_U);
    R u2 = (R)rng_u01(seed, pix, s, db | SLOT_BSDF_V);
    R ox = (R)2 * u1 - (R)1, oy = (R)2 * u2 - (R)1;
    R dx = 0, dy = 0;
    if (!(std::fabs(ox) < (R)1e-6 && std::fabs(oy) < (R)1e-6)) {
        R r, th;
        if (std::fabs(ox) > std::fabs(oy)) { r = ox; th = (Consts<R>::pi / (R)4) * (oy / ox); }
        else { r = oy; th = (Consts<R>::pi / (R)2) - (Consts<R>::pi / (R)4) * (ox / oy); }
        R sn, cs;
        sincos_r(th, sn, cs);
        dx = r * cs;
        dy = r * sn;
    }
    R lz = std::sqrt(std::max((R)0, ((R)1 - dx * dx) - dy * dy));
    mpdf = lz * Consts<R>::inv_pi;
    if (g_plant == PLANT_BOUNCE_PDF && mpdf > (R)0) mpdf = Consts<R>::inv_pi / (R)2;
    if (!(mpdf > (R)0)) return false;
    nd = V3<R>{(tg.x * dx + bt.x * dy) + N.x * lz, (tg.y * dx + bt.y * dy) + N.y * lz,
               (tg.z * dx + bt.z * dy) + N.z * lz};
    return true;
}

template <class R>
struct PixelAcc {
    Kahan<R> rad[NB], dir[NB];
    R alb[NB] = {0};
    V3<R> nrm{0, 0, 0};
    R depth = 0;
    uint32_t count = 0;
    uint32_t tri = 0xffffffffu;
    uint32_t mesh = 0xffffffffu, material = 0xffffffffu;
    uint32_t inst = 0xffffffffu;
};

// Camera-relative placement of an instance: the FP64 model matrix minus the
// camera position, its inverse and normal matrix, rounded to R (the same FP64
// operations as the GPU's rebase kernel).
template <class R>
struct Placement {
    R l2w[12], w2l[12], nm[9];
};

template <class R>
Placement<R> rebase_instance(const double* M, const double c[3]) {
    double A[3][3];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) A[r][k] = M[k * 4 + r];
    const double t[3] = {M[12] - c[0], M[13] - c[1], M[14] - c[2]};
    const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    const double det = (A[0][0] * c00 + A[0][1] * c01) + A[0][2] * c02;
    const double id = 1.0 / det;
    double inv[3][3];
    inv[0][0] = c00 * id;
    inv[1][0] = c01 * id;
    inv[2][0] = c02 * id;
    inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
    inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
    inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
    inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
    inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
    inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
    Placement<R> P;
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) {
            P.l2w[r * 4 + k] = (R)A[r][k];
            P.w2l[r * 4 + k] = (R)inv[r][k];
            P.nm[r * 3 + k] = g_plant == PLANT_NORMAL_MATRIX ? (R)inv[r][k] : (R)inv[k][r];
        }
        P.l2w[r * 4 + 3] = (R)t[r];
        P.w2l[r * 4 + 3] = (R)(-((inv[r][0] * t[0] + inv[r][1] * t[1]) + inv[r][2] * t[2]));
    }
    return P;
}

template <class R>
inline V3<R> xform_point(const R* m, const V3<R>& p) {
    return {((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3], ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7],
            ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11]};
}
template <class R>
inline V3<R> xform_dir(const R* m, const V3<R>& d) {
    return {(m[0] * d.x + m[1] * d.y) + m[2] * d.z, (m[4] * d.x + m[5] * d.y) + m[6] * d.z,
            (m[8] * d.x + m[9] * d.y) + m[10] * d.z};
}
template <class R>
inline V3<R> xform_normal(const R* n, const V3<R>& v) {
    return {(n[0] * v.x + n[1] * v.y) + n[2] * v.z, (n[3] * v.x + n[4] * v.y) + n[5] * v.z,
            (n[6] * v.x + n[7] * v.y) + n[8] * v.z};
}

template <class R>
void render_impl(const InstanceO* inst, int ninst, const Camera& cam_in, const Light* lights_in, int nl,
                 const Options& opt, Outputs& out, int nthreads, const int* pixel_list, size_t npix) {
    const size_t total = pixel_list ? npix : (size_t)cam_in.W * cam_in.H;
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> g_cnt[7];
    for (auto& c : g_cnt) c = 0;
    static const double kIdentity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const bool rebased = !(ninst == 1 && std::memcmp(inst[0].M, kIdentity, sizeof(kIdentity)) == 0);
    // camera-relative frame for instanced scenes: camera at the origin, lights moved in FP64
    Camera cam = cam_in;
    std::vector<Light> lights(lights_in, lights_in + nl);
    std::vector<Placement<R>> place(ninst);
    if (rebased) {
        for (int i = 0; i < ninst; ++i) place[i] = rebase_instance<R>(inst[i].M, cam_in.pos);
        for (auto& L : lights)
            for (int k = 0; k < 3; ++k) L.pos[k] = L.pos[k] - cam_in.pos[k];
        cam.pos[0] = cam.pos[1] = cam.pos[2] = 0.0;
    }
    // closest hit over the instances (ties keep the lower instance), or occlusion
    auto trace = [&](const RayT<R>& ray, bool any, R tmax, TravCount* tc, uint32_t& hi) {
        Hit<R> best{(R)INFINITY, 0xffffffffu, 0, 0};
        hi = 0;
        for (int i = 0; i < ninst; ++i) {
            RayT<R> r = ray;
            if (rebased) r = RayT<R>{xform_point(place[i].w2l, ray.o), xform_dir(place[i].w2l, ray.d)};
            Hit<R> h = bvh_trace<R>(*inst[i].m, *inst[i].b, r, any, tmax, tc);
            if (h.id == 0xffffffffu) continue;
            if (best.id == 0xffffffffu || h.t < best.t) { best = h; hi = (uint32_t)i; }
            if (any) break;
        }
        return best;
    };

    // adaptive-sampling settings: 0 stands for the reference's defaults (cpu_path_tracer.hpp:39-42)
    const int batch_n = opt.samples_per_batch > 0 ? opt.samples_per_batch : 30;
    const int miss_n = opt.samples_to_detect_miss > 0 ? opt.samples_to_detect_miss : 30;
    const float tol = opt.sampling_tolerance > 0.0f ? opt.sampling_tolerance : 0.05f;
    auto worker = [&]() {
        uint64_t cnt[7] = {0};
        TravCount tc_closest, tc_any;
        while (true) {
            size_t k0 = next.fetch_add(64);
            if (k0 >= total) break;
            size_t k1 = std::min(total, k0 + 64);
            for (size_t kk = k0; kk < k1; ++kk) {
                int pix = pixel_list ? pixel_list[kk] : (int)kk;
                int pi = pix % cam.W, pj = pix / cam.W;
                PixelAcc<R> acc;
                R s1 = 0, s2 = 0;   // adaptive sampling sums (cpu_path_tracer.ipp:198-227)
                int nsamp = opt.spp;
                for (int s = 0; s < opt.spp; ++s) {
                    // ---- camera ray (cpu_path_tracer.ipp:178-192, camera.ipp:1126-1163,1738-1749)
                    R X = (R)pi, Y = (R)pj;
                    if (s > 0) {
                        X = X + (R)rng_u01(opt.seed, (uint32_t)pix, (uint32_t)s, SLOT_JITTER_X);
                        Y = Y + (R)rng_u01(opt.seed, (uint32_t)pix, (uint32_t)s, SLOT_JITTER_Y);
                    }
                    R hx = ((R)cam.kinv[0] * X + (R)cam.kinv[2] * Y) + (R)cam.kinv[4];
                    R hy = ((R)cam.kinv[1] * X + (R)cam.kinv[3] * Y) + (R)cam.kinv[5];
                    V3<R> dl{hx, (R)cam.zdir * hy, (R)cam.zdir};
                    const float* Rm = cam.rot;
                    V3<R> dg{((R)Rm[0] * dl.x + (R)Rm[1] * dl.y) + (R)Rm[2] * dl.z,
                             ((R)Rm[3] * dl.x + (R)Rm[4] * dl.y) + (R)Rm[5] * dl.z,
                             ((R)Rm[6] * dl.x + (R)Rm[7] * dl.y) + (R)Rm[8] * dl.z};
                    RayT<R> ray{{(R)cam.pos[0], (R)cam.pos[1], (R)cam.pos[2]}, normalize(dg)};
                    cnt[0]++;
                    R beta[NB];
                    for (int k = 0; k < NB; ++k) beta[k] = 1;
                    R path[NB] = {0}, direct[NB] = {0};
                    uint32_t* rec = out.records ? out.records + 4 * ((size_t)s * cam.W * cam.H + (size_t)pix)
                                                : nullptr;
                    if (rec) { rec[0] = 0xffffffffu; rec[1] = rec[2] = rec[3] = 0; }
                    for (int bounce = 0; bounce <= opt.bounces; ++bounce) {
                        uint32_t hinst = 0;
                        Hit<R> h = trace(ray, false, (R)INFINITY, &tc_closest, hinst);
                        if (h.id == 0xffffffffu) {
                            // miss: background radiance when shown (cpu_path_tracer.ipp:257-270)
                            if (opt.show_background) {
                                R bg[NB];
                                background_radiance<R>(opt, ray.d, bg);
                                for (int k = 0; k < NB; ++k) path[k] = path[k] + beta[k] * (bg[k] * (R)1);
                            }
                            break;
                        }
                        const Mesh& m = *inst[hinst].m;
                        const Material& mat = inst[hinst].mat;
                        auto vpos = [&](uint32_t v) { return mesh_pos<R>(m, v); };
                        auto vnrm = [&](uint32_t v) {
                            return V3<R>{(R)m.nrm[3 * v], (R)m.nrm[3 * v + 1], (R)m.nrm[3 * v + 2]};
                        };
                        if (rec && bounce == 0) {
                            float fb1 = (float)h.b1, fb2 = (float)h.b2;
                            rec[0] = h.id;
                            std::memcpy(&rec[1], &fb1, 4);
                            std::memcpy(&rec[2], &fb2, 4);
                        }
                        // ---- processIntersection (cpu_path_tracer.ipp:279-469)
                        uint32_t i0 = m.idx[3 * h.id], i1 = m.idx[3 * h.id + 1], i2 = m.idx[3 * h.id + 2];
                        V3<R> p0 = vpos(i0), p1 = vpos(i1), p2 = vpos(i2);
                        R w0 = h.b1, w1 = h.b2, w2 = ((R)1 - h.b1) - h.b2;
                        R uvu = 0, uvv = 0;
                        if (!m.uv.empty()) {
                            uvu = (w0 * (R)m.uv[2 * i1] + w1 * (R)m.uv[2 * i2]) + w2 * (R)m.uv[2 * i0];
                            uvv = (w0 * (R)m.uv[2 * i1 + 1] + w1 * (R)m.uv[2 * i2 + 1]) + w2 * (R)m.uv[2 * i0 + 1];
                        }
                        // FP32 twin: the mesh's stored face normal (FP64-derived, as the kernels read it);
                        // FP64: from the FP64 positions directly
                        V3<R> fn;
                        if constexpr (std::is_same<R, float>::value)
                            fn = V3<R>{m.fn[4 * h.id], m.fn[4 * h.id + 1], m.fn[4 * h.id + 2]};
                        else
                            fn = normalize(cross(p1 - p0, p2 - p0));
                        V3<R> n0 = vnrm(i0), n1 = vnrm(i1), n2 = vnrm(i2);
                        V3<R> Nl = fn;
                        if (mat.smooth) Nl = normalize((w0 * n1 + w1 * n2) + w2 * n0);
                        // instances: normal matrix on the normals (not renormalised, cpu_path_tracer.ipp:310)
                        const V3<R> N_frame = rebased ? xform_normal(place[hinst].nm, Nl) : Nl;
                        const V3<R> fng = rebased ? xform_normal(place[hinst].nm, fn) : fn;
                        V3<R> tg, bt;
                        tangent_frame(N_frame, tg, bt);
                        // Material::getNormal (material.ipp:129-143): tangent-space map through [tg bt N]
                        V3<R> N = N_frame;
                        if (mat.nmap) {
                            R nt[3];
                            image_bilinear<R>(mat.nmap, mat.nmap_w, mat.nmap_h, 3, uvu, uvv, nt);
                            const V3<R> mapped{(tg.x * nt[0] + bt.x * nt[1]) + N_frame.x * nt[2],
                                               (tg.y * nt[0] + bt.y * nt[1]) + N_frame.y * nt[2],
                                               (tg.z * nt[0] + bt.z * nt[1]) + N_frame.z * nt[2]};
                            N = g_plant == PLANT_NMAP_UNNORMALISED ? mapped : normalize(mapped);
                        }
                        // albedo: vertex albedo x Material::getAlbedo (material.ipp:114-116)
                        R mvs[NB];
                        if (mat.tex && mat.tex_channels == 8) {
                            image_bilinear<R>(mat.tex, mat.tex_w, mat.tex_h, NB, uvu, uvv, mvs);
                        } else {
                            R texv = mat.tex ? tex_bilinear<R>(mat, uvu, uvv) : (R)1;
                            for (int k = 0; k < NB; ++k) mvs[k] = mat.tex ? texv * (R)mat.albedo[k] : (R)mat.albedo[k];
                        }
                        R alb[NB];
                        for (int k = 0; k < NB; ++k) {
                            R av = (w0 * (R)m.alb[NB * i1 + k] + w1 * (R)m.alb[NB * i2 + k]) + w2 * (R)m.alb[NB * i0 + k];
                            alb[k] = av * mvs[k];
                        }
                        V3<R> P = (w0 * p1 + w1 * p2) + w2 * p0;
                        {  // computeShadingPoint (cpu_path_tracer.ipp:481-511)
                            V3<R> q0 = P - dot(P - p1, n1) * n1;
                            V3<R> q1 = P - dot(P - p2, n2) * n2;
                            V3<R> q2 = P - dot(P - p0, n0) * n0;
                            V3<R> sp = (q0 * w0 + q1 * w1) + q2 * w2;
                            V3<R> diff = sp - P;
                            if (dot(diff, Nl) > (R)0) P = sp;
                        }
                        if (rebased) P = xform_point(place[hinst].l2w, P);  // model matrix (cpu_path_tracer.ipp:344)
                        if (bounce == 0) {
                            if (acc.tri == 0xffffffffu) {
                                acc.tri = h.id;
                                acc.inst = hinst;
                                acc.mesh = m.id;
                                acc.material = mat.material_id;
                            }
                            for (int k = 0; k < NB; ++k) acc.alb[k] = acc.alb[k] + alb[k];
                            acc.nrm = acc.nrm + N;
                            acc.depth = acc.depth + h.t;
                            acc.count++;
                        }
                        V3<R> org = offset_point(P, fng);
                        V3<R> V = {-ray.d.x, -ray.d.y, -ray.d.z};
                        (void)V;
                        R vrad[NB] = {0};
                        if (bounce == 0 && opt.has_ambient) {  // cpu_path_tracer.ipp:389-396, material.ipp:67
                            for (int k = 0; k < NB; ++k)
                                vrad[k] = vrad[k] + beta[k] * ((R)opt.ambient[k] *
                                                               (g_plant == PLANT_AMBIENT_NO_ALBEDO ? (R)1 : alb[k]));
                        }
                        for (int l = 0; l < nl; ++l) {
                            V3<R> ld;
                            R dist, lpdf, lrad[NB];
                            sample_light<R>(lights[l], org, opt.seed, (uint32_t)pix, (uint32_t)s,
                                            ((uint32_t)bounce << 8) | (SLOT_LIGHT0 + 2 * l), ld, dist, lpdf, lrad);
                            if (!(lpdf > 0)) continue;
                            cnt[1]++;
                            RayT<R> sray{org, ld};
                            uint32_t si = 0;
                            Hit<R> sh = trace(sray, true, dist, &tc_any, si);
                            if (sh.id != 0xffffffffu) continue;  // occluded
                            if (rec && bounce == 0) rec[3] |= 1u << l;
                            R ndl = dot(N, ld);
                            R mx = std::max(ndl, (R)0);
                            R cth = g_plant == PLANT_DROP_COSINE ? (R)1 : std::max((R)0, dot(ld, N));
                            R mpdf = (ndl <= (R)0) ? (R)0 : ndl * Consts<R>::inv_pi;
                            R wgt = power_heuristic(lpdf, mpdf);
                            for (int k = 0; k < NB; ++k) {
                                R bsdf = (alb[k] * Consts<R>::inv_pi) * mx;
                                vrad[k] = vrad[k] + ((((beta[k] * bsdf) * lrad[k]) * cth) * wgt) / lpdf;
                            }
                        }
                        for (int k = 0; k < NB; ++k) path[k] = path[k] + vrad[k];
                        if (bounce == 0)
                            for (int k = 0; k < NB; ++k) direct[k] = direct[k] + vrad[k];
                        if (bounce >= opt.bounces) break;
                        // ---- BSDF sampling (material.ipp:8-33, material_sampling.hpp:15-30, sampling.ipp:53-84)
                        uint32_t db = (uint32_t)bounce << 8;
                        V3<R> nd;
                        R mpdf;
                        if (!cosine_sample(opt.seed, (uint32_t)pix, (uint32_t)s, (uint32_t)bounce, N_frame, tg, bt, nd, mpdf))
                            break;
                        R ndd = dot(N, nd);
                        R mx = std::max(ndd, (R)0);
                        R cth = std::max((R)0, dot(nd, N));
                        R lp = 0;
                        for (int l = 0; l < nl; ++l) lp = lp + light_pdf<R>(lights[l], org, nd);
                        if (nl > 0) lp = lp / (R)nl;
                        R wgt = power_heuristic(mpdf, lp);
                        R qmax = 0;
                        for (int k = 0; k < NB; ++k) {
                            R bsdf = (alb[k] * Consts<R>::inv_pi) * mx;
                            beta[k] = beta[k] * (((bsdf * cth) * wgt) / mpdf);
                            qmax = std::max(qmax, beta[k]);
                        }
                        if (opt.rr && bounce + 1 >= opt.rr_start) {
                            R q = std::min((R)1, qmax);
                            R ur = (R)rng_u01(opt.seed, (uint32_t)pix, (uint32_t)s, db | SLOT_RR);
                            if (ur >= q) break;
                            if (g_plant != PLANT_RR_NO_REWEIGHT)
                                for (int k = 0; k < NB; ++k) beta[k] = beta[k] / q;
                        }
                        ray = RayT<R>{org, nd};
                        cnt[2]++;
                    }
                    for (int k = 0; k < NB; ++k) { acc.rad[k].add(path[k]); acc.dir[k].add(direct[k]); }
                    if (opt.adaptive) {
                        // cpu_path_tracer.ipp:198-227; a stop at sample s leaves s as the divisor
                        R sq = 0;
                        for (int k = 0; k < NB; ++k) sq = sq + path[k] * path[k];
                        const R illum = std::sqrt(sq);
                        s1 = s1 + illum;
                        s2 = s2 + illum * illum;
                        bool stop = false;
                        if (s % miss_n == 0 && s != 0 && acc.depth == (R)0) stop = true;
                        if (!stop && s % batch_n == 0 && s != 0) {
                            if (s1 == (R)0) {
                                stop = true;
                            } else {
                                const R n = (R)s + (R)1;
                                const R mu = s1 / n;
                                const R var = ((R)1 / (n - (R)1)) * (s2 - (s1 * s1) / n);
                                const R I = (std::is_same<R, float>::value ? (R)1.96f : (R)1.96) * std::sqrt(var / n);
                                if (I <= (R)tol * mu) stop = true;
                            }
                        }
                        if (stop) {
                            nsamp = g_plant == PLANT_ADAPTIVE_DIVISOR ? s + 1 : s;
                            break;
                        }
                    }
                }
                // ---- averaging and passes (cpu_path_tracer.ipp:229-239, passes.ipp:81-127)
                R sf = (R)nsamp;
                size_t o = (size_t)pix;
                for (int k = 0; k < NB; ++k) {
                    R rad = acc.rad[k].s / sf, dr = acc.dir[k].s / sf;
                    if (out.radiance) out.radiance[NB * o + k] = (float)rad;
                    if (out.direct) out.direct[NB * o + k] = (float)dr;
                    if (out.indirect) out.indirect[NB * o + k] = (float)(rad - dr);
                    if (out.albedo) out.albedo[NB * o + k] = (float)(acc.alb[k] / sf);
                }
                {
                    V3<R> n = acc.nrm;
                    if (acc.count > 0) n = normalize(n);
                    if (out.normal) {
                        out.normal[3 * o] = (float)n.x; out.normal[3 * o + 1] = (float)n.y; out.normal[3 * o + 2] = (float)n.z;
                    }
                    if (out.normal_camera) {  // normalize(getViewNormalMatrix() * normal_global) (cpu_path_tracer.ipp:237-239)
                        V3<R> c{0, 0, 0};
                        if (acc.count > 0) {
                            const float* vn = cam_in.view_normal;
                            float t[9];
                            bool given = false;
                            for (int k = 0; k < 9; ++k) given = given || vn[k] != 0.0f;
                            for (int r = 0; r < 3; ++r)
                                for (int q = 0; q < 3; ++q) t[r * 3 + q] = given ? vn[r * 3 + q] : cam_in.rot[q * 3 + r];
                            c = normalize(V3<R>{((R)t[0] * n.x + (R)t[1] * n.y) + (R)t[2] * n.z,
                                                ((R)t[3] * n.x + (R)t[4] * n.y) + (R)t[5] * n.z,
                                                ((R)t[6] * n.x + (R)t[7] * n.y) + (R)t[8] * n.z});
                        }
                        out.normal_camera[3 * o] = (float)c.x; out.normal_camera[3 * o + 1] = (float)c.y;
                        out.normal_camera[3 * o + 2] = (float)c.z;
                    }
                }
                if (out.depth) out.depth[o] = acc.count ? (float)(acc.depth / (R)acc.count) : INFINITY;
                if (out.alpha) out.alpha[o] = (float)acc.count / (float)nsamp;
                if (out.mesh_id) out.mesh_id[o] = acc.tri == 0xffffffffu ? 0xffffffffu : acc.mesh;
                if (out.material_id) out.material_id[o] = acc.tri == 0xffffffffu ? 0xffffffffu : acc.material;
                if (out.tri_id) out.tri_id[o] = acc.tri;
                if (out.instance_id) out.instance_id[o] = acc.tri == 0xffffffffu ? acc.tri : acc.inst;
            }
        }
        cnt[3] = tc_closest.internal; cnt[4] = tc_closest.tris;
        cnt[5] = tc_any.internal; cnt[6] = tc_any.tris;
        for (int k = 0; k < 7; ++k) g_cnt[k] += cnt[k];
    };
    int nt = std::max(1, nthreads);
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
    if (out.counts)
        for (int k = 0; k < 7; ++k) out.counts[k] = g_cnt[k];
}

}  // namespace

void pixel_solid_angles(const float kinv[6], float zdir, int W, int H, float* out) {
    using D = V3<double>;
    const double k[6] = {kinv[0], kinv[1], kinv[2], kinv[3], kinv[4], kinv[5]};
    const double zd = zdir;
    auto corner = [&](double x, double y) {
        const double hx = (k[0] * x + k[2] * y) + k[4];
        const double hy = (k[1] * x + k[3] * y) + k[5];
        return normalize(D{hx, zd * hy, zd});
    };
    auto tangent = [](const D& p0, const D& p1) { return normalize(cross(cross(p0, p1 - p0), p0)); };
    auto tri = [&](const D& c0, const D& c1, const D& c2) {
        const double a0 = std::acos(dot(tangent(c0, c1), tangent(c0, c2)));
        const double a1 = std::acos(dot(tangent(c1, c0), tangent(c1, c2)));
        const double a2 = std::acos(dot(tangent(c2, c0), tangent(c2, c1)));
        return ((a0 + a1) + a2) - 3.14159265358979323846;
    };
    for (int j = 0; j < H; ++j)
        for (int i = 0; i < W; ++i) {
            const D c0 = corner(i, j), c1 = corner(i + 1, j), c2 = corner(i + 1, j + 1), c3 = corner(i, j + 1);
            out[(size_t)j * W + i] = (float)(tri(c0, c1, c2) + tri(c0, c2, c3));
        }
}

void psf_kernel(const PSFO& P, int K, float* out) {
    const int S = P.supersample;
    const float center = ((float)K - 1.0f) / 2.0f;
    const float ang = (3.14159265358979323846f / 180.0f) * P.angle_deg;
    const float cr = std::cos(ang), sr = std::sin(ang);
    auto eval = [&](float px, float py, float v[NB]) {
        if (P.kind == 0) {
            const float rx = cr * px + sr * py, ry = -sr * px + cr * py;
            for (int k = 0; k < NB; ++k) {
                const float xn = rx / P.sigma_x[k], yn = ry / P.sigma_y[k];
                const float ex = -0.5f * ((xn * xn) + (yn * yn));
                const float coef = 1.0f / (((2.0f * 3.14159265358979323846f) * P.sigma_x[k]) * P.sigma_y[k]);
                v[k] = coef * std::exp(ex);
            }
            return;
        }
        const float qx = px * P.pixel_size[0], qy = py * P.pixel_size[1];
        const double radius = (double)std::sqrt(qx * qx + qy * qy);
        if (radius < 1e-20) {
            for (int k = 0; k < NB; ++k) v[k] = 1.0f;
            return;
        }
        const double coefficient = 3.141592653589793 * (double)(P.aperture_diameter / P.focal_length);
        for (int k = 0; k < NB; ++k) {
            const double x = coefficient * radius / (double)P.wavelength[k];
            const float a = (float)(2.0 * ::j1(x) / x);
            v[k] = a * a;
        }
    };
    for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) {
            float acc[NB] = {0};
            for (int si = 0; si < S; ++si)
                for (int sj = 0; sj < S; ++sj) {
                    const float ox = ((float)si + 0.5f) / (float)S, oy = ((float)sj + 0.5f) / (float)S;
                    const float spx = ((float)i + ox) - 0.5f, spy = ((float)j + oy) - 0.5f;
                    float v[NB];
                    eval(spx - center, spy - center, v);
                    for (int k = 0; k < NB; ++k) acc[k] = acc[k] + v[k];
                }
            for (int k = 0; k < NB; ++k) out[((size_t)i + (size_t)j * K) * NB + k] = acc[k] / (float)(S * S);
        }
}

void convolve_spatial(const float* in, int W, int H, const float* kern, int K, float* out) {
    const int h = K / 2;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float sum[NB] = {0};
            for (int ky = 0; ky < K; ++ky) {
                const int iy = y + ky - h;
                if (iy < 0 || iy >= H) continue;
                for (int kx = 0; kx < K; ++kx) {
                    const int ix = x + kx - h;
                    if (ix < 0 || ix >= W) continue;
                    const float* src = in + ((size_t)iy * W + ix) * NB;
                    const float* kv = kern + ((size_t)ky * K + kx) * NB;
                    for (int k = 0; k < NB; ++k) sum[k] = sum[k] + src[k] * kv[k];
                }
            }
            for (int k = 0; k < NB; ++k) out[((size_t)y * W + x) * NB + k] = sum[k];
        }
}

static float magnitude_f(const float* v) {
    float q = 0.0f;
    for (int k = 0; k < NB; ++k) q += v[k] * v[k];
    return std::sqrt(q);
}

void render_stars(const float rot[9], int W, int H, const float eta[NB], float area, const StarPSFO& psf,
                  const float* dirs, const float* irr, size_t n, float* img) {
    // initKernels: the border pixel of largest magnitude, kernel(0, x),
    // kernel(size-1, x), then kernel(y, 0), kernel(y, size-1); std::max keeps
    // the first of equal magnitudes.
    float edge[4][NB];
    for (int i = 0; i < psf.count; ++i) {
        const int K = psf.size[i];
        const float* kern = psf.kernel[i];
        auto at = [&](int a, int b) { return kern + ((size_t)a + (size_t)b * K) * NB; };
        float best[NB] = {0};
        auto consider = [&](const float* v) {
            if (magnitude_f(best) < magnitude_f(v))
                for (int k = 0; k < NB; ++k) best[k] = v[k];
        };
        for (int x = 0; x < K; ++x) {
            consider(at(0, x));
            consider(at(K - 1, x));
        }
        for (int y = 1; y < K - 1; ++y) {
            consider(at(y, 0));
            consider(at(y, K - 1));
        }
        for (int k = 0; k < NB; ++k) edge[i][k] = best[k];
    }
    // camera-frame vectors (cpu_unresolved_renderer.ipp:81-126)
    std::vector<float> v(3 * n);
    for (size_t s = 0; s < n; ++s) {
        const float* d = dirs + 3 * s;
        v[3 * s] = (rot[0] * d[0] + rot[3] * d[1]) + rot[6] * d[2];
        v[3 * s + 1] = (rot[1] * d[0] + rot[4] * d[1]) + rot[7] * d[2];
        v[3 * s + 2] = (rot[2] * d[0] + rot[5] * d[1]) + rot[8] * d[2];
    }
    const float* Km = psf.intrinsic;
    const int tile = 512;
    for (int ty = 0; ty < (H + tile - 1) / tile; ++ty)
        for (int tx = 0; tx < (W + tile - 1) / tile; ++tx) {
            const int rx0 = tx * tile, ry0 = ty * tile;
            const int rx1 = std::min(W, (tx + 1) * tile), ry1 = std::min(H, (ty + 1) * tile);
            for (size_t s = 0; s < n; ++s) {
                const float* c = &v[3 * s];
                if (psf.cull_behind && !(c[2] > 0.0f)) continue;
                const float hx = c[0] / c[2], hy = c[1] / c[2];
                const float X = (Km[0] * hx + Km[2] * hy) + Km[4];
                const float Y = (Km[1] * hx + Km[3] * hy) + Km[5];
                const float radius = 41;
                if (!((X + radius) >= (float)rx0 && (X - radius) < (float)rx1 && (Y + radius) >= (float)ry0 &&
                      (Y - radius) < (float)ry1))
                    continue;
                float P[NB];
                for (int k = 0; k < NB; ++k) P[k] = (eta[k] * area) * irr[NB * s + k];
                if (!(magnitude_f(P) > psf.minimum_power)) continue;
                if (psf.count == 0) {  // no PSF: processRegion assigns the power to one pixel
                    const int px = (int)X, py = (int)Y;
                    if (px >= rx0 && px < rx1 && py >= ry0 && py < ry1)
                        for (int k = 0; k < NB; ++k) img[((size_t)py * W + px) * NB + k] = P[k];
                    continue;
                }
                // getKernel (psf.ipp:89-110)
                int sel = psf.count - 1;
                if (psf.minimum_power != 0.0f) {
                    for (int i = 0; i < psf.count; ++i) {
                        sel = i;
                        float e[NB];
                        for (int k = 0; k < NB; ++k) e[k] = edge[i][k] * P[k];
                        if (magnitude_f(e) < psf.minimum_power) break;
                    }
                }
                const int K = psf.size[sel];
                // getResponse: received_power * kernel
                std::vector<float> resp((size_t)K * K * NB);
                for (size_t q = 0; q < (size_t)K * K; ++q)
                    for (int k = 0; k < NB; ++k) resp[q * NB + k] = P[k] * psf.kernel[sel][q * NB + k];
                auto R = [&](int a, int b, int k) { return resp[((size_t)a + (size_t)b * K) * NB + k]; };
                // addImage within the ROI
                const int ix = (int)std::floor(X), iy = (int)std::floor(Y);
                const float rxf = (float)K / 2.0f, ryf = (float)K / 2.0f;
                const int sx = std::max(rx0, ix - (int)rxf), sy = std::max(ry0, iy - (int)ryf);
                const int ex = std::min(rx1, ix + (int)rxf + 1), ey = std::min(ry1, iy + (int)ryf + 1);
                for (int i = sx; i < ex; ++i)
                    for (int j = sy; j < ey; ++j) {
                        const float kx = (float)i - (X - rxf), ky = (float)j - (Y - ryf);
                        const int kx0 = (int)std::floor(kx), ky0 = (int)std::floor(ky);
                        const int kx1 = kx0 + 1, ky1 = ky0 + 1;
                        const float wx1 = kx - (float)kx0, wy1 = ky - (float)ky0;
                        const float wx0 = 1.0f - wx1, wy0 = 1.0f - wy1;
                        if (!(kx0 >= 0 && kx0 < K && ky0 >= 0 && ky0 < K)) continue;
                        float* dst = img + ((size_t)j * W + i) * NB;
                        for (int k = 0; k < NB; ++k) {
                            float val = 0.0f;
                            if (kx1 < K && ky1 < K) {
                                val += R(kx0, ky0, k) * wx0 * wy0;
                                val += R(kx1, ky0, k) * wx1 * wy0;
                                val += R(kx0, ky1, k) * wx0 * wy1;
                                val += R(kx1, ky1, k) * wx1 * wy1;
                            } else if (kx1 < K) {
                                val += R(kx0, ky0, k) * wx0;
                                val += R(kx1, ky0, k) * wx1;
                            } else if (ky1 < K) {
                                val += R(kx0, ky0, k) * wy0;
                                val += R(kx0, ky1, k) * wy1;
                            } else {
                                val = R(kx0, ky0, k);
                            }
                            dst[k] += val;
                        }
                    }
            }
        }
}

// One level pass of the reference's filterPixelSeparable over a whole image.
// pix(a, b) is the linear pixel of loop coordinates (a, b); the pass moves b
// (first pass) or a (second pass), clamped to [0, nb-1] / [0, na-1].
template <class Pix>
static void eatwt_pass(const DenoiseO& o, int na, int nb, int step, bool move_b, Pix pix, const float* in,
                       float* out, const float* depth, const float* normal) {
    static const float K1[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) {
            const size_t c = pix(a, b);
            float sum[NB] = {0};
            float ws = 0.0f;
            const float cd = depth[c];
            const float* cn = normal + 3 * c;
            for (int k = -2; k <= 2; ++k) {
                int sa = move_b ? a : a + k * step;
                int sb = move_b ? b + k * step : b;
                sa = std::clamp(sa, 0, na - 1);
                sb = std::clamp(sb, 0, nb - 1);
                const size_t s = pix(sa, sb);
                const float sd = depth[s];
                const float ratio = std::max(cd, sd) / (std::min(cd, sd) + o.eps);
                const float dw = ratio < (1.0f + o.depth_threshold) ? 1.0f
                                                                     : std::exp(-(ratio - 1.0f) / o.depth_threshold);
                const float* sn = normal + 3 * s;
                const float dot = cn[0] * sn[0] + cn[1] * sn[1] + cn[2] * sn[2];
                const float sim = std::max(0.0f, dot);
                const float nw = std::max(o.min_normal_weight,
                                          std::exp(-std::pow(1.0f - sim, 2.0f) / o.normal_threshold));
                const float w = K1[k + 2] * dw * nw;
                for (int q = 0; q < NB; ++q) sum[q] += in[s * NB + q] * w;
                ws += w;
            }
            if (ws > o.eps)
                for (int q = 0; q < NB; ++q) out[c * NB + q] = sum[q] / ws;
            else
                for (int q = 0; q < NB; ++q) out[c * NB + q] = in[c * NB + q];
        }
}

void denoise_eatwt(const DenoiseO& o, int W, int H, float* direct, float* indirect, const float* albedo,
                   const float* depth, const float* normal, bool literal) {
    const size_t n = (size_t)W * H * NB;
    std::vector<float> illum[2] = {std::vector<float>(n), std::vector<float>(n)};
    float* img[2] = {direct, indirect};
    for (int m = 0; m < 2; ++m)
        for (size_t q = 0; q < n; ++q) illum[m][q] = img[m][q] / std::max(albedo[q], o.eps);
    const int levels[2] = {o.levels_direct, o.levels_indirect};
    for (int m = 0; m < 2; ++m) {
        std::vector<float> tmp(n);
        float* cur = illum[m].data();
        float* nxt = tmp.data();
        for (int level = 0; level < levels[m]; ++level) {
            const int step = 1 << level;
            for (int half = 0; half < 2; ++half) {
                if (literal)  // loops y over H, x over W; image(y, x) = data[y + x*W]; "horizontal" moves x first
                    eatwt_pass(o, H, W, step, half == 0, [W](int y, int x) { return (size_t)y + (size_t)x * W; },
                               cur, nxt, depth, normal);
                else          // loops i over W, j over H; the first pass moves j, the second i
                    eatwt_pass(o, W, H, step, half == 0, [W](int i, int j) { return (size_t)i + (size_t)j * W; },
                               cur, nxt, depth, normal);
                std::swap(cur, nxt);
            }
        }
        if (cur != illum[m].data()) std::copy(cur, cur + n, illum[m].data());
    }
    for (int m = 0; m < 2; ++m)
        for (size_t q = 0; q < n; ++q) img[m][q] = illum[m][q] * albedo[q];
}

static InstanceO identity_instance(const Mesh& m, const BVH& b, const Material& mat) {
    InstanceO I{&m, &b, mat, {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}};
    return I;
}

void render_f32(const Mesh& m, const BVH& b, const Camera& cam, const Light* lights, int nl,
                const Material& mat, const Options& opt, Outputs& out, int nthreads,
                const int* pixel_list, size_t npix) {
    InstanceO I = identity_instance(m, b, mat);
    render_impl<float>(&I, 1, cam, lights, nl, opt, out, nthreads, pixel_list, npix);
}
void render_f64(const Mesh& m, const BVH& b, const Camera& cam, const Light* lights, int nl,
                const Material& mat, const Options& opt, Outputs& out, int nthreads,
                const int* pixel_list, size_t npix) {
    InstanceO I = identity_instance(m, b, mat);
    render_impl<double>(&I, 1, cam, lights, nl, opt, out, nthreads, pixel_list, npix);
}
void render_scene(int precision, const InstanceO* inst, int n, const Camera& cam, const Light* lights, int nl,
                  const Options& opt, Outputs& out, int nthreads, const int* pixel_list, size_t npix) {
    if (precision == 64) render_impl<double>(inst, n, cam, lights, nl, opt, out, nthreads, pixel_list, npix);
    else render_impl<float>(inst, n, cam, lights, nl, opt, out, nthreads, pixel_list, npix);
}

// ---------------------------------------------------------------- single-function hooks
// The path's sampling functions one at a time, for the tests' closed-form and
// distribution pins (the renders call the very same templates).
namespace {
template <class R>
void light_sample_n(const Light& L, const double p[3], uint32_t seed, uint32_t pix, uint32_t dim0, size_t n,
                    double* dirs, double* dist, double* pdf, double* rad) {
    const V3<R> P{(R)p[0], (R)p[1], (R)p[2]};
    for (size_t i = 0; i < n; ++i) {
        V3<R> d;
        R ds, pd, rd[NB];
        sample_light<R>(L, P, seed, pix, (uint32_t)i, dim0, d, ds, pd, rd);
        dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
        dist[i] = ds;
        pdf[i] = pd;
        for (int k = 0; k < NB; ++k) rad[NB * i + k] = rd[k];
    }
}
template <class R>
void light_pdf_n(const Light& L, const double p[3], const double* dirs, size_t n, double* out) {
    const V3<R> P{(R)p[0], (R)p[1], (R)p[2]};
    for (size_t i = 0; i < n; ++i) out[i] = light_pdf<R>(L, P, V3<R>{(R)dirs[3 * i], (R)dirs[3 * i + 1], (R)dirs[3 * i + 2]});
}
template <class R>
void bounce_dirs_n(const double N_[3], uint32_t seed, uint32_t pix, uint32_t bounce, size_t n, double* dirs,
                   double* pdf) {
    const V3<R> N{(R)N_[0], (R)N_[1], (R)N_[2]};
    V3<R> tg, bt;
    tangent_frame(N, tg, bt);
    for (size_t i = 0; i < n; ++i) {
        V3<R> d{0, 0, 0};
        R mp = 0;
        if (!cosine_sample(seed, pix, (uint32_t)i, bounce, N, tg, bt, d, mp)) mp = 0;
        dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
        pdf[i] = mp;
    }
}
template <class R>
void tex_n(const Material& m, const double* uv, size_t n, double* out) {
    for (size_t i = 0; i < n; ++i) out[i] = tex_bilinear<R>(m, (R)uv[2 * i], (R)uv[2 * i + 1]);
}
}  // namespace

void hook_light_sample(int precision, const Light& L, const double p[3], uint32_t seed, uint32_t pix, uint32_t dim0,
                       size_t n, double* dirs, double* dist, double* pdf, double* rad) {
    if (precision == 64) light_sample_n<double>(L, p, seed, pix, dim0, n, dirs, dist, pdf, rad);
    else light_sample_n<float>(L, p, seed, pix, dim0, n, dirs, dist, pdf, rad);
}
void hook_light_pdf(int precision, const Light& L, const double p[3], const double* dirs, size_t n, double* out) {
    if (precision == 64) light_pdf_n<double>(L, p, dirs, n, out);
    else light_pdf_n<float>(L, p, dirs, n, out);
}
void hook_bounce_dirs(int precision, const double N[3], uint32_t seed, uint32_t pix, uint32_t bounce, size_t n,
                      double* dirs, double* pdf) {
    if (precision == 64) bounce_dirs_n<double>(N, seed, pix, bounce, n, dirs, pdf);
    else bounce_dirs_n<float>(N, seed, pix, bounce, n, dirs, pdf);
}
void hook_tex_bilinear(int precision, const Material& m, const double* uv, size_t n, double* out) {
    if (precision == 64) tex_n<double>(m, uv, n, out);
    else tex_n<float>(m, uv, n, out);
}
float hook_kahan_sum_f32(const float* x, size_t n) {
    Kahan<float> k;
    for (size_t i = 0; i < n; ++i) k.add(x[i]);
    return k.s;
}

}  // namespace vo

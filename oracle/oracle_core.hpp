// oracle_core.hpp -- CPU restatement of vira's ray-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for the HIP
// renderer in vira_amd/csrc.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it, and only as the checker.  The
// product library never links it.
//
// The reference (crgnam/vira v0.8.3) cannot be compiled here: its working path
// is Embree 3.13.5 + GLM + oneTBB (include/vira/scene.hpp:11,
// include/vira/vec.hpp:8).  This file re-expresses, in plain C++, the
// algorithms of the reference files cited at each function.  It is templated
// on the arithmetic type:
//   Real = float  : the FP32 twin.  Every operation is written in the order the
//                   HIP kernels use, compiled with -ffp-contract=off, so that
//                   primary-hit triangle ids, hit distances and deterministic
//                   radiance can be compared bit for bit.
//   Real = double : the FP64 reference answer, used to state the FP32 error.
//
// Deliberate, documented deviations from the reference (all new features the
// reference does not have, or replacements for non-reproducible parts):
//   * counter-based RNG (lowbias32 hash of seed/pixel/sample/dimension) in
//     place of the unseeded thread_local std::mt19937
//     (include/implementation/rendering/cpu_path_tracer.ipp:67-70);
//   * watertight ray/triangle test of Woop, Benthin & Wald (JCGT 2013) in
//     place of Embree's robust rtcIntersect1 (embree_tlas.ipp:128-173), with
//     ties on exactly equal t resolved to the lowest triangle id so that the
//     closest hit is independent of the BVH;
//   * conservative slab test (Ize 2013, "Robust BVH ray traversal") in place
//     of aabb.ipp:87-115, so that no valid hit is culled by rounding;
//   * shadow rays are any-hit queries with tmax = distance; the reference's
//     closest-hit "hit.t > distance" test (cpu_path_tracer.ipp:409-411) lights
//     exactly the same points;
//   * optional Russian roulette (the reference has none);
//   * Kahan-compensated FP32 accumulation over samples (the reference sums
//     naively, cpu_path_tracer.ipp:195);
//   * a bounce whose cosine sample has pdf == 0 ends the path (the reference
//     re-traces the old ray, cpu_path_tracer.ipp:437-460);
//   * the sphere light's cone (sphere_light.ipp:31-153) uses
//     1 - cos(theta_max) = x / (1 + sqrt(1 - x)), x = (r/d)^2, and samples
//     cos(theta) = 1 - zeta (1 - cos(theta_max)), sin(theta) from it: the same
//     quantities as the reference's 1 - sqrt(d^2 - r^2)/d and acos/sin/cos,
//     without the FP32 cancellation that biases a 1 AU sun by ~0.3 %.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace vo {

constexpr int NB = 8;  // spectral bins: SpectralData<8, UNIFORM_RANGE, 400, 1000>

// ---------------------------------------------------------------- hashing / RNG
inline uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Uniform float in [0,1) keyed by (seed, pixel, sample, dim).  Replaces the
// per-thread std::mt19937 of cpu_path_tracer.ipp:67-71.
inline float rng_u01(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
    uint32_t h = lowbias32(seed ^ 0x9e3779b9U);
    h = lowbias32(h ^ pixel);
    h = lowbias32(h ^ sample);
    h = lowbias32(h ^ dim);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// RNG dimension slots per bounce: dim = (bounce << 8) | slot.
enum : uint32_t {
    SLOT_JITTER_X = 0,
    SLOT_JITTER_Y = 1,
    SLOT_BSDF_U = 2,
    SLOT_BSDF_V = 3,
    SLOT_RR = 4,
    SLOT_LIGHT0 = 16,  // light l uses 16 + 2l and 17 + 2l
};

// ---------------------------------------------------------------- fBm DEM synthesis
// Value-noise fBm, 6 octaves, lacunarity 2, gain 0.5, normalised to
// +-amplitude.  Only +,-,*,/ and floor in FP64, so CPU and GPU agree bit for
// bit.  (The reference reads DEMs from files; this generator defines the
// synthetic inputs of the benchmark scenes.)
inline double fbm_lattice(int64_t ix, int64_t iy, uint32_t octave, uint32_t seed) {
    uint32_t h = lowbias32(seed);
    h = lowbias32(octave ^ h);
    h = lowbias32((uint32_t)iy ^ h);
    h = lowbias32((uint32_t)ix ^ h);
    return (double)(h >> 8) * (1.0 / 8388608.0) - 1.0;
}

inline double fbm_fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
inline double fbm_lerp(double a, double b, double t) { return a + t * (b - a); }

inline double fbm_height(int i, int j, uint32_t seed, double amplitude, int octaves,
                         double base_cells) {
    double sum = 0.0, amp = 1.0, norm = 0.0;
    double scale = 1.0 / base_cells;
    for (int o = 0; o < octaves; ++o) {
        double x = (double)i * scale;
        double y = (double)j * scale;
        double fx0 = std::floor(x), fy0 = std::floor(y);
        int64_t ix = (int64_t)fx0, iy = (int64_t)fy0;
        double fx = x - fx0, fy = y - fy0;
        double v00 = fbm_lattice(ix, iy, (uint32_t)o, seed);
        double v10 = fbm_lattice(ix + 1, iy, (uint32_t)o, seed);
        double v01 = fbm_lattice(ix, iy + 1, (uint32_t)o, seed);
        double v11 = fbm_lattice(ix + 1, iy + 1, (uint32_t)o, seed);
        double sx = fbm_fade(fx), sy = fbm_fade(fy);
        double n = fbm_lerp(fbm_lerp(v00, v10, sx), fbm_lerp(v01, v11, sx), sy);
        sum = sum + amp * n;
        norm = norm + amp;
        amp = amp * 0.5;
        scale = scale * 2.0;
    }
    return amplitude * (sum / norm);
}

// Per-vertex grey albedo 0.12 + U(-0.02, 0.02).
inline float vertex_albedo(uint32_t idx, uint32_t seed) {
    uint32_t h = lowbias32(idx ^ lowbias32(seed ^ 0xa5a5a5a5U));
    double u = (double)(h >> 8) * (1.0 / 16777216.0);
    return (float)(0.12 + 0.02 * (2.0 * u - 1.0));
}

// ---------------------------------------------------------------- small vectors
template <class T>
struct V3 {
    T x, y, z;
    T operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
};
template <class T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
// dot as GLM: (x*x' + y*y') + z*z'
template <class T> inline T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// cross as GLM (glm/detail/func_geometric.inl)
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// normalize as GLM: v * (1 / sqrt(dot(v, v)))
template <class T> inline V3<T> normalize(V3<T> a) {
    T inv = T(1) / std::sqrt(dot(a, a));
    return a * inv;
}

// ---------------------------------------------------------------- tessellation
struct Projection {
    int kind;           // 0 = flat grid, 1 = gnomonic drape on a sphere
    double spacing;     // metres per DEM cell
    double radius;      // sphere radius (kind 1)
    double center[3];   // body centre in the world frame (kind 1)
    double offset[3];   // rebasing origin subtracted in FP64 before the FP32 store
};

// FP64 world position of grid node (i, j) with height h.
// Flat: x = (i - cx) s, y = (cy - j) s, z = h (the non-OGR branch of
// DEMProjection::transformToWorld, dem_projection.ipp:130-141, with the DEM
// centred on the origin).  Sphere: gnomonic tile centred on +x.
inline V3<double> dem_world(const Projection& p, int W, int H, int i, int j, double h) {
    double cx = (double)(W - 1) * 0.5, cy = (double)(H - 1) * 0.5;
    if (p.kind == 0) {
        return {((double)i - cx) * p.spacing, (cy - (double)j) * p.spacing, h};
    }
    double a = p.spacing / p.radius;
    V3<double> d{1.0, ((double)i - cx) * a, (cy - (double)j) * a};
    double inv = 1.0 / std::sqrt(dot(d, d));
    double r = p.radius + h;
    double s = r * inv;
    return {p.center[0] + d.x * s, p.center[1] + d.y * s, p.center[2] + d.z * s};
}

// Index buffer in the order of imageToIndexBuffer (image_utils.ipp:364-412):
// i outer over W-1, j inner over H-1, two triangles per quad.  Triangles
// touching a non-finite height are dropped (the reference drops only +-inf).
inline void dem_indices(const float* heights, int W, int H, std::vector<uint32_t>& idx) {
    idx.clear();
    idx.reserve((size_t)(W - 1) * (H - 1) * 6);
    auto ok = [&](uint32_t v) { return std::isfinite(heights[v]); };
    for (uint32_t i = 0; i < (uint32_t)(W - 1); ++i) {
        for (uint32_t j = 0; j < (uint32_t)(H - 1); ++j) {
            uint32_t f0 = i + (j + 1) * (uint32_t)W, f1 = i + 1 + j * (uint32_t)W, f2 = i + j * (uint32_t)W;
            if (ok(f0) && ok(f1) && ok(f2)) { idx.push_back(f0); idx.push_back(f1); idx.push_back(f2); }
            f0 = i + 1 + j * (uint32_t)W; f1 = i + (j + 1) * (uint32_t)W; f2 = i + 1 + (j + 1) * (uint32_t)W;
            if (ok(f0) && ok(f1) && ok(f2)) { idx.push_back(f0); idx.push_back(f1); idx.push_back(f2); }
        }
    }
}

// Unit face normal normalize(cross(q1 - q0, q2 - q0)) (triangle.ipp:36-44) in
// FP64 from the FP64 positions the mesh was built from (pos64, or the FP32
// positions widened), rounded to FP32 -- the face normal the kernels store per
// triangle.  Returns false where the reference skips the triangle (a corner at
// infinity, or a degenerate triangle: NaN normal, mesh.ipp:211-216).
inline bool face_normal64(const double* q0, const double* q1, const double* q2, float out[3]) {
    if (std::isinf(q0[0]) || std::isinf(q1[0]) || std::isinf(q2[0])) return false;
    const double ax = q1[0] - q0[0], ay = q1[1] - q0[1], az = q1[2] - q0[2];
    const double bx = q2[0] - q0[0], by = q2[1] - q0[1], bz = q2[2] - q0[2];
    const double cx = ay * bz - by * az, cy = az * bx - bz * ax, cz = ax * by - bx * ay;
    const double inv = 1.0 / std::sqrt((cx * cx + cy * cy) + cz * cz);
    const double nx = cx * inv, ny = cy * inv, nz = cz * inv;
    if (std::isnan(nx) || std::isnan(ny) || std::isnan(nz)) return false;
    out[0] = (float)nx;
    out[1] = (float)ny;
    out[2] = (float)nz;
    return true;
}

// Face normals of every triangle: fn[4t..4t+2] the normal, fn[4t+3] = 1 (or 0 where skipped).
inline void face_normals(const float* pos, const double* pos64, const uint32_t* idx, size_t nt, float* fn) {
    for (size_t t = 0; t < nt; ++t) {
        double q[3][3];
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = idx[3 * t + k];
            for (int r = 0; r < 3; ++r) q[k][r] = pos64 ? pos64[3 * v + r] : (double)pos[3 * v + r];
        }
        float n[3] = {0.0f, 0.0f, 0.0f};
        const bool ok = face_normal64(q[0], q[1], q[2], n);
        fn[4 * t] = n[0];
        fn[4 * t + 1] = n[1];
        fn[4 * t + 2] = n[2];
        fn[4 * t + 3] = ok ? 1.0f : 0.0f;
    }
}

// Vertex normals as Mesh::calculateNormals (mesh.ipp:199-233): the sum of the
// unit face normals (face_normals) in triangle order, in FP32, then normalised.
inline void vertex_normals(const float* pos, const double* pos64, size_t nv, const uint32_t* idx, size_t nt,
                           float* nrm) {
    std::fill(nrm, nrm + 3 * nv, 0.0f);
    std::vector<float> fn(4 * nt);
    face_normals(pos, pos64, idx, nt, fn.data());
    for (size_t t = 0; t < nt; ++t) {
        if (fn[4 * t + 3] == 0.0f) continue;
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = idx[3 * t + k];
            nrm[3 * v] = nrm[3 * v] + fn[4 * t];
            nrm[3 * v + 1] = nrm[3 * v + 1] + fn[4 * t + 1];
            nrm[3 * v + 2] = nrm[3 * v + 2] + fn[4 * t + 2];
        }
    }
    for (size_t v = 0; v < nv; ++v) {
        V3<float> n{nrm[3 * v], nrm[3 * v + 1], nrm[3 * v + 2]};
        if (length(n) != 0.0f) {
            n = normalize(n);
            nrm[3 * v] = n.x; nrm[3 * v + 1] = n.y; nrm[3 * v + 2] = n.z;
        }
    }
}

// ---------------------------------------------------------------- watertight test
template <class R>
struct RayT {
    V3<R> o, d;
};

template <class R>
struct WTRay {  // per-ray precomputation of Woop et al. 2013, section 3
    int kx, ky, kz;
    R Sx, Sy, Sz;
};

template <class R>
inline WTRay<R> wt_setup(const V3<R>& d) {
    WTRay<R> w;
    R ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > (kz == 0 ? ax : ay)) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < R(0)) std::swap(kx, ky);
    w.kx = kx; w.ky = ky; w.kz = kz;
    w.Sx = d[kx] / d[kz];
    w.Sy = d[ky] / d[kz];
    w.Sz = R(1) / d[kz];
    return w;
}

template <class R>
struct Hit {
    R t;          // +inf on miss
    uint32_t id;  // triangle id in mesh order; 0xffffffff on miss
    R b1, b2;     // Embree-convention barycentrics: p = (1-b1-b2) v0 + b1 v1 + b2 v2
};

// Returns true and fills t/b1/b2 if the ray hits the triangle at t >= 0.
template <class R>
inline bool wt_intersect(const RayT<R>& ray, const WTRay<R>& w, const V3<R>& v0, const V3<R>& v1,
                         const V3<R>& v2, R& t_out, R& b1_out, R& b2_out) {
    V3<R> A = v0 - ray.o, B = v1 - ray.o, C = v2 - ray.o;
    R Ax = A[w.kx] - w.Sx * A[w.kz];
    R Ay = A[w.ky] - w.Sy * A[w.kz];
    R Bx = B[w.kx] - w.Sx * B[w.kz];
    R By = B[w.ky] - w.Sy * B[w.kz];
    R Cx = C[w.kx] - w.Sx * C[w.kz];
    R Cy = C[w.ky] - w.Sy * C[w.kz];
    R U = Cx * By - Cy * Bx;
    R V = Ax * Cy - Ay * Cx;
    R W = Bx * Ay - By * Ax;
    if constexpr (std::is_same<R, float>::value) {
        if (U == 0.0f || V == 0.0f || W == 0.0f) {  // fall back to FP64 edge tests
            double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
            U = (float)(CxBy - CyBx);
            double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
            V = (float)(AxCy - AyCx);
            double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
            W = (float)(BxAy - ByAx);
        }
    }
    if ((U < R(0) || V < R(0) || W < R(0)) && (U > R(0) || V > R(0) || W > R(0))) return false;
    R det = U + V + W;
    if (det == R(0)) return false;
    R Az = w.Sz * A[w.kz];
    R Bz = w.Sz * B[w.kz];
    R Cz = w.Sz * C[w.kz];
    R T = (U * Az + V * Bz) + W * Cz;
    // sign test: T must have the sign of det (t >= 0)
    if ((det < R(0) && T > R(0)) || (det > R(0) && T < R(0))) return false;
    R rcp = R(1) / det;
    R t = T * rcp;
    if (!(t >= R(0)) || std::isinf(t)) return false;
    t_out = t;
    b1_out = V * rcp;
    b2_out = W * rcp;
    return true;
}

// ---------------------------------------------------------------- BVH2
// Binned SAH exactly as ViraBLAS (vira_blas.ipp:44-55,132-254,287-292):
// longest axis of the node box, 8 bins over the centroid bounds on that axis,
// cost nL*AL + nR*AR, split iff cost < N*SA, leaf when N <= 2, in-place
// partition, abort if one side is empty.  FP64 bounds and centroids over the
// FP32 vertex positions.
struct BNode {
    double lo[3], hi[3];
    uint32_t leftFirst;  // child index (internal) or first triIdx slot (leaf)
    uint32_t triCount;   // 0 = internal
};

struct Mesh {
    std::vector<float> pos;       // 3*nv (FP32 positions as stored on the GPU)
    std::vector<double> pos64;    // 3*nv (FP64 positions; empty if not provided)
    std::vector<uint32_t> idx;    // 3*nt
    std::vector<float> nrm;       // 3*nv
    std::vector<float> alb;       // NB*nv
    std::vector<float> uv;        // 2*nv (may be empty)
    std::vector<float> fn;        // 4*nt face normals from the FP64 positions (face_normals)
    size_t nv = 0, nt = 0;
    uint32_t id = 0;              // Mesh::getID, written to the mesh_id pass
};

struct BVH {
    std::vector<BNode> nodes;
    std::vector<uint32_t> triIdx;
    // float copies of the boxes for the FP32 traversal (exact: bounds of floats)
    std::vector<float> flo, fhi;
};

inline void bvh_tri_bounds(const Mesh& m, uint32_t t, double lo[3], double hi[3], double c[3]) {
    for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
    double s[3] = {0, 0, 0};
    for (int v = 0; v < 3; ++v) {
        uint32_t vi = m.idx[3 * t + v];
        for (int k = 0; k < 3; ++k) {
            double p = (double)m.pos[3 * vi + k];
            lo[k] = std::min(lo[k], p);
            hi[k] = std::max(hi[k], p);
            s[k] += p;
        }
    }
    for (int k = 0; k < 3; ++k) c[k] = (1.0 / 3.0) * s[k];
}

BVH bvh_build(const Mesh& m, int nthreads);

// Traversal counters for the algorithmic-bytes model.
struct TravCount {
    uint64_t internal = 0;  // internal nodes whose two children were tested
    uint64_t tris = 0;      // ray/triangle tests
};

// Conservative slab test (Ize 2013): returns the entry distance or +inf.
// e[k] is 0 for ordinary direction components; for |d_k| < 1e-30 it is
// FLT_MIN and inv_k = 1/d_k is +-inf, which makes a ray lying in a box face
// count as inside the closed box (an exact 0 * inf would give NaN).
constexpr float kSlabGrowF = 1.0000005f;   // 1 + 2*gamma(3) rounded up
constexpr double kSlabGrowD = 1.0 + 1e-15;
constexpr float kTinyDir = 1e-30f;

template <class R>
inline R slab(const R lo[3], const R hi[3], const V3<R>& o, const V3<R>& inv, const V3<R>& e, R best) {
    R grow = std::is_same<R, float>::value ? (R)kSlabGrowF : (R)kSlabGrowD;
    R tx1 = ((lo[0] - o.x) - e.x) * inv.x, tx2 = ((hi[0] - o.x) + e.x) * inv.x;
    R tmin = std::min(tx1, tx2), tmax = std::max(tx1, tx2);
    R ty1 = ((lo[1] - o.y) - e.y) * inv.y, ty2 = ((hi[1] - o.y) + e.y) * inv.y;
    tmin = std::max(tmin, std::min(ty1, ty2));
    tmax = std::min(tmax, std::max(ty1, ty2));
    R tz1 = ((lo[2] - o.z) - e.z) * inv.z, tz2 = ((hi[2] - o.z) + e.z) * inv.z;
    tmin = std::max(tmin, std::min(tz1, tz2));
    tmax = std::min(tmax, std::max(tz1, tz2));
    tmin = std::max(tmin, R(0));
    tmax = tmax * grow;
    R lim = best * grow;
    if (tmin <= tmax && tmin <= lim) return tmin;
    return (R)INFINITY;
}

template <class R>
inline V3<R> slab_eps(const V3<R>& d) {
    auto f = [](R c) -> R { return std::fabs(c) < (R)kTinyDir ? (R)std::numeric_limits<float>::min() : R(0); };
    return {f(d.x), f(d.y), f(d.z)};
}

template <class R>
inline V3<R> safe_inv(const V3<R>& d) {
    return {R(1) / d.x, R(1) / d.y, R(1) / d.z};
}

template <class R>
inline V3<R> mesh_pos(const Mesh& m, uint32_t v) {
    if constexpr (std::is_same<R, double>::value) {
        if (!m.pos64.empty()) return {m.pos64[3 * v], m.pos64[3 * v + 1], m.pos64[3 * v + 2]};
    }
    return {(R)m.pos[3 * v], (R)m.pos[3 * v + 1], (R)m.pos[3 * v + 2]};
}

// Closest hit (anyhit == false) or occlusion within [0, tmax] (anyhit == true).
// Traversal order as ViraBLAS::intersect (vira_blas.ipp:58-108): test both
// children, descend into the nearer, push the farther.
template <class R>
inline Hit<R> bvh_trace(const Mesh& m, const BVH& b, const RayT<R>& ray, bool anyhit, R tmax,
                        TravCount* cnt) {
    Hit<R> h{(R)INFINITY, 0xffffffffu, 0, 0};
    R best = anyhit ? tmax : (R)INFINITY;
    WTRay<R> w = wt_setup(ray.d);
    V3<R> inv = safe_inv(ray.d);
    V3<R> eps = slab_eps(ray.d);
    auto box = [&](uint32_t n, R lo[3], R hi[3]) {
        for (int k = 0; k < 3; ++k) {
            if constexpr (std::is_same<R, float>::value) { lo[k] = b.flo[3 * n + k]; hi[k] = b.fhi[3 * n + k]; }
            else { lo[k] = (R)b.nodes[n].lo[k]; hi[k] = (R)b.nodes[n].hi[k]; }
        }
    };
    auto test_leaf = [&](const BNode& nd) -> bool {
        for (uint32_t i = 0; i < nd.triCount; ++i) {
            uint32_t t = b.triIdx[nd.leftFirst + i];
            if (cnt) cnt->tris++;
            R tt, b1, b2;
            V3<R> v0 = mesh_pos<R>(m, m.idx[3 * t]), v1 = mesh_pos<R>(m, m.idx[3 * t + 1]),
                  v2 = mesh_pos<R>(m, m.idx[3 * t + 2]);
            if (!wt_intersect(ray, w, v0, v1, v2, tt, b1, b2)) continue;
            if (anyhit) {
                if (tt <= tmax) { h.t = tt; h.id = t; h.b1 = b1; h.b2 = b2; return true; }
            } else if (tt < h.t || (tt == h.t && t < h.id)) {
                h.t = tt; h.id = t; h.b1 = b1; h.b2 = b2; best = tt;
            }
        }
        return false;
    };
    R lo[3], hi[3];
    if (b.nodes.empty()) return h;
    box(0, lo, hi);
    if (std::isinf(slab(lo, hi, ray.o, inv, eps, best))) return h;
    uint32_t stack[128];
    int sp = 0;
    uint32_t node = 0;
    while (true) {
        const BNode& nd = b.nodes[node];
        if (nd.triCount > 0) {
            if (test_leaf(nd)) return h;
            if (sp == 0) break;
            node = stack[--sp];
            continue;
        }
        if (cnt) cnt->internal++;
        uint32_t c1 = nd.leftFirst, c2 = nd.leftFirst + 1;
        box(c1, lo, hi);
        R d1 = slab(lo, hi, ray.o, inv, eps, best);
        box(c2, lo, hi);
        R d2 = slab(lo, hi, ray.o, inv, eps, best);
        if (d1 > d2) { std::swap(d1, d2); std::swap(c1, c2); }
        if (std::isinf(d1)) {
            if (sp == 0) break;
            node = stack[--sp];
        } else {
            node = c1;
            if (!std::isinf(d2)) stack[sp++] = c2;
        }
    }
    return h;
}

template <class R>
inline Hit<R> brute_trace(const Mesh& m, const RayT<R>& ray, bool anyhit, R tmax) {
    Hit<R> h{(R)INFINITY, 0xffffffffu, 0, 0};
    WTRay<R> w = wt_setup(ray.d);
    for (uint32_t t = 0; t < (uint32_t)m.nt; ++t) {
        R tt, b1, b2;
        if (!wt_intersect(ray, w, mesh_pos<R>(m, m.idx[3 * t]), mesh_pos<R>(m, m.idx[3 * t + 1]),
                          mesh_pos<R>(m, m.idx[3 * t + 2]), tt, b1, b2))
            continue;
        if (anyhit) {
            if (tt <= tmax) { h.t = tt; h.id = t; h.b1 = b1; h.b2 = b2; return h; }
        } else if (tt < h.t || (tt == h.t && t < h.id)) {
            h.t = tt; h.id = t; h.b1 = b1; h.b2 = b2;
        }
    }
    return h;
}

// ---------------------------------------------------------------- scene description
struct Camera {
    int W, H;
    float kinv[6];   // K^-1 as GLM mat3x2 [col][row]: k00 k01 | k10 k11 | k20 k21
    float rot[9];    // row-major R (columns = camera x, y, z axes in world)
    double pos[3];   // camera position in the rebased world frame
    float zdir;      // +1 (computer-vision frame) or -1 (Blender frame)
    float view_normal[9];  // Camera::getViewNormalMatrix, row-major; all zero = transpose(rot)
};

struct Light {
    int kind;           // 0 = point, 1 = sphere
    double pos[3];      // rebased world position
    float radius;       // sphere radius
    float value[NB];    // point: spectral intensity P/4pi; sphere: spectral radiance
};

struct Material {
    float albedo[NB];          // constant albedo (multiplies the texture and vertex albedo)
    int tex_w = 0, tex_h = 0;  // optional grey albedo texture, row-major i + j*W
    const float* tex = nullptr;
    int smooth = 0;            // Mesh::smoothShading (default false, mesh.hpp:144)
    int tex_channels = 1;      // 1: grey texture x albedo[k]; 8: spectral albedo map (Image<TSpectral>)
    const float* nmap = nullptr;  // tangent-space normal map, 3 floats per texel (Material::getNormal)
    int nmap_w = 0, nmap_h = 0;
    uint32_t material_id = 0;  // Material::getID, written to the material_id pass
};

struct Options {
    int spp = 1;
    int bounces = 0;
    int rr = 0;          // Russian roulette on bounce rays
    int rr_start = 1;    // first bounce index whose continuation is subject to RR
    uint32_t seed = 0;
    // CPUPathTracerOptions (cpu_path_tracer.hpp:30-43)
    int show_background = 0;
    int adaptive = 0;
    int samples_per_batch = 30;
    float sampling_tolerance = 0.05f;
    int samples_to_detect_miss = 30;
    // Scene::getAmbient / hasAmbient, getBackgroundEmission (scene.hpp:100-113)
    float ambient[NB] = {0, 0, 0, 0, 0, 0, 0, 0};
    int has_ambient = 0;
    const float* background = nullptr;  // W*H*NB spectral image, or null (zero background)
    int bg_w = 0, bg_h = 0;
};

struct Outputs {  // all optional (nullptr = skip); pixel index i + j*W
    float* radiance = nullptr;  // NB per pixel
    float* direct = nullptr;    // NB
    float* indirect = nullptr;  // NB
    float* albedo = nullptr;    // NB
    float* normal = nullptr;    // 3
    float* depth = nullptr;
    float* alpha = nullptr;
    uint32_t* tri_id = nullptr;
    uint64_t* counts = nullptr;  // [0] primary rays, [1] shadow rays, [2] bounce rays,
                                 // [3..4] internal/tris closest, [5..6] internal/tris any-hit
    uint32_t* records = nullptr; // 4 per (sample, pixel) at [s * W*H + pixel]: first-hit id,
                                 // b1 and b2 bits, bit l = light l sampled and unoccluded
    uint32_t* instance_id = nullptr;
    float* normal_camera = nullptr;  // 3
    uint32_t* mesh_id = nullptr;
    uint32_t* material_id = nullptr;
};

// One instance of a two-level scene (EmbreeTLAS::newInstance, embree_tlas.ipp:77-115):
// a mesh with its BVH and material placed by an FP64 column-major model matrix.
struct InstanceO {
    const Mesh* m;
    const BVH* b;
    Material mat;
    double M[16];
};

// Test instrument: g_plant != 0 plants one known mistake in the oracle, so
// that the tests can show that each pin fails on it (tests/test_oracle_plants.py).
// 0 everywhere else; set only from the tests, before a render starts.
enum : int {
    PLANT_NONE = 0,
    PLANT_MIS_BALANCE = 1,        // balance heuristic instead of the power heuristic
    PLANT_DROP_COSINE = 2,        // the light term loses cos(theta) at the surface
    PLANT_CONE_OMEGA = 3,         // sphere-light solid angle pi (1 - cos) instead of 2 pi (1 - cos)
    PLANT_CONE_SQUARE = 4,        // cone samples with 1 - cos = zeta^2 q (not uniform in cos)
    PLANT_BOUNCE_PDF = 5,         // cosine-sampled bounce weighted with the uniform pdf 1/(2 pi)
    PLANT_RR_NO_REWEIGHT = 6,     // a path that survives Russian roulette is not divided by q
    PLANT_TEX_OFF_BY_ONE = 7,     // bilinear lookup one texel to the right
    PLANT_TEX_TRANSPOSED = 8,     // bilinear lookup with u and v swapped
    PLANT_KAHAN_OFF = 9,          // Kahan accumulation without the compensation
    PLANT_NORMAL_MATRIX = 10,     // instances: normal matrix not transposed
    PLANT_BACKGROUND_SWAP = 11,   // background u = atan2(x, y) / 2 pi
    PLANT_ADAPTIVE_DIVISOR = 12,  // adaptive stop at sample s divides by s + 1
    PLANT_NMAP_UNNORMALISED = 13, // normal-mapped normal not renormalised
    PLANT_AMBIENT_NO_ALBEDO = 14, // ambient light not multiplied by the albedo
};
extern int g_plant;

template <class R>
struct Kahan {
    R s = 0, c = 0;
    void add(R x) {
        if (g_plant == PLANT_KAHAN_OFF) {
            s = s + x;
            return;
        }
        R y = x - c;
        R t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

V3<float> offset_point_f32(V3<float> p, V3<float> n);

void render_f32(const Mesh& m, const BVH& b, const Camera& cam, const Light* lights, int nl,
                const Material& mat, const Options& opt, Outputs& out, int nthreads,
                const int* pixel_list, size_t npix);
void render_f64(const Mesh& m, const BVH& b, const Camera& cam, const Light* lights, int nl,
                const Material& mat, const Options& opt, Outputs& out, int nthreads,
                const int* pixel_list, size_t npix);
// Pixel solid angles of Camera::initializePixelSolidAngle (camera.ipp:1594-1666):
// FP64 corner directions K^-1 (x, y, 1) (the FP32 K^-1 widened), two spherical
// triangles by Girard's theorem (camera.ipp:1578-1592), stored FP32.
void pixel_solid_angles(const float kinv[6], float zdir, int W, int H, float* out);

// Point-spread functions (same field layout as the library's vira_gpu_psf):
// GaussianPSF::evaluate (gaussian_psf.ipp), AiryDiskPSF::evaluate
// (airy_disk.ipp), kernels by PointSpreadFunction::makeKernel (psf.ipp).
struct PSFO {
    int kind;  // 0 Gaussian, 1 Airy disk
    float sigma_x[NB], sigma_y[NB];
    float angle_deg;
    float focal_length, aperture_diameter;
    float pixel_size[2];
    float wavelength[NB];
    int supersample;
};
void psf_kernel(const PSFO& p, int K, float* out);  // out[(x + y*K)*NB + k]
// Image::convolveSpatial (image.ipp:1267-1330), FP32, zero padding, ky outer / kx inner.
void convolve_spatial(const float* in, int W, int H, const float* kern, int K, float* out);

// Unresolved point sources (CPUUnresolvedRenderer::render / findPointsInRegion /
// processRegion, cpu_unresolved_renderer.ipp:25-200): camera-frame direction
// R^T d (rot row-major), projection K (x/z, y/z, 1) (camera.ipp:1062-1081),
// power (eta*A)*E (camera.ipp:1187-1191), 512-pixel ROIs each taking the
// sources within 41 px in source order, kernel choice by the edge maxima
// that PointSpreadFunction::initKernels finds (psf.ipp:134-162, computed here
// from the kernels), and Image::addImage (image.ipp:1074-1150) into img
// (W*H*NB, accumulated).  cull_behind = 1 skips camera-frame z <= 0.
struct StarPSFO {
    int count;
    int size[4];
    const float* kernel[4];  // [(x + y*size)*NB + k]
    float minimum_power;
    float intrinsic[6];      // GLM mat3x2 [col][row]
    int cull_behind;
};
void render_stars(const float rot[9], int W, int H, const float eta[NB], float area, const StarPSFO& psf,
                  const float* dirs, const float* irr, size_t n, float* img);

// Edge-avoiding a-trous wavelet denoiser (denoiseSpectralRadianceEATWT,
// cpu_denoise.ipp:10-187; options cpu_denoise.hpp:9-18).  literal = true
// follows the reference's image(y, x) = data[y + x*W] indexing (square images
// only); literal = false filters first along j (clamped to H-1), then along i
// (clamped to W-1), which is the same thing for W == H and defined for any
// W x H.  direct and indirect (W*H*NB) are denoised in place.
struct DenoiseO {
    int levels_direct, levels_indirect;
    float depth_threshold, normal_threshold, min_normal_weight, eps;
};
void denoise_eatwt(const DenoiseO& o, int W, int H, float* direct, float* indirect, const float* albedo,
                   const float* depth, const float* normal, bool literal);

// Two-level scene.  One instance with the identity matrix renders in the
// caller's frame; anything else renders camera-relative: model matrices and
// light positions minus the FP64 camera position, rounded to the working
// precision as the GPU does (FP32) or kept (FP64).
void render_scene(int precision, const InstanceO* inst, int n, const Camera& cam, const Light* lights, int nl,
                  const Options& opt, Outputs& out, int nthreads, const int* pixel_list, size_t npix);

// Single-function hooks for the tests' pins (precision 32 = FP32 twin, 64 = FP64):
// Light::sample for samples s = 0..n-1 at point p (RNG dimensions dim0, dim0+1);
// Light::getPDF of n directions; the cosine-weighted bounce direction and pdf in the
// shading frame of normal N for samples 0..n-1; the bilinear albedo lookup at n uvs;
// the Kahan sum the pixel accumulators use.
void hook_light_sample(int precision, const Light& L, const double p[3], uint32_t seed, uint32_t pix, uint32_t dim0,
                       size_t n, double* dirs, double* dist, double* pdf, double* rad);
void hook_light_pdf(int precision, const Light& L, const double p[3], const double* dirs, size_t n, double* out);
void hook_bounce_dirs(int precision, const double N[3], uint32_t seed, uint32_t pix, uint32_t bounce, size_t n,
                      double* dirs, double* pdf);
void hook_tex_bilinear(int precision, const Material& m, const double* uv, size_t n, double* out);
float hook_kahan_sum_f32(const float* x, size_t n);

}  // namespace vo

// device_math.hpp -- gfx950 device helpers shared by the vira_gpu kernels.
//
// All code is compiled with -ffp-contract=off and correctly rounded FP32
// division and square root: every +, -, *, / and sqrt below is a
// separately rounded IEEE operation in the order written, which is what makes
// primary-hit ids and deterministic radiance reproducible bit for bit against
// a CPU restatement of the same formulas.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vg {

constexpr int NB = 8;  // spectral bins

// __syncthreads() that first waits for the calling wave's LDS operations.
// For a workgroup barrier the compiler waits for LDS only where other memory
// has to be ordered as well, and a wave's LDS store still in flight at the
// barrier was seen on gfx950 to land after another wave's load issued past
// it (a per-bin count stored by one wave just before the barrier was read
// back stale by the others).  Used wherever one wave's LDS stores just
// before a barrier are read by other waves after it.
__device__ __forceinline__ void block_sync() {
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt and expcnt left alone
    __syncthreads();
}
// The same for global memory written by one wave and read by the others after
// the barrier: every memory operation of the wave completes first (the
// compiler's workgroup barrier does not wait for global stores either).
__device__ __forceinline__ void block_sync_all() {
    __builtin_amdgcn_s_waitcnt(0);  // vmcnt(0) expcnt(0) lgkmcnt(0)
    __syncthreads();
}

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 mul(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
// GLM order: (x*x' + y*y') + z*z'
__device__ __forceinline__ float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// GLM cross (glm/detail/func_geometric.inl)
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ float length(f3 a) { return __builtin_sqrtf(dot(a, a)); }
// GLM normalize: v * (1 / sqrt(dot(v, v)))
__device__ __forceinline__ f3 normalize(f3 a) {
    float inv = __fdiv_rn(1.0f, __builtin_sqrtf(dot(a, a)));
    return mul(a, inv);
}
__device__ __forceinline__ float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// ---------------------------------------------------------------- hashing / RNG
__device__ __forceinline__ uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ float rng_u01(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
    uint32_t h = lowbias32(seed ^ 0x9e3779b9U);
    h = lowbias32(h ^ pixel);
    h = lowbias32(h ^ sample);
    h = lowbias32(h ^ dim);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
enum : uint32_t {
    SLOT_JITTER_X = 0,
    SLOT_JITTER_Y = 1,
    SLOT_BSDF_U = 2,
    SLOT_BSDF_V = 3,
    SLOT_RR = 4,
    SLOT_LIGHT0 = 16,
};

// ---------------------------------------------------------------- order-preserving float keys
__device__ __forceinline__ int f2key(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __forceinline__ float key2f(int k) { return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff)); }

// ---------------------------------------------------------------- watertight ray/triangle
// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013.
struct WT {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ __forceinline__ WT wt_setup(f3 d) {
    WT w;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > (kz == 0 ? ax : ay)) kz = 2;
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    if (comp(d, kz) < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    w.kx = kx;
    w.ky = ky;
    w.kz = kz;
    float dz = comp(d, kz);
    w.Sx = __fdiv_rn(comp(d, kx), dz);
    w.Sy = __fdiv_rn(comp(d, ky), dz);
    w.Sz = __fdiv_rn(1.0f, dz);
    return w;
}

// Returns true with t >= 0 and Embree barycentrics (b1, b2) on a hit.
__device__ __forceinline__ bool wt_intersect(f3 o, const WT& w, f3 v0, f3 v1, f3 v2, float& t_out,
                                             float& b1_out, float& b2_out) {
    f3 A = sub(v0, o), B = sub(v1, o), C = sub(v2, o);
    float Akx = comp(A, w.kx), Aky = comp(A, w.ky), Akz = comp(A, w.kz);
    float Bkx = comp(B, w.kx), Bky = comp(B, w.ky), Bkz = comp(B, w.kz);
    float Ckx = comp(C, w.kx), Cky = comp(C, w.ky), Ckz = comp(C, w.kz);
    float Ax = Akx - w.Sx * Akz;
    float Ay = Aky - w.Sy * Akz;
    float Bx = Bkx - w.Sx * Bkz;
    float By = Bky - w.Sy * Bkz;
    float Cx = Ckx - w.Sx * Ckz;
    float Cy = Cky - w.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = w.Sz * Akz;
    float Bz = w.Sz * Bkz;
    float Cz = w.Sz * Ckz;
    float T = (U * Az + V * Bz) + W * Cz;
    if ((det < 0.0f && T > 0.0f) || (det > 0.0f && T < 0.0f)) return false;
    float rcp = __fdiv_rn(1.0f, det);
    float t = T * rcp;
    if (!(t >= 0.0f) || isinf(t)) return false;
    t_out = t;
    b1_out = V * rcp;
    b2_out = W * rcp;
    return true;
}

// ---------------------------------------------------------------- conservative slab test
// Ize, "Robust BVH Ray Traversal", JCGT 2(2), 2013: the far distance is grown
// by 1 + 2*gamma(3) so rounding never culls a box the ray touches.
constexpr float kSlabGrow = 1.0000005f;
constexpr float kTinyDir = 1e-30f;

// inv = 1/d (IEEE: +-inf for a zero component).  e_k is FLT_MIN where
// |d_k| < 1e-30, else 0: with SAFE, the bounds are moved out by e_k before the
// multiply, so a ray lying in a box face counts as inside the closed box
// instead of producing 0 * inf = NaN.  SAFE is only needed for rays with such
// a component; without it the e terms are left out.
template <bool SAFE>
__device__ __forceinline__ float slab(float lox, float loy, float loz, float hix, float hiy, float hiz, f3 o,
                                      f3 inv, f3 e, float best) {
    float ax = lox - o.x, bx = hix - o.x, ay = loy - o.y, by = hiy - o.y, az = loz - o.z, bz = hiz - o.z;
    if (SAFE) {
        ax = ax - e.x; bx = bx + e.x;
        ay = ay - e.y; by = by + e.y;
        az = az - e.z; bz = bz + e.z;
    }
    float tx1 = ax * inv.x, tx2 = bx * inv.x;
    float tmin = fminf(tx1, tx2), tmax = fmaxf(tx1, tx2);
    float ty1 = ay * inv.y, ty2 = by * inv.y;
    tmin = fmaxf(tmin, fminf(ty1, ty2));
    tmax = fminf(tmax, fmaxf(ty1, ty2));
    float tz1 = az * inv.z, tz2 = bz * inv.z;
    tmin = fmaxf(tmin, fminf(tz1, tz2));
    tmax = fminf(tmax, fmaxf(tz1, tz2));
    tmin = fmaxf(tmin, 0.0f);
    tmax = tmax * kSlabGrow;
    float lim = best * kSlabGrow;
    return (tmin <= tmax && tmin <= lim) ? tmin : __int_as_float(0x7f800000);
}

__device__ __forceinline__ float slab_eps(float c) { return fabsf(c) < kTinyDir ? 1.17549435e-38f : 0.0f; }

// ---------------------------------------------------------------- BVH layout
// Binary node holding its two children's boxes (64 B, one cache line).
// Child reference: bit 31 clear -> internal node index; bit 31 set -> leaf
// with first triangle slot (ref >> 4) & 0x7ffffff and count (ref & 15) + 1;
// 0xffffffff -> empty child.
struct __attribute__((aligned(16))) Node2 {
    float4 x;  // c0.lo.x c0.hi.x c1.lo.x c1.hi.x
    float4 y;
    float4 z;
    uint4 ref; // ref0 ref1 - -
};
constexpr uint32_t kEmptyRef = 0xffffffffu;

// Four-wide node (128 B, two cache lines): per axis one float4 of the four
// children's lower bounds and one of their upper bounds, then the four child
// references (same encoding as Node2).
struct __attribute__((aligned(16))) Node4 {
    float4 lox, hix, loy, hiy, loz, hiz;
    uint4 ref;
    uint4 pad;
};
static_assert(sizeof(Node4) == 128, "Node4 is two cache lines");
// Four-wide node with quantised child boxes (64 B, one cache line; traversal
// width code 5).  Per axis a the node keeps an FP32 origin o[a] and a scale
// 2^(ex[a] - 127); child c's bounds on that axis are the bytes c of qlo[a] and
// qhi[a], decoded as fmaf(q, scale, origin) in FP32.  The compressing kernel
// chooses every byte with that same instruction so that the decoded box
// contains the child's box, rounding outward, so no ray that reaches a child
// is refused by its decoded box and the conservative slab test keeps its
// guarantee.
struct __attribute__((aligned(64))) Node4Q {
    uint4 ref;
    float ox, oy, oz;
    uint32_t ex;       // byte a: biased exponent of axis a's scale
    uint32_t qlo[3];   // per axis: byte c = child c's lower bound
    uint32_t qhi[3];   // per axis: byte c = child c's upper bound
    uint32_t pad[2];
};
static_assert(sizeof(Node4Q) == 64, "Node4Q is one cache line");
constexpr int kWidthQ = 5;  // traversal width code of Node4Q trees

__device__ __forceinline__ float q_scale(uint32_t ex, int a) { return __uint_as_float(((ex >> (8 * a)) & 0xffu) << 23); }
__device__ __forceinline__ float q_decode(uint32_t q, int c, float scale, float origin) {
    return __builtin_fmaf((float)((q >> (8 * c)) & 0xffu), scale, origin);
}

__device__ __host__ __forceinline__ bool ref_is_leaf(uint32_t r) { return (r & 0x80000000u) != 0; }
__device__ __host__ __forceinline__ uint32_t leaf_first(uint32_t r) { return (r >> 4) & 0x7ffffffu; }
__device__ __host__ __forceinline__ uint32_t leaf_count(uint32_t r) { return (r & 15u) + 1u; }
__device__ __host__ __forceinline__ uint32_t make_leaf(uint32_t first, uint32_t count) {
    return 0x80000000u | (first << 4) | (count - 1u);
}

// Vertex record: 64 B, the Vertex<TSpectral=8 bins, float> of vertex.hpp:18-48
// (position, spectral albedo, normal, uv) so one hit fetches 3 cache lines.
struct __attribute__((aligned(64))) GVertex {
    float px, py, pz;
    float alb[NB];
    float nx, ny, nz;
    float u, v;
};
static_assert(sizeof(GVertex) == 64, "vertex record is one cache line");

// Adds v over the wave's active lanes into *counter with one atomic.
__device__ __forceinline__ void wave_add(unsigned long long* counter, uint32_t v) {
    // one atomic per wave: reduce over the active lanes
    unsigned long long active = __ballot(1);
    uint32_t sum = v;
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t o = __shfl_xor(sum, off, 64);
        // lanes that exited contribute nothing: their shuffled value is read from an inactive lane
        if (active & (1ull << (__lane_id() ^ off))) sum += o;
    }
    if (__lane_id() == (unsigned)(__ffsll((long long)active) - 1)) atomicAdd(counter, (unsigned long long)sum);
}

// The same with 64-bit sums (counts that can pass 2^32 within one wave).
__device__ __forceinline__ void wave_add64(unsigned long long* counter, unsigned long long v) {
    unsigned long long active = __ballot(1);
    unsigned long long sum = v;
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(sum, off, 64);
        if (active & (1ull << (__lane_id() ^ off))) sum += o;
    }
    if (__lane_id() == (unsigned)(__ffsll((long long)active) - 1)) atomicAdd(counter, sum);
}

// Maximum over the wave's active lanes, one atomic per wave.
__device__ __forceinline__ void wave_max(unsigned long long* counter, uint32_t v) {
    unsigned long long active = __ballot(1);
    uint32_t m = v;
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t o = __shfl_xor(m, off, 64);
        if ((active & (1ull << (__lane_id() ^ off))) && o > m) m = o;
    }
    if (__lane_id() == (unsigned)(__ffsll((long long)active) - 1)) atomicMax(counter, (unsigned long long)m);
}

__device__ __forceinline__ f3 vpos(const GVertex* V, uint32_t i) { return f3{V[i].px, V[i].py, V[i].pz}; }

// Unit face normal normalize(cross(q1 - q0, q2 - q0)) (triangle.ipp:36-44,
// mesh.ipp:211-214) evaluated in FP64 from the FP64 positions the mesh was
// built from, stored FP32: x, y, z and w = 1, or w = 0 where the reference
// skips the triangle (a corner at infinity or a degenerate triangle).  FP32
// positions tens of kilometres from the origin are quantised to millimetres,
// which would tilt the normals of 30 m facets by ~1e-4 rad; from FP64 they
// keep FP32 precision.
__device__ __forceinline__ float4 face_normal64(const double* q0, const double* q1, const double* q2) {
    if (isinf(q0[0]) || isinf(q1[0]) || isinf(q2[0])) return make_float4(0.f, 0.f, 0.f, 0.f);
    const double ax = q1[0] - q0[0], ay = q1[1] - q0[1], az = q1[2] - q0[2];
    const double bx = q2[0] - q0[0], by = q2[1] - q0[1], bz = q2[2] - q0[2];
    const double cx = ay * bz - by * az, cy = az * bx - bz * ax, cz = ax * by - bx * ay;
    const double inv = __ddiv_rn(1.0, __dsqrt_rn((cx * cx + cy * cy) + cz * cz));
    const double nx = cx * inv, ny = cy * inv, nz = cz * inv;
    if (isnan(nx) || isnan(ny) || isnan(nz)) return make_float4(0.f, 0.f, 0.f, 0.f);
    return make_float4((float)nx, (float)ny, (float)nz, 1.0f);
}

}  // namespace vg

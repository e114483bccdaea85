// instance.hpp -- two-level scenes: instances of meshes placed by FP64 model
// matrices, rendered in the camera's frame.
//
// EmbreeTLAS::newInstance / intersect (embree_tlas.ipp:77-173) with
// Instance::getModelMatrix / getModelNormalMatrix in processIntersection
// (cpu_path_tracer.ipp:288-292,303-310,341-344,386).  The reference keeps FP32
// model matrices in the global frame; here each model matrix stays FP64 until
// a kernel (k_rebase in render.hip) subtracts the FP64 camera position and
// rounds the camera-relative matrix, its inverse and its normal matrix to
// FP32, so bodies at 1.5e8 km are placed to FP32 precision relative to the
// camera rather than to the global origin.
#pragma once

#include "shading.hpp"
#include "traverse.hpp"

namespace vg {

constexpr int kMaxInstances = 1 << 16;

// TLAS node in the camera-relative frame, in depth-first order with skip
// links (a stackless walk): inst >= 0 marks a leaf holding that instance,
// whose box is the instance's placed box; an internal node's box is the union
// of the leaves in [index + 1, skip), and its first child follows it.
struct TNode {
    float lo[3], hi[3];
    int32_t inst;
    int32_t skip;  // the node after this subtree
};
static_assert(sizeof(TNode) == 32, "TLAS node is 32 bytes");

struct InstG {
    const void* nodes;
    const float4* tris;
    const GVertex* verts;
    const uint32_t* idx;
    MatF mat;
    float l2w[12];  // row-major 3x4: mesh frame -> camera-relative frame
    float w2l[12];  // its inverse
    float nm[9];    // normal matrix: transpose(inverse(linear part)), row-major
    float blo[3], bhi[3];  // the BLAS's box in the mesh frame (lo > hi: empty)
    float wlo[3], whi[3];  // that box placed in the camera-relative frame, rounded outward and padded
    uint32_t mesh_id, material_id;  // for the mesh_id and material_id passes
};

// GLM order: ((m0 x + m1 y) + m2 z) + m3 per row.
__device__ __forceinline__ f3 xform_point(const float* m, f3 p) {
    return mk3(((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3], ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7],
               ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11]);
}
__device__ __forceinline__ f3 xform_dir(const float* m, f3 d) {
    return mk3((m[0] * d.x + m[1] * d.y) + m[2] * d.z, (m[4] * d.x + m[5] * d.y) + m[6] * d.z,
               (m[8] * d.x + m[9] * d.y) + m[10] * d.z);
}
__device__ __forceinline__ f3 xform_normal(const float* n, f3 v) {
    return mk3((n[0] * v.x + n[1] * v.y) + n[2] * v.z, (n[3] * v.x + n[4] * v.y) + n[5] * v.z,
               (n[6] * v.x + n[7] * v.y) + n[8] * v.z);
}

// Camera-relative FP32 matrices of one instance from its FP64 column-major
// model matrix M and the FP64 camera position c.
__device__ __forceinline__ void rebase_instance(const double* M, double cx, double cy, double cz, InstG& I) {
    double A[3][3];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) A[r][k] = M[k * 4 + r];
    const double t[3] = {M[12] - cx, M[13] - cy, M[14] - cz};
    const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    const double det = (A[0][0] * c00 + A[0][1] * c01) + A[0][2] * c02;
    const double id = __ddiv_rn(1.0, det);
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
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) {
            I.l2w[r * 4 + k] = (float)A[r][k];
            I.w2l[r * 4 + k] = (float)inv[r][k];
            I.nm[r * 3 + k] = (float)inv[k][r];
        }
        I.l2w[r * 4 + 3] = (float)t[r];
        I.w2l[r * 4 + 3] = (float)(-((inv[r][0] * t[0] + inv[r][1] * t[1]) + inv[r][2] * t[2]));
    }
    // World box of the mesh box (FP64 corners), padded by 1e-5 of its size and
    // position so that the FP32 ray transforms of the traversal can never
    // reach a triangle the box test has already rejected.
    if (!(I.blo[0] <= I.bhi[0])) {
        for (int k = 0; k < 3; ++k) { I.wlo[k] = 1.f; I.whi[k] = -1.f; }
        return;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(double)((c & 1) ? I.bhi[0] : I.blo[0]), (double)((c & 2) ? I.bhi[1] : I.blo[1]),
                             (double)((c & 4) ? I.bhi[2] : I.blo[2])};
        for (int r = 0; r < 3; ++r) {
            const double w = ((A[r][0] * p[0] + A[r][1] * p[1]) + A[r][2] * p[2]) + t[r];
            lo[r] = fmin(lo[r], w);
            hi[r] = fmax(hi[r], w);
        }
    }
    double pad = 0.0;
    for (int r = 0; r < 3; ++r) pad = fmax(pad, fmax(fabs(lo[r]), fabs(hi[r])) + (hi[r] - lo[r]));
    pad *= 1e-5;
    for (int r = 0; r < 3; ++r) {
        I.wlo[r] = __double2float_rd(lo[r] - pad);
        I.whi[r] = __double2float_ru(hi[r] + pad);
    }
}

// Closest hit (or occlusion) over the instances through the TLAS: the ray is
// carried into each reached instance's frame (t is the same in both frames)
// and its BLAS is traversed with the best t so far.  Subtrees whose box the
// ray misses, or enters beyond the best hit, are skipped.  Ties in t go to the
// lower instance, so the hit does not depend on the visiting order.
template <int W, bool STATS = false>
__device__ __forceinline__ HitRec trace_instances(const InstG* __restrict__ inst, const TNode* __restrict__ tn,
                                                  int nn, f3 o, f3 d, float tmax, bool any, uint32_t* lds_ref,
                                                  float* lds_dist, int stride, uint32_t& hit_inst,
                                                  TravCount* tc = nullptr) {
    HitRec best;
    best.t = __builtin_huge_valf();
    best.id = 0xffffffffu;
    best.b1 = best.b2 = 0.f;
    hit_inst = 0;
    float bt = tmax;
    const f3 inv = mk3(__fdiv_rn(1.0f, d.x), __fdiv_rn(1.0f, d.y), __fdiv_rn(1.0f, d.z));
    const f3 e = mk3(slab_eps(d.x), slab_eps(d.y), slab_eps(d.z));
    int i = 0;
    while (i < nn) {
        const TNode N = tn[i];
        if (STATS) tc->tlas++;
        if (!(slab<true>(N.lo[0], N.lo[1], N.lo[2], N.hi[0], N.hi[1], N.hi[2], o, inv, e, bt) <
              __builtin_huge_valf())) {
            i = N.skip;
            continue;
        }
        if (N.inst < 0) {
            i++;
            continue;
        }
        const InstG& I = inst[N.inst];
        const f3 ol = xform_point(I.w2l, o), dl = xform_dir(I.w2l, d);
        const HitRec h = traverse<W, STATS>(I.nodes, I.tris, ol, dl, bt, any, lds_ref, lds_dist, stride, tc);
        if (h.id != 0xffffffffu &&
            (best.id == 0xffffffffu || h.t < best.t || (h.t == best.t && (uint32_t)N.inst < hit_inst))) {
            best = h;
            hit_inst = (uint32_t)N.inst;
            bt = h.t;
            if (any) break;
        }
        i = N.skip;
    }
    return best;
}

// A hit on instance I: local hit, then model matrix for the point and normal
// matrix for the normals (not renormalised, as the reference).
template <bool FEAT = true>
__device__ __forceinline__ Surface surface_instance(const InstG& I, uint32_t id, float b1, float b2) {
    LocalHit L;
    local_hit<FEAT>(I.verts, I.idx, I.mat, id, b1, b2, L);
    return finish_surface<FEAT>(xform_point(I.l2w, L.P), xform_normal(I.nm, L.Nl), xform_normal(I.nm, L.fn), L.alb,
                                I.mat, L.u, L.v);
}

}  // namespace vg

// traverse.hpp -- BVH traversal with watertight intersection (row a5, a6).
//
// Replaces rtcIntersect1 (embree_tlas.ipp:128-173) and restates the order of
// ViraBLAS::intersect (vira_blas.ipp:58-108): both child boxes are tested,
// the nearer child is entered and the farther one pushed.  The stack's top 16
// entries live in LDS (one column per lane, so a wave's pushes and pops hit 64
// consecutive dwords); deeper entries spill to private memory, which the
// tree depth (< 63) bounds.
#pragma once

#include "device_math.hpp"

namespace vg {

constexpr int kStackLds = 16;

struct HitRec {
    float t;
    uint32_t id;
    float b1, b2;
};

// lds_ref/lds_dist point at this lane's column: entry k is at [k * stride].
// any != 0: occlusion query within [0, tmax] (returns on the first hit);
// otherwise closest hit within [0, tmax], ties on t resolved to the lowest id.
template <bool SAFE>
__device__ __forceinline__ HitRec traverse_impl(const Node2* __restrict__ nodes, const float4* __restrict__ tris,
                                                f3 o, f3 d, float tmax, bool any, uint32_t* lds_ref,
                                                float* lds_dist, int stride) {
    HitRec h;
    h.t = __int_as_float(0x7f800000);
    h.id = 0xffffffffu;
    h.b1 = 0.f;
    h.b2 = 0.f;
    float best = tmax;
    const WT w = wt_setup(d);
    const f3 inv = mk3(__fdiv_rn(1.0f, d.x), __fdiv_rn(1.0f, d.y), __fdiv_rn(1.0f, d.z));
    const f3 e = mk3(slab_eps(d.x), slab_eps(d.y), slab_eps(d.z));
    uint32_t spill_ref[64 - kStackLds];
    float spill_dist[64 - kStackLds];
    int sp = 0;
    uint32_t cur = 0;  // root node
    while (true) {
        // ---- descend through internal nodes until a leaf (or nothing) is reached
        while (!ref_is_leaf(cur)) {
            const Node2 nd = nodes[cur];
            float d0 = nd.ref.x != kEmptyRef
                           ? slab<SAFE>(nd.x.x, nd.y.x, nd.z.x, nd.x.y, nd.y.y, nd.z.y, o, inv, e, best)
                           : __int_as_float(0x7f800000);
            float d1 = nd.ref.y != kEmptyRef
                           ? slab<SAFE>(nd.x.z, nd.y.z, nd.z.z, nd.x.w, nd.y.w, nd.z.w, o, inv, e, best)
                           : __int_as_float(0x7f800000);
            bool h0 = d0 < __int_as_float(0x7f800000), h1 = d1 < __int_as_float(0x7f800000);
            if (h0 && h1) {
                uint32_t nearr = nd.ref.x, farr = nd.ref.y;
                float fard = d1;
                if (d1 < d0) {
                    nearr = nd.ref.y;
                    farr = nd.ref.x;
                    fard = d0;
                }
                // push far
                int slot = sp & (kStackLds - 1);
                if (sp >= kStackLds) {
                    spill_ref[sp - kStackLds] = lds_ref[slot * stride];
                    spill_dist[sp - kStackLds] = lds_dist[slot * stride];
                }
                lds_ref[slot * stride] = farr;
                lds_dist[slot * stride] = fard;
                sp++;
                cur = nearr;
            } else if (h0) {
                cur = nd.ref.x;
            } else if (h1) {
                cur = nd.ref.y;
            } else {
                cur = kEmptyRef;
                break;
            }
        }
        // ---- leaf
        if (cur != kEmptyRef) {
            uint32_t first = leaf_first(cur), cnt = leaf_count(cur);
            for (uint32_t i = 0; i < cnt; ++i) {
                const float4 a = tris[3 * (first + i)];
                const float4 b = tris[3 * (first + i) + 1];
                const float4 c = tris[3 * (first + i) + 2];
                float t, b1, b2;
                if (!wt_intersect(o, w, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), t, b1, b2))
                    continue;
                uint32_t id = __float_as_uint(a.w);
                if (any) {
                    if (t <= tmax) {
                        h.t = t;
                        h.id = id;
                        h.b1 = b1;
                        h.b2 = b2;
                        return h;
                    }
                } else if (t <= tmax && (t < h.t || (t == h.t && id < h.id))) {
                    h.t = t;
                    h.id = id;
                    h.b1 = b1;
                    h.b2 = b2;
                    best = t;
                }
            }
        }
        // ---- pop the next entry that can still hold a closer hit
        bool found = false;
        while (sp > 0) {
            sp--;
            int slot = sp & (kStackLds - 1);
            uint32_t r = lds_ref[slot * stride];
            float dist = lds_dist[slot * stride];
            if (sp >= kStackLds) {
                lds_ref[slot * stride] = spill_ref[sp - kStackLds];
                lds_dist[slot * stride] = spill_dist[sp - kStackLds];
            }
            if (dist <= best * kSlabGrow) {
                cur = r;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return h;
}

// Rays with a (near-)zero direction component take the SAFE slab test.
__device__ __forceinline__ HitRec traverse(const Node2* __restrict__ nodes, const float4* __restrict__ tris, f3 o,
                                           f3 d, float tmax, bool any, uint32_t* lds_ref, float* lds_dist,
                                           int stride) {
    if (fabsf(d.x) < kTinyDir || fabsf(d.y) < kTinyDir || fabsf(d.z) < kTinyDir)
        return traverse_impl<true>(nodes, tris, o, d, tmax, any, lds_ref, lds_dist, stride);
    return traverse_impl<false>(nodes, tris, o, d, tmax, any, lds_ref, lds_dist, stride);
}

}  // namespace vg

// traverse.hpp -- BVH traversal with watertight intersection.
//
// Replaces rtcIntersect1 (embree_tlas.ipp:128-173) and follows the order of
// ViraBLAS::intersect (vira_blas.ipp:58-108): all child boxes of a node are
// tested, the nearest child is entered and the others are pushed far-to-near.
// Three node layouts: Node2 (binary, 64 B), Node4 (four-wide, 128 B) and
// Node4Q (four-wide with quantised boxes, 64 B; width code kWidthQ).
// The stack's top 8 entries live in LDS (one column per lane, so a wave's
// pushes and pops hit 64 consecutive dwords); deeper entries spill to private
// memory.  Capacity: a binary tree of depth < 63 (the builder's bound) needs
// at most 62 entries; its four-wide collapse at most 3 per level, 186.  On
// the 33.5M-triangle DEM frame a push past 8 entries happens on 4 in 10,000
// camera and bounce rays (the visit counters below), and 16 LDS entries were
// no faster, so the LDS they would take holds the radiance sums instead.
//
// STATS = true counts what a traversal fetches (opt-in visit counters; the
// reference's analogue is the Vulkan path tracer's debug counters,
// ray_generation.glsl:274-276): internal nodes fetched, leaves entered,
// triangles tested, stack entries moved out of LDS into private memory, and
// the deepest stack.
#pragma once

#include "device_math.hpp"

namespace vg {

constexpr int kLeafBatch = 2;  // leaf triangles fetched together (4 spilled registers and measured slower)
constexpr int kStackLds = 8;     // power of two (ring index = sp & (kStackLds - 1))
constexpr int kStackSpill = 184;  // kStackLds + kStackSpill >= 186

struct HitRec {
    float t;
    uint32_t id;
    float b1, b2;
};

// Per-lane visit counts of the rays one lane traced (STATS traversals).
struct TravCount {
    uint32_t nodes;   // internal nodes fetched (BLAS)
    uint32_t leaves;  // leaves entered
    uint32_t tris;    // triangles tested
    uint32_t spills;  // pushes that moved an entry from the LDS ring to private memory
    uint32_t maxsp;   // deepest stack
    uint32_t tlas;    // TLAS nodes whose box was tested (instanced scenes)
    uint32_t wsteps;  // node-loop iterations of the wave (counted once per iteration, by its lowest active lane)
};

struct TravStack {
    uint32_t* lds_ref;
    float* lds_dist;
    int stride;
    int sp;
    uint32_t spill_ref[kStackSpill];
    float spill_dist[kStackSpill];

    __device__ __forceinline__ void push(uint32_t r, float d) {
        int slot = sp & (kStackLds - 1);
        if (sp >= kStackLds) {  // the ring slot holds entry sp - 16: move it to private memory
            spill_ref[sp - kStackLds] = lds_ref[slot * stride];
            spill_dist[sp - kStackLds] = lds_dist[slot * stride];
        }
        lds_ref[slot * stride] = r;
        lds_dist[slot * stride] = d;
        sp++;
    }
    __device__ __forceinline__ void pop(uint32_t& r, float& d) {
        sp--;
        int slot = sp & (kStackLds - 1);
        r = lds_ref[slot * stride];
        d = lds_dist[slot * stride];
        if (sp >= kStackLds) {  // bring entry sp - 16 back into the freed ring slot
            lds_ref[slot * stride] = spill_ref[sp - kStackLds];
            lds_dist[slot * stride] = spill_dist[sp - kStackLds];
        }
    }
};

// Compare-exchange by selects (no branch: the sorting network stays straight-line code).
__device__ __forceinline__ void cswap(float& da, uint32_t& ra, float& db, uint32_t& rb) {
    const bool s = db < da;
    const float lo = s ? db : da, hi = s ? da : db;
    const uint32_t rl = s ? rb : ra, rh = s ? ra : rb;
    da = lo;
    db = hi;
    ra = rl;
    rb = rh;
}

// any: occlusion query within [0, tmax] (returns on the first hit);
// otherwise closest hit within [0, tmax], ties on t resolved to the lowest id.
template <int W, bool SAFE, bool STATS>
__device__ __forceinline__ HitRec traverse_impl(const void* __restrict__ nodes_v, const float4* __restrict__ tris,
                                                f3 o, f3 d, float tmax, bool any, uint32_t* lds_ref, float* lds_dist,
                                                int stride, TravCount* tc) {
    constexpr float kInfF = __builtin_huge_valf();
    HitRec h;
    h.t = kInfF;
    h.id = 0xffffffffu;
    h.b1 = 0.f;
    h.b2 = 0.f;
    float best = tmax;
    const WT w = wt_setup(d);
    const f3 inv = mk3(__fdiv_rn(1.0f, d.x), __fdiv_rn(1.0f, d.y), __fdiv_rn(1.0f, d.z));
    const f3 e = mk3(slab_eps(d.x), slab_eps(d.y), slab_eps(d.z));
    TravStack st;
    st.lds_ref = lds_ref;
    st.lds_dist = lds_dist;
    st.stride = stride;
    st.sp = 0;
    auto push = [&](uint32_t r, float dist) {
        if (STATS && st.sp >= kStackLds) tc->spills++;
        st.push(r, dist);
        if (STATS && (uint32_t)st.sp > tc->maxsp) tc->maxsp = (uint32_t)st.sp;
    };
    uint32_t cur = 0;  // root node
    while (true) {
        // ---- descend through internal nodes until a leaf (or nothing) is reached
        while (!ref_is_leaf(cur)) {
            if (STATS) {
                tc->nodes++;
                if (__lane_id() == (unsigned)(__ffsll((long long)__ballot(1)) - 1)) tc->wsteps++;
            }
            // Every child's slab test runs, and an empty child's result is then replaced by a
            // select: a test guarded by the child's reference made the compiler issue that
            // child's bound loads only after the references arrived (two memory round trips
            // per node) and branch around each test.
            if (W == 2) {
                const Node2 nd = reinterpret_cast<const Node2*>(nodes_v)[cur];
                float d0 = slab<SAFE>(nd.x.x, nd.y.x, nd.z.x, nd.x.y, nd.y.y, nd.z.y, o, inv, e, best);
                float d1 = slab<SAFE>(nd.x.z, nd.y.z, nd.z.z, nd.x.w, nd.y.w, nd.z.w, o, inv, e, best);
                d0 = nd.ref.x != kEmptyRef ? d0 : kInfF;
                d1 = nd.ref.y != kEmptyRef ? d1 : kInfF;
                bool h0 = d0 < kInfF, h1 = d1 < kInfF;
                if (h0 && h1) {
                    uint32_t nearr = nd.ref.x, farr = nd.ref.y;
                    float fard = d1;
                    if (d1 < d0) {
                        nearr = nd.ref.y;
                        farr = nd.ref.x;
                        fard = d0;
                    }
                    push(farr, fard);
                    cur = nearr;
                } else if (h0) {
                    cur = nd.ref.x;
                } else if (h1) {
                    cur = nd.ref.y;
                } else {
                    cur = kEmptyRef;
                    break;
                }
            } else {
                // the four children's boxes: FP32 (Node4) or decoded from bytes (Node4Q)
                float lx[4], ly[4], lz[4], hx[4], hy[4], hz[4];
                uint32_t r0, r1, r2, r3;
                if (W == kWidthQ) {
                    const Node4Q nd = reinterpret_cast<const Node4Q*>(nodes_v)[cur];
                    r0 = nd.ref.x; r1 = nd.ref.y; r2 = nd.ref.z; r3 = nd.ref.w;
                    const float sx = q_scale(nd.ex, 0), sy = q_scale(nd.ex, 1), sz = q_scale(nd.ex, 2);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        lx[k] = q_decode(nd.qlo[0], k, sx, nd.ox);
                        hx[k] = q_decode(nd.qhi[0], k, sx, nd.ox);
                        ly[k] = q_decode(nd.qlo[1], k, sy, nd.oy);
                        hy[k] = q_decode(nd.qhi[1], k, sy, nd.oy);
                        lz[k] = q_decode(nd.qlo[2], k, sz, nd.oz);
                        hz[k] = q_decode(nd.qhi[2], k, sz, nd.oz);
                    }
                } else {
                    const Node4 nd = reinterpret_cast<const Node4*>(nodes_v)[cur];
                    r0 = nd.ref.x; r1 = nd.ref.y; r2 = nd.ref.z; r3 = nd.ref.w;
                    lx[0] = nd.lox.x; lx[1] = nd.lox.y; lx[2] = nd.lox.z; lx[3] = nd.lox.w;
                    ly[0] = nd.loy.x; ly[1] = nd.loy.y; ly[2] = nd.loy.z; ly[3] = nd.loy.w;
                    lz[0] = nd.loz.x; lz[1] = nd.loz.y; lz[2] = nd.loz.z; lz[3] = nd.loz.w;
                    hx[0] = nd.hix.x; hx[1] = nd.hix.y; hx[2] = nd.hix.z; hx[3] = nd.hix.w;
                    hy[0] = nd.hiy.x; hy[1] = nd.hiy.y; hy[2] = nd.hiy.z; hy[3] = nd.hiy.w;
                    hz[0] = nd.hiz.x; hz[1] = nd.hiz.y; hz[2] = nd.hiz.z; hz[3] = nd.hiz.w;
                }
                float d0 = slab<SAFE>(lx[0], ly[0], lz[0], hx[0], hy[0], hz[0], o, inv, e, best);
                float d1 = slab<SAFE>(lx[1], ly[1], lz[1], hx[1], hy[1], hz[1], o, inv, e, best);
                float d2 = slab<SAFE>(lx[2], ly[2], lz[2], hx[2], hy[2], hz[2], o, inv, e, best);
                float d3 = slab<SAFE>(lx[3], ly[3], lz[3], hx[3], hy[3], hz[3], o, inv, e, best);
                d0 = r0 != kEmptyRef ? d0 : kInfF;
                d1 = r1 != kEmptyRef ? d1 : kInfF;
                d2 = r2 != kEmptyRef ? d2 : kInfF;
                d3 = r3 != kEmptyRef ? d3 : kInfF;
                // sorting network: d0 <= d1 <= d2 <= d3
                cswap(d0, r0, d1, r1);
                cswap(d2, r2, d3, r3);
                cswap(d0, r0, d2, r2);
                cswap(d1, r1, d3, r3);
                cswap(d1, r1, d2, r2);
                if (!(d0 < kInfF)) {
                    cur = kEmptyRef;
                    break;
                }
                if (d3 < kInfF) push(r3, d3);
                if (d2 < kInfF) push(r2, d2);
                if (d1 < kInfF) push(r1, d1);
                cur = r0;
            }
        }
        // ---- leaf: its triangles are fetched kLeafBatch at a time, every load of a batch issued
        // before the first test (one memory round trip per batch, not per triangle); slots past
        // the leaf load its last triangle again, so every load is in bounds
        if (cur != kEmptyRef) {
            const uint32_t first = leaf_first(cur), cnt = leaf_count(cur);
            if (STATS) tc->leaves++;
            for (uint32_t i0 = 0; i0 < cnt; i0 += kLeafBatch) {
                float4 ta[kLeafBatch], tb[kLeafBatch], tcv[kLeafBatch];
#pragma unroll
                for (int k = 0; k < kLeafBatch; ++k) {
                    const uint32_t i = first + min(i0 + (uint32_t)k, cnt - 1u);
                    ta[k] = tris[3 * i];
                    tb[k] = tris[3 * i + 1];
                    tcv[k] = tris[3 * i + 2];
                }
#pragma unroll
                for (int k = 0; k < kLeafBatch; ++k) {
                    if (i0 + (uint32_t)k >= cnt) break;
                    if (STATS) tc->tris++;
                    const float4 a = ta[k], b = tb[k], c = tcv[k];
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
        }
        // ---- pop the next entry that can still hold a closer hit
        bool found = false;
        while (st.sp > 0) {
            uint32_t r;
            float dist;
            st.pop(r, dist);
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
// STATS: add this ray's visits to *tc.
template <int W, bool STATS = false>
__device__ __forceinline__ HitRec traverse(const void* __restrict__ nodes, const float4* __restrict__ tris, f3 o, f3 d,
                                           float tmax, bool any, uint32_t* lds_ref, float* lds_dist, int stride,
                                           TravCount* tc = nullptr) {
    if (fabsf(d.x) < kTinyDir || fabsf(d.y) < kTinyDir || fabsf(d.z) < kTinyDir)
        return traverse_impl<W, true, STATS>(nodes, tris, o, d, tmax, any, lds_ref, lds_dist, stride, tc);
    return traverse_impl<W, false, STATS>(nodes, tris, o, d, tmax, any, lds_ref, lds_dist, stride, tc);
}

}  // namespace vg

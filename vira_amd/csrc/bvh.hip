// bvh.hip -- binned-SAH BVH build on the GPU.
//
// Replaces Mesh::buildBVH -> EmbreeBLAS (mesh.ipp:396-418,
// embree_blas.ipp:14-71) and restates the reference's own binned SAH
// (ViraBLAS, vira_blas.ipp:132-254) for a GPU:
//   * large nodes (> kSmall triangles) are split level by level, every
//     decision taken on the GPU: each node is cut into 4096-slot chunks, one
//     workgroup bins a chunk in LDS and merges it into the node's global bins,
//     one thread per node sweeps the bins (SAH), and a stable chunked
//     partition moves the triangle ids from one id buffer to the other; the
//     host waits only for each level's node count.  Every step is
//     deterministic.
//   * small nodes are finished by one wave each from LDS: the wave splits
//     nodes of more than 32 triangles together, then every lane splits the
//     smaller ones alone (binned SAH, in-place partition), the "subdivide"
//     of the reference.
//   * the binary tree is then rewritten into 64-byte nodes that hold both
//     children's boxes, and the triangles into leaf order (48 B each).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "context.hpp"

namespace vg {

void mesh_upload_pack(Ctx& c, MeshG& m, const float* pos, const float* alb, const float* uv, const float* nrm);
void compress_blas(Ctx& c, Blas& B);

namespace {

// method 0 (auto) takes the grid build for grid meshes at width 4: on the
// 33.5M-triangle DEM frame it builds in 3.5 ms against 33 ms for the binned
// SAH and renders the same image faster (fewer node visits per ray)
constexpr bool kGridByDefault = true;
constexpr uint32_t kChunk = 4096;      // slots per binning / partition workgroup
constexpr int kChunkThreads = 256;
constexpr int kChunkItems = kChunk / kChunkThreads;
constexpr uint32_t kSmall = 256;       // nodes at or below this size are built one wave each (k_build_small)
constexpr int kMaxBins = 32;
constexpr int kSmallBins = 8;
constexpr int kBinVals = 13;           // count, box lo/hi (6), centroid lo/hi (6)
constexpr int kMedianDepth = 48;       // deeper nodes are split at the object median

struct BNodeG {  // binary build node: children of an internal node are adjacent
    float lo[3], hi[3];
    uint32_t left;   // first child (internal) or first slot (leaf)
    uint32_t count;  // 0 = internal
};

// A triangle's box (24 B, read at random by the binning passes); its centroid
// is recomputed where needed as 0.5 (lo + hi), the same value every time.
struct __attribute__((aligned(8))) PrimBox {
    float lo[3], hi[3];
};
__device__ __forceinline__ float centroid(const PrimBox& b, int a) { return 0.5f * (b.lo[a] + b.hi[a]); }
// the same for a run-time axis, by selection (no indexing into registers)
__device__ __forceinline__ float centroid_on(const PrimBox& b, int a) {
    return a == 0 ? centroid(b, 0) : (a == 1 ? centroid(b, 1) : centroid(b, 2));
}
__device__ __forceinline__ float pick3(const float v[3], int a) { return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]); }

struct Chunk {
    uint32_t task, begin, end;
    uint32_t buf;  // the node's triangle ids are in buffer buf (0: prims, 1: the other buffer)
};
struct TaskParam {
    int axis;
    float cmin;
    float scale;
    int split;        // bins <= split go left (set by k_sweep_large)
    uint32_t begin;   // first slot of the node
    uint32_t chunk0;  // first chunk of the node
    uint32_t nchunk;  // chunks of the node
    uint32_t pad;
};
// The SAH decision of one binned node: split bin, left count, the children's
// boxes and centroid bounds (0 = left, 1 = right).
struct SweepOut {
    int split;  // -1: no bin boundary separates the centroids
    uint32_t nleft;
    float lo[2][3], hi[2][3];
    float clo[2][3], chi[2][3];
};
struct ChunkBase {
    uint32_t left, right;
};
struct SmallTask {
    uint32_t node, begin, count, depth;
    uint32_t base;  // first node of the task's scratch region (2 (count - 1) nodes)
    uint32_t buf;   // buffer holding the task's triangle ids (0: prims, 1: the other buffer)
};

__device__ __forceinline__ int bin_of(float c, float cmin, float scale, int nb) {
    float f = (c - cmin) * scale;
    int b = (int)f;
    b = b < 0 ? 0 : b;
    return b >= nb ? nb - 1 : b;
}

__global__ void k_prim_boxes(size_t nt, const uint32_t* idx, const GVertex* V, PrimBox* pb, uint32_t* prims) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const GVertex& a = V[idx[3 * t]];
    const GVertex& b = V[idx[3 * t + 1]];
    const GVertex& c = V[idx[3 * t + 2]];
    PrimBox o;
    o.lo[0] = fminf(fminf(a.px, b.px), c.px); o.lo[1] = fminf(fminf(a.py, b.py), c.py); o.lo[2] = fminf(fminf(a.pz, b.pz), c.pz);
    o.hi[0] = fmaxf(fmaxf(a.px, b.px), c.px); o.hi[1] = fmaxf(fmaxf(a.py, b.py), c.py); o.hi[2] = fmaxf(fmaxf(a.pz, b.pz), c.pz);
    pb[t] = o;
    prims[t] = (uint32_t)t;
}

// Root box and centroid box: LDS reduction, then one set of global atomics per block.
__global__ void __launch_bounds__(256) k_reduce_bounds(size_t n, const PrimBox* pb, int* keys) {
    __shared__ int s[12];
    if (threadIdx.x < 12) s[threadIdx.x] = (threadIdx.x % 6) < 3 ? INT_MAX : INT_MIN;
    __syncthreads();
    int l[12];
    for (int k = 0; k < 12; ++k) l[k] = (k % 6) < 3 ? INT_MAX : INT_MIN;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const PrimBox p = pb[i];
        const float cx = centroid(p, 0), cy = centroid(p, 1), cz = centroid(p, 2);
        float v[12] = {p.lo[0], p.lo[1], p.lo[2], p.hi[0], p.hi[1], p.hi[2], cx, cy, cz, cx, cy, cz};
        for (int k = 0; k < 12; ++k) {
            int key = f2key(v[k]);
            l[k] = (k % 6) < 3 ? min(l[k], key) : max(l[k], key);
        }
    }
    for (int k = 0; k < 12; ++k) {
        if ((k % 6) < 3) atomicMin(&s[k], l[k]);
        else atomicMax(&s[k], l[k]);
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        int k = threadIdx.x;
        if ((k % 6) < 3) atomicMin(&keys[k], s[k]);
        else atomicMax(&keys[k], s[k]);
    }
}

// ---------------------------------------------------------------- large nodes, level by level
// Every level runs on the GPU from device-resident counts, so the host only
// queues launches: a plan per node (axis, bins, chunks), the node's chunks,
// binning, the SAH sweep, the stable partition, and the emission of the
// children (2 per node, numbered base + 2 * node of the level) to the next
// level or to the small-subtree list.  Grids are fixed and stride over the
// level's nodes or chunks.
struct LTask {  // a node of more than kSmall triangles awaiting its split
    uint32_t node, begin, count, depth;
    float clo[3], chi[3];
    uint32_t buf;  // buffer holding the node's triangle ids (0: prims, 1: the other buffer)
};
// Device counters of the level loop (one word each).
enum : int {
    LC_NT0 = 0,      // nodes of the even levels
    LC_NT1 = 1,      // nodes of the odd levels
    LC_NCH = 2,      // chunks of the current level
    LC_NODE = 3,     // next free binary node
    LC_SMALL = 4,    // small subtrees emitted so far
    LC_LSMALL = 5,   // small subtrees emitted by the current level
    LC_DEPTH = 6,    // deepest level node
    LC_ERR = 7,      // 1: a binned node without a valid split
    LC_WORDS = 8
};
constexpr int kLevelBlocks = 1024;  // grid of the per-node and per-chunk passes

// Root node and root task from the reduced bounds (keys: box lo/hi, centroid lo/hi).
__global__ void k_lv_root(const int* keys, uint32_t n, BNodeG* nodes, LTask* task, SmallTask* small,
                          uint32_t* cnt) {
    BNodeG r;
    LTask t;
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = key2f(keys[k]);
        r.hi[k] = key2f(keys[3 + k]);
        t.clo[k] = key2f(keys[6 + k]);
        t.chi[k] = key2f(keys[9 + k]);
    }
    r.left = 0;
    r.count = n;
    nodes[0] = r;
    t.node = 0;
    t.begin = 0;
    t.count = n;
    t.depth = 0;
    t.buf = 0;
    const bool large = n > kSmall;
    if (large) task[0] = t;
    else small[0] = SmallTask{0u, 0u, n, 0u, 0u, 0u};
    for (int k = 0; k < LC_WORDS; ++k) cnt[k] = 0;
    cnt[LC_NT0] = large ? 1u : 0u;
    cnt[LC_NODE] = 1;
    cnt[LC_SMALL] = large ? 0u : 1u;
}

// Per node: axis, bin mapping and chunk count; nodes whose centroids coincide
// (or that are too deep) get no chunks and are split at the object median.
__global__ void k_lv_plan(const LTask* __restrict__ tasks, const uint32_t* __restrict__ d_nt, int nb,
                          TaskParam* __restrict__ params, uint32_t* __restrict__ nchunk, uint32_t* cnt) {
    const uint32_t nt = *d_nt;
    uint32_t md = 0;
    for (uint32_t ti = blockIdx.x * blockDim.x + threadIdx.x; ti < nt; ti += gridDim.x * blockDim.x) {
        const LTask t = tasks[ti];
        md = max(md, t.depth);
        const float ext[3] = {t.chi[0] - t.clo[0], t.chi[1] - t.clo[1], t.chi[2] - t.clo[2]};
        int axis = (ext[1] > ext[0]) ? 1 : 0;
        if (ext[2] > pick3(ext, axis)) axis = 2;
        const float scale = (float)nb / pick3(ext, axis);
        const bool median = !(pick3(ext, axis) > 0.f) || !isfinite(scale) || t.depth >= (uint32_t)kMedianDepth;
        const uint32_t nch = median ? 0u : (t.count + kChunk - 1) / kChunk;
        params[ti] = TaskParam{axis, pick3(t.clo, axis), scale, -1, t.begin, 0u, nch, 0u};
        nchunk[ti] = nch;
    }
    for (int off = 32; off > 0; off >>= 1) md = max(md, (uint32_t)__shfl_xor((int)md, off));
    if ((threadIdx.x & 63) == 0 && md > 0) atomicMax(&cnt[LC_DEPTH], md);
}

// Each node's chunks (one wave per node): first chunk from the scan of the counts.
__global__ void k_lv_chunks(const LTask* __restrict__ tasks, const uint32_t* __restrict__ d_nt,
                            TaskParam* __restrict__ params, const uint32_t* __restrict__ chunk0,
                            Chunk* __restrict__ chunks) {
    const uint32_t nt = *d_nt, lane = threadIdx.x & 63;
    const uint32_t nwaves = gridDim.x * (blockDim.x / 64);
    for (uint32_t ti = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); ti < nt; ti += nwaves) {
        const uint32_t c0 = chunk0[ti], nch = params[ti].nchunk;
        const uint32_t b = tasks[ti].begin, e = b + tasks[ti].count;
        if (lane == 0) params[ti].chunk0 = c0;
        for (uint32_t k = lane; k < nch; k += 64)
            chunks[c0 + k] = Chunk{ti, b + k * kChunk, min(b + (k + 1) * kChunk, e), tasks[ti].buf};
    }
}

__global__ void k_init_bins(int* bins, const uint32_t* d_nt, int nb) {
    const size_t nvals = (size_t)(*d_nt) * nb * kBinVals;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvals; i += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(i % kBinVals);
        bins[i] = v == 0 ? 0 : (((v - 1) % 6) < 3 ? INT_MAX : INT_MIN);
    }
}

// Binning of the level's chunks: a workgroup bins its chunks into LDS and
// merges them into their nodes' global bins; the bin of every slot and each
// chunk's count per bin are kept for the partition.  Each wave takes 1024
// slots of a chunk in stages
// of 256: the lanes load slots k*64 + lane (consecutive lanes, consecutive
// slots, so the loads of the ids and of the mostly consecutive boxes
// coalesce), bin them and stage the boxes in LDS; then every lane reads back
// kStageItems consecutive slots, whose triangles are mostly neighbours in the
// same bin, accumulates each run of equal bins in registers and adds a run to
// the LDS bins with one set of atomics, so neither lanes of a wave nor slots
// of a thread contend for one bin address slot by slot.  The stages are
// pipelined: the next stage's boxes and the one after's ids are loaded while
// a stage is accumulated.
constexpr int kStageItems = 4;                           // consecutive slots per lane and stage
constexpr int kStageSlots = kStageItems * 64;            // slots per wave and stage
constexpr int kStageWords = kStageSlots + kStageSlots / 32;
// LDS word of staged slot p: one word of skew per 32 slots, so that lanes
// reading slots lane * kStageItems + j hit distinct banks
__device__ __forceinline__ int stage_word(int p) { return p + (p >> 5); }

// A run of acc[0] slots of bin b: its count to the chunk's counts, its box and
// centroid keys to the node's bins.
__device__ __forceinline__ void bin_flush(int* s, int* s_cnt, int b, const int acc[kBinVals]) {
    int* sb = &s[b * kBinVals];
    atomicAdd(&s_cnt[b], acc[0]);
#pragma unroll
    for (int k = 1; k < kBinVals; ++k) {
        if (((k - 1) % 6) < 3) atomicMin(&sb[k], acc[k]);
        else atomicMax(&sb[k], acc[k]);
    }
}

// The block's bins of one node into the node's global bins (after a barrier).
__device__ __forceinline__ void bins_to_node(const int* s, int nb, int* node_bins) {
    for (int i = threadIdx.x; i < nb * kBinVals; i += blockDim.x) {
        const int v = i % kBinVals;
        if (s[i - v] == 0) continue;  // empty bin
        if (v == 0) atomicAdd(&node_bins[i], s[i]);
        else if (((v - 1) % 6) < 3) atomicMin(&node_bins[i], s[i]);
        else atomicMax(&node_bins[i], s[i]);
    }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A workgroup takes a contiguous range of the level's chunks, so the chunks
// of a node of many chunks mostly share a workgroup: it keeps the node's bins
// in LDS across them and merges them into the node's global bins once per
// node (the top levels' few nodes would otherwise take one set of atomics per
// chunk on the same addresses).
__global__ void __launch_bounds__(kChunkThreads) k_bin_large(const Chunk* chunks, const uint32_t* d_nch,
                                                             const TaskParam* params, int nb, const uint32_t* ids0,
                                                             const uint32_t* ids1, const PrimBox* pb, int* bins,
                                                             uint32_t* chunk_counts, uint8_t* slot_bin) {
    __shared__ int s[kMaxBins * kBinVals];  // the current node's bins
    __shared__ int s_cnt[kMaxBins];         // the current chunk's counts
    __shared__ float s_box[kChunkThreads / 64][6][kStageWords];  // per wave: lo x,y,z, hi x,y,z
    __shared__ uint8_t s_bin[kChunkThreads / 64][kStageSlots];
    const uint32_t nch = *d_nch;
    const uint32_t c_first = (uint32_t)(((uint64_t)nch * blockIdx.x) / gridDim.x);
    const uint32_t c_last = (uint32_t)(((uint64_t)nch * (blockIdx.x + 1)) / gridDim.x);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    float (*bx)[kStageWords] = s_box[w];
    uint8_t* bn = s_bin[w];
    uint32_t node = 0xffffffffu;  // task whose bins s holds
    for (uint32_t ci = c_first; ci < c_last; ++ci) {
        const Chunk ch = chunks[ci];
        block_sync();  // the previous chunk's bins and counts are complete
        if (ch.task != node) {  // the same for the whole block
            if (node != 0xffffffffu) {
                bins_to_node(s, nb, &bins[(size_t)node * nb * kBinVals]);
                block_sync();
            }
            for (int i = threadIdx.x; i < nb * kBinVals; i += blockDim.x) {
                const int v = i % kBinVals;
                s[i] = v == 0 ? 0 : (((v - 1) % 6) < 3 ? INT_MAX : INT_MIN);
            }
            node = ch.task;
        }
        if (threadIdx.x < (unsigned)nb) s_cnt[threadIdx.x] = 0;
        block_sync();
        const TaskParam tp = params[ch.task];
        const uint32_t* prims = ch.buf ? ids1 : ids0;
        int rb = -1;  // bin of the run being accumulated
        int acc[kBinVals];
#pragma unroll
        for (int k = 0; k < kBinVals; ++k) acc[k] = 0;
        // the chunk's slots split over the waves in whole stages, so a small chunk keeps every wave busy
        const uint32_t per = (ch.end - ch.begin + (kChunkThreads / 64) * kStageSlots - 1) /
                             ((kChunkThreads / 64) * kStageSlots) * kStageSlots;
        const uint32_t wfirst = ch.begin + (uint32_t)w * per;
        const uint32_t wend = min(wfirst + per, ch.end);
        constexpr int kStages = kChunkItems / kStageItems;
        uint32_t idn[kStageItems];
        PrimBox pn[kStageItems];
#pragma unroll
        for (int k = 0; k < kStageItems; ++k) {
            const uint32_t slot = wfirst + k * 64 + lane;
            idn[k] = slot < wend ? prims[slot] : 0u;
        }
#pragma unroll
        for (int k = 0; k < kStageItems; ++k)
            if (wfirst + k * 64 + lane < wend) pn[k] = pb[idn[k]];
#pragma unroll
        for (int k = 0; k < kStageItems; ++k) {
            const uint32_t slot = wfirst + kStageSlots + k * 64 + lane;
            idn[k] = slot < wend ? prims[slot] : 0u;
        }
        for (int st = 0; st < kStages; ++st) {
            const uint32_t wbase = wfirst + (uint32_t)st * kStageSlots;
            if (wbase >= wend) break;  // the same for the whole wave
            PrimBox p[kStageItems];
#pragma unroll
            for (int k = 0; k < kStageItems; ++k) p[k] = pn[k];
#pragma unroll
            for (int k = 0; k < kStageItems; ++k) {
                const uint32_t slot = wbase + k * 64 + lane;
                if (slot < wend) {
                    const float c[3] = {centroid(p[k], 0), centroid(p[k], 1), centroid(p[k], 2)};
                    const int b = bin_of(pick3(c, tp.axis), tp.cmin, tp.scale, nb);
                    slot_bin[slot] = (uint8_t)b;
                    const int q = stage_word(k * 64 + lane);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        bx[a][q] = p[k].lo[a];
                        bx[3 + a][q] = p[k].hi[a];
                    }
                    bn[k * 64 + lane] = (uint8_t)b;
                }
            }
            // the next stage's boxes and the one after's ids are in flight while this one is accumulated
            if (st + 1 < kStages) {
#pragma unroll
                for (int k = 0; k < kStageItems; ++k)
                    if (wbase + kStageSlots + k * 64 + lane < wend) pn[k] = pb[idn[k]];
#pragma unroll
                for (int k = 0; k < kStageItems; ++k) {
                    const uint32_t slot = wbase + 2 * kStageSlots + k * 64 + lane;
                    idn[k] = (st + 2 < kStages && slot < wend) ? prims[slot] : 0u;
                }
            }
            wave_sync();
#pragma unroll
            for (int j = 0; j < kStageItems; ++j) {
                const int pos = lane * kStageItems + j;
                if (wbase + pos >= wend) break;
                const int q = stage_word(pos);
                PrimBox pj;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    pj.lo[a] = bx[a][q];
                    pj.hi[a] = bx[3 + a][q];
                }
                const int b = bn[pos];
                int v[kBinVals];
                v[0] = 1;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float c = centroid(pj, a);
                    v[1 + a] = f2key(pj.lo[a]);
                    v[4 + a] = f2key(pj.hi[a]);
                    v[7 + a] = f2key(c);
                    v[10 + a] = f2key(c);
                }
                if (b != rb) {
                    if (rb >= 0) bin_flush(s, s_cnt, rb, acc);
                    rb = b;
#pragma unroll
                    for (int q2 = 0; q2 < kBinVals; ++q2) acc[q2] = v[q2];
                } else {
                    acc[0] += 1;
#pragma unroll
                    for (int q2 = 1; q2 < kBinVals; ++q2)
                        acc[q2] = ((q2 - 1) % 6) < 3 ? min(acc[q2], v[q2]) : max(acc[q2], v[q2]);
                }
            }
            wave_sync();  // the stage is read before the next one overwrites it
        }
        if (rb >= 0) bin_flush(s, s_cnt, rb, acc);
        block_sync();
        if (threadIdx.x < (unsigned)nb) {
            const int c = s_cnt[threadIdx.x];
            chunk_counts[(size_t)ci * nb + threadIdx.x] = (uint32_t)c;
            s[threadIdx.x * kBinVals] += c;
        }
    }
    block_sync();
    if (node != 0xffffffffu) bins_to_node(s, nb, &bins[(size_t)node * nb * kBinVals]);
}

// number of set bits of m below this lane
__device__ __forceinline__ uint32_t rank_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Stable partition of one chunk (striding over the level's chunks) from the
// node's buffer into the other one: lefts to the chunk's left base, rights to
// its right base, in slot order.  Wave w takes the chunk's slots
// w*1024 + k*64 + lane (coalesced); a slot's rank among the lefts is the
// lefts of the waves before, of the steps before and of the lanes below it in
// its step's ballot.
__global__ void __launch_bounds__(kChunkThreads) k_partition_large(const Chunk* chunks, const uint32_t* d_nch,
                                                                   const TaskParam* params, const ChunkBase* bases,
                                                                   uint32_t* ids0, uint32_t* ids1,
                                                                   const uint8_t* slot_bin) {
    __shared__ uint32_t lds[kChunkThreads / 64];
    const uint32_t nch = *d_nch;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t ci = blockIdx.x; ci < nch; ci += gridDim.x) {
        const Chunk ch = chunks[ci];
        const TaskParam tp = params[ch.task];
        const ChunkBase cb = bases[ci];
        const uint32_t* prims = ch.buf ? ids1 : ids0;
        uint32_t* tmp = ch.buf ? ids0 : ids1;
        // the chunk's slots split over the waves in whole steps of 64, so a small chunk keeps every
        // wave busy; wfirst: chunk-local index of the wave's first slot
        const uint32_t per = (ch.end - ch.begin + (kChunkThreads / 64) * 64 - 1) / ((kChunkThreads / 64) * 64) * 64;
        const uint32_t wfirst = (uint32_t)w * per;
        const uint32_t wend = min(ch.begin + wfirst + per, ch.end);
        const uint32_t s0 = ch.begin + wfirst + lane;
        uint32_t ids[kChunkItems];
        uint64_t lm[kChunkItems];  // lefts of each step
        uint32_t nl = 0;
#pragma unroll
        for (int k = 0; k < kChunkItems; ++k) {
            const uint32_t slot = s0 + k * 64;
            bool left = false;
            ids[k] = 0;
            if (slot < wend) {
                ids[k] = prims[slot];
                left = (int)slot_bin[slot] <= tp.split;
            }
            lm[k] = __ballot(left);
            nl += __popcll(lm[k]);
        }
        if (lane == 0) lds[w] = nl;
        __syncthreads();
        uint32_t lbefore = 0;
        for (int q = 0; q < w; ++q) lbefore += lds[q];
#pragma unroll
        for (int k = 0; k < kChunkItems; ++k) {
            const uint32_t slot = s0 + k * 64;
            if (slot < wend) {
                const uint32_t l = lbefore + rank_below(lm[k]);
                if ((lm[k] >> lane) & 1ull) tmp[cb.left + l] = ids[k];
                else tmp[cb.right + (wfirst + k * 64 + lane - l)] = ids[k];
            }
            lbefore += __popcll(lm[k]);
        }
        __syncthreads();  // lds is rewritten by the next chunk
    }
}

__device__ __forceinline__ float half_area(const float lo[3], const float hi[3]) {
    float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    return ex * ey + ey * ez + ez * ex;
}

// SAH sweep of one binned node's bins (vira_blas.ipp:185-254 with bins):
// right-to-left areas, then the cheapest left-to-right boundary; the split is
// written to the node's TaskParam for the partition, and the children's boxes
// and centroid bounds are gathered from the bins.  A node without chunks is
// split at the object median: both children keep its box and centroid bounds.
// Also counts the node's children that stay large (nlarge) and that go to the
// small-subtree builder (nsmall).
__global__ void k_sweep_large(const int* __restrict__ bins, int nb, const LTask* __restrict__ tasks,
                              const uint32_t* __restrict__ d_nt, const BNodeG* __restrict__ nodes, TaskParam* params,
                              SweepOut* out, uint32_t* nlarge, uint32_t* nsmall, uint32_t* cnt) {
    const uint32_t nt = *d_nt;
    for (uint32_t ti = blockIdx.x * blockDim.x + threadIdx.x; ti < nt; ti += gridDim.x * blockDim.x) {
        const LTask t = tasks[ti];
        SweepOut o;
        if (params[ti].nchunk == 0) {
            const BNodeG par = nodes[t.node];
            o.split = -1;
            o.nleft = t.count / 2;
            for (int h = 0; h < 2; ++h)
                for (int k = 0; k < 3; ++k) {
                    o.lo[h][k] = par.lo[k];
                    o.hi[h][k] = par.hi[k];
                    o.clo[h][k] = t.clo[k];
                    o.chi[h][k] = t.chi[k];
                }
        } else {
            const int* tb = bins + (size_t)ti * nb * kBinVals;
            float rarea[kMaxBins];
            uint32_t rcnt[kMaxBins];
            float rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
            uint32_t rc = 0;
            for (int b = nb - 1; b > 0; --b) {
                const int* bb = tb + b * kBinVals;
                rc += (uint32_t)bb[0];
                if (bb[0] > 0)
                    for (int k = 0; k < 3; ++k) {
                        rl[k] = fminf(rl[k], key2f(bb[1 + k]));
                        rh[k] = fmaxf(rh[k], key2f(bb[4 + k]));
                    }
                rarea[b] = rc ? half_area(rl, rh) : 0.f;
                rcnt[b] = rc;
            }
            float ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY};
            uint32_t lc = 0;
            float best = INFINITY;
            int split = -1;
            for (int b = 0; b < nb - 1; ++b) {
                const int* bb = tb + b * kBinVals;
                lc += (uint32_t)bb[0];
                if (bb[0] > 0)
                    for (int k = 0; k < 3; ++k) {
                        ll[k] = fminf(ll[k], key2f(bb[1 + k]));
                        lh[k] = fmaxf(lh[k], key2f(bb[4 + k]));
                    }
                if (lc == 0 || rcnt[b + 1] == 0) continue;
                const float cost = (float)lc * half_area(ll, lh) + (float)rcnt[b + 1] * rarea[b + 1];
                if (cost < best) {
                    best = cost;
                    split = b;
                }
            }
            params[ti].split = split;
            for (int h = 0; h < 2; ++h)
                for (int k = 0; k < 3; ++k) {
                    o.lo[h][k] = o.clo[h][k] = INFINITY;
                    o.hi[h][k] = o.chi[h][k] = -INFINITY;
                }
            uint32_t nl = 0;
            for (int b = 0; b < nb; ++b) {
                const int* bb = tb + b * kBinVals;
                if (bb[0] == 0) continue;
                const int h = b <= split ? 0 : 1;
                if (h == 0) nl += (uint32_t)bb[0];
                for (int k = 0; k < 3; ++k) {
                    o.lo[h][k] = fminf(o.lo[h][k], key2f(bb[1 + k]));
                    o.hi[h][k] = fmaxf(o.hi[h][k], key2f(bb[4 + k]));
                    o.clo[h][k] = fminf(o.clo[h][k], key2f(bb[7 + k]));
                    o.chi[h][k] = fmaxf(o.chi[h][k], key2f(bb[10 + k]));
                }
            }
            o.split = split;
            o.nleft = nl;
        }
        out[ti] = o;
        uint32_t nbig = 0, nsm = 0;
        if (params[ti].nchunk != 0 && o.split < 0) {
            atomicOr(&cnt[LC_ERR], 1u);  // no bin boundary separates the centroids: the node is dropped
        } else {
            const uint32_t nl = o.nleft, nr = t.count - o.nleft;
            nbig = (nl > kSmall ? 1u : 0u) + (nr > kSmall ? 1u : 0u);
            nsm = 2u - nbig;
        }
        nlarge[ti] = nbig;
        nsmall[ti] = nsm;
    }
}

// Per chunk: how many of its triangles go left (bins <= the node's split).
__global__ void k_chunk_left(const Chunk* __restrict__ chunks, const uint32_t* __restrict__ d_nch,
                             const TaskParam* __restrict__ params, const uint32_t* __restrict__ counts, int nb,
                             uint32_t* left) {
    const uint32_t nch = *d_nch;
    for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < nch; ci += gridDim.x * blockDim.x) {
        const int split = params[chunks[ci].task].split;
        uint32_t cl = 0;
        for (int b = 0; b <= split; ++b) cl += counts[(size_t)ci * nb + b];
        left[ci] = cl;
    }
}

// Where each chunk writes its left and right triangles in the stable
// partition: the node's chunks in order, lefts from the node's first slot,
// rights after all its lefts.  lscan = exclusive scan of k_chunk_left over all
// chunks; a node's chunks are consecutive, so the lefts before chunk ci within
// its node are lscan[ci] - lscan[first chunk], and the rights the remaining
// slots before it.
__global__ void k_chunk_bases(const Chunk* __restrict__ chunks, const uint32_t* __restrict__ d_nch,
                              const TaskParam* __restrict__ params, const SweepOut* __restrict__ sw,
                              const uint32_t* __restrict__ lscan, ChunkBase* bases) {
    const uint32_t nch = *d_nch;
    for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < nch; ci += gridDim.x * blockDim.x) {
        const Chunk ch = chunks[ci];
        const TaskParam tp = params[ch.task];
        const uint32_t lrun = lscan[ci] - lscan[tp.chunk0];
        const uint32_t rrun = (ch.begin - tp.begin) - lrun;
        bases[ci] = ChunkBase{tp.begin + lrun, tp.begin + sw[ch.task].nleft + rrun};
    }
}

// The children of every node of the level: binary nodes base + 2 ti and
// base + 2 ti + 1 (boxes from the sweep, triangle ranges as leaves until
// split), the parent turned internal, and each child queued as a large node of
// the next level or as a small subtree (positions from the scans of nlarge and
// nsmall).  The children's ids are in the other buffer after a binned split
// (the partition moved them) and still in the node's after an object-median
// split.
__global__ void k_lv_emit(const LTask* __restrict__ tasks, const uint32_t* __restrict__ d_nt,
                          const TaskParam* __restrict__ params, const SweepOut* __restrict__ sw, const uint32_t* __restrict__ nlarge,
                          const uint32_t* __restrict__ nsmall, const uint32_t* __restrict__ loff,
                          const uint32_t* __restrict__ soff, BNodeG* nodes,
                          LTask* __restrict__ next, SmallTask* __restrict__ small, const uint32_t* __restrict__ cnt) {
    const uint32_t nt = *d_nt, base = cnt[LC_NODE], sbase = cnt[LC_SMALL];
    for (uint32_t ti = blockIdx.x * blockDim.x + threadIdx.x; ti < nt; ti += gridDim.x * blockDim.x) {
        const LTask t = tasks[ti];
        const SweepOut o = sw[ti];
        if (nlarge[ti] + nsmall[ti] == 0) continue;  // dropped by the sweep (error flagged)
        const uint32_t c = base + 2 * ti;
        const uint32_t nl = o.nleft, nr = t.count - o.nleft;
        BNodeG L, R;
        for (int k = 0; k < 3; ++k) {
            L.lo[k] = o.lo[0][k];
            L.hi[k] = o.hi[0][k];
            R.lo[k] = o.lo[1][k];
            R.hi[k] = o.hi[1][k];
        }
        L.left = t.begin;
        L.count = nl;
        R.left = t.begin + nl;
        R.count = nr;
        nodes[c] = L;
        nodes[c + 1] = R;
        nodes[t.node].left = c;
        nodes[t.node].count = 0;
        uint32_t lp = loff[ti], sp = sbase + soff[ti];
        const uint32_t depth = t.depth + 1;
        const uint32_t buf = params[ti].nchunk > 0 ? t.buf ^ 1u : t.buf;
        for (int h = 0; h < 2; ++h) {
            const uint32_t n = h == 0 ? nl : nr;
            const uint32_t b = h == 0 ? t.begin : t.begin + nl;
            if (n > kSmall) {
                LTask x;
                x.node = c + h;
                x.begin = b;
                x.count = n;
                x.depth = depth;
                x.buf = buf;
                for (int k = 0; k < 3; ++k) {
                    x.clo[k] = o.clo[h][k];
                    x.chi[k] = o.chi[h][k];
                }
                next[lp++] = x;
            } else {
                small[sp++] = SmallTask{c + h, b, n, depth, 2 * b, buf};
            }
        }
    }
}

// After a level: its nodes' children are numbered, its small subtrees counted
// and their range recorded (rec[0] first, rec[1] count) for their builder.
__global__ void k_lv_advance(uint32_t* cnt, int parity, uint32_t* rec) {
    rec[0] = cnt[LC_SMALL];
    rec[1] = cnt[LC_LSMALL];
    cnt[LC_NODE] += 2 * cnt[parity ? LC_NT1 : LC_NT0];
    cnt[LC_SMALL] += cnt[LC_LSMALL];
}

// Small subtrees (nodes of at most kSmall triangles; the reference's
// recursive "subdivide", vira_blas.ipp:132-182), one wave per subtree.  The
// wave stages its triangles' boxes in LDS once and works from LDS only:
//   * nodes of more than kCoop triangles are split by the whole wave, up to
//     kFront of them per round (breadth first): their items are binned into
//     LDS with atomics (8 bins of order-preserving integer keys per node),
//     lane f sweeps node f's bins with the same arithmetic as a serial build,
//     a stable ballot partition reorders each node's items, and the
//     children's centroid bounds come from the bins;
//   * queued nodes of kGroupMin..kCoop triangles are split by groups of 8
//     lanes, 8 nodes per round: lane r of a group holds bin r, every lane
//     reads each item once, the sweep runs as prefix unions across the
//     group, and the group partitions the items stably by its ballots;
//   * smaller queued nodes are split one per lane (bins in registers, the
//     same in-place partition).
// Child pairs are numbered by prefix sums over a round's splitting nodes, so
// the nodes of a task fill [0, used) of its scratch region without gaps, and
// k_compact_small moves them behind the large-level nodes.  Every decision is
// an exact function of a node's triangle set (integer counts, min/max bounds,
// one fixed sweep), so the tree is the same as a serial build's whatever the
// order in which nodes are split; the item order inside a node, which the
// object-median fallback and the leaves' triangle order follow, depends only
// on which partition (stable, or the lanes' in-place one) a node of its size
// takes.
constexpr uint32_t kTaskRoot = 0xffffffffu;
constexpr uint32_t kCoop = 32;              // nodes above this are split by the whole wave
constexpr uint32_t kQueueCap = kSmall / 2;  // queued nodes hold >= 2 triangles each, disjoint
constexpr uint32_t kGroupMin = 9;           // queued nodes of kGroupMin..kCoop triangles: 8 lanes each
constexpr uint32_t kBigCap = kSmall / kGroupMin + 1;  // such nodes are disjoint
constexpr int kFront = 4;                   // whole-wave nodes split per round
constexpr uint32_t kFrontCap = kSmall / (kCoop + 1) + 1;  // whole-wave nodes are disjoint
static_assert(kSmall <= 256, "item numbers of a subtree are bytes");
static_assert(kCoop > 15, "wave-split nodes are never leaves (max_leaf <= 15)");
static_assert(kSmallBins == 8, "the wave-split binning puts one bin on each lane of a group of 8");

// A node's box is not queued: it is the union of the boxes of the items in
// its box range, which is the node's own item range for a child of a binned
// split (its box is the union of its bins', i.e. of its items') and its
// parent's box range for a child of an object-median split (whose box is its
// parent's); root: the box is the task root's (from the large levels).
struct BoxRef {
    uint32_t b, n;
    bool root;
};
// c ? a : b, field by field (a select of whole structs would go through scratch memory)
__device__ __forceinline__ BoxRef pick_ref(bool c, const BoxRef& a, const BoxRef& b) {
    return BoxRef{c ? a.b : b.b, c ? a.n : b.n, c ? a.root : b.root};
}
struct QEnt {  // queued node: slot (kTaskRoot or a scratch index), first item, items, depth, box range; the
               // box, filled in from the box range by the node's splitter
    uint32_t slot, b, n, depth;
    BoxRef ref;
    float lo[3], hi[3];
};
struct CEnt {  // wave-split node: slot, first item, items, depth, box range (packed), centroid bounds
    uint32_t slot, b, n, depth;
    uint32_t ref;  // b | n << 8 | root << 17
    float clo[3], chi[3];
};
__device__ __forceinline__ uint32_t ref_pack(const BoxRef& r) { return r.b | (r.n << 8) | (r.root ? 1u << 17 : 0u); }
__device__ __forceinline__ BoxRef ref_unpack(uint32_t w) { return BoxRef{w & 0xffu, (w >> 8) & 0x1ffu, ((w >> 17) & 1u) != 0u}; }

// A queued node as kept in LDS (8 bytes): slot (10 bits, the task root as all ones), first item,
// items and the root flag in w0; depth (14 bits) and the box range in w1.
struct QEntL {
    uint32_t w0, w1;
};
constexpr uint32_t kSlotRootPacked = 0x3ffu;
__device__ __forceinline__ QEntL q_pack(uint32_t slot, uint32_t b, uint32_t n, uint32_t depth, const BoxRef& r) {
    const uint32_t ps = slot == kTaskRoot ? kSlotRootPacked : slot;
    return QEntL{ps | (b << 10) | (n << 18) | (r.root ? 1u << 27 : 0u), depth | (r.b << 14) | (r.n << 22)};
}
__device__ __forceinline__ QEnt q_unpack(const QEntL& l) {
    QEnt q;
    const uint32_t ps = l.w0 & 0x3ffu;
    q.slot = ps == kSlotRootPacked ? kTaskRoot : ps;
    q.b = (l.w0 >> 10) & 0xffu;
    q.n = (l.w0 >> 18) & 0x1ffu;
    q.depth = l.w1 & 0x3fffu;
    q.ref = BoxRef{(l.w1 >> 14) & 0xffu, (l.w1 >> 22) & 0x1ffu, ((l.w0 >> 27) & 1u) != 0u};
    for (int k = 0; k < 3; ++k) q.lo[k] = q.hi[k] = 0.0f;
    return q;
}

// Item j's box, kept in LDS as three float2 (24 bytes): (lo.x, lo.y), (lo.z, hi.x), (hi.y, hi.z).
struct Box6 {
    float lo[3], hi[3];
};
__device__ __forceinline__ Box6 box_at(const float2* s_box, uint32_t j) {
    const float2 a = s_box[3 * j], b = s_box[3 * j + 1], c = s_box[3 * j + 2];
    return Box6{{a.x, a.y, b.x}, {b.y, c.x, c.y}};
}

// Centroid bounds of items [b, b + n) of the wave's list (every lane gets them).
// A whole-wave node's bin mapping from its centroid bounds: the axis of
// largest extent, its low bound and bins per unit; not binned (object
// median) when the centroids coincide on it or the node is too deep.
struct WaveBinning {
    bool binned;
    int axis;
    float cmin, scale;
};
__device__ __forceinline__ WaveBinning wave_binning(const float clo[3], const float chi[3], uint32_t depth) {
    const float ext[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
    int axis = (ext[1] > ext[0]) ? 1 : 0;
    if (ext[2] > pick3(ext, axis)) axis = 2;
    const float eaxis = pick3(ext, axis);
    const float scale = (float)kSmallBins / eaxis;
    return WaveBinning{eaxis > 0.f && isfinite(scale) && depth < kMedianDepth, axis, pick3(clo, axis), scale};
}

__device__ __forceinline__ void wave_centroid_bounds(const uint8_t* s_perm, const float2* s_box, uint32_t b,
                                                     uint32_t n, uint32_t lane,
                                                     float clo[3], float chi[3]) {
    float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = lane; i < n; i += 64) {
        const Box6 x = box_at(s_box, s_perm[b + i]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float c = 0.5f * (x.lo[k] + x.hi[k]);
            l[k] = fminf(l[k], c);
            h[k] = fmaxf(h[k], c);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            l[k] = fminf(l[k], __shfl_xor(l[k], off));
            h[k] = fmaxf(h[k], __shfl_xor(h[k], off));
        }
        clo[k] = l[k];
        chi[k] = h[k];
    }
}

// SAH sweep over 8 bins (vira_blas.ipp:185-254 with bins): right-to-left
// areas, then the cheapest left-to-right boundary; split = -1 when no
// boundary has triangles on both sides.
__device__ __forceinline__ void sweep8(const uint32_t bcnt[kSmallBins], const float blo[kSmallBins][3],
                                       const float bhi[kSmallBins][3], int& split, float& best) {
    float rarea[kSmallBins];
    uint32_t rcnt[kSmallBins];
    float rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t rc = 0;
#pragma unroll
    for (int b = kSmallBins - 1; b > 0; --b) {
        rc += bcnt[b];
        for (int k = 0; k < 3; ++k) { rl[k] = fminf(rl[k], blo[b][k]); rh[k] = fmaxf(rh[k], bhi[b][k]); }
        rarea[b] = rc ? half_area(rl, rh) : 0.f;
        rcnt[b] = rc;
    }
    float ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t lc = 0;
    best = INFINITY;
    split = -1;
#pragma unroll
    for (int b = 0; b < kSmallBins - 1; ++b) {
        lc += bcnt[b];
        for (int k = 0; k < 3; ++k) { ll[k] = fminf(ll[k], blo[b][k]); lh[k] = fmaxf(lh[k], bhi[b][k]); }
        if (lc == 0 || rcnt[b + 1] == 0) continue;
        const float cost = (float)lc * half_area(ll, lh) + (float)rcnt[b + 1] * rarea[b + 1];
        if (cost < best) { best = cost; split = b; }
    }
}

// The children's boxes (and counts) of a binned split: the union of the bins
// on each side.
__device__ __forceinline__ void split_boxes8(const uint32_t bcnt[kSmallBins], const float blo[kSmallBins][3],
                                             const float bhi[kSmallBins][3], int split, uint32_t& nl, float Llo[3],
                                             float Lhi[3], float Rlo[3], float Rhi[3]) {
    nl = 0;
    for (int k = 0; k < 3; ++k) { Llo[k] = Rlo[k] = INFINITY; Lhi[k] = Rhi[k] = -INFINITY; }
#pragma unroll
    for (int b = 0; b < kSmallBins; ++b) {
        const bool left = b <= split;
        nl += left ? bcnt[b] : 0u;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            Llo[k] = left ? fminf(Llo[k], blo[b][k]) : Llo[k];
            Lhi[k] = left ? fmaxf(Lhi[k], bhi[b][k]) : Lhi[k];
            Rlo[k] = left ? Rlo[k] : fminf(Rlo[k], blo[b][k]);
            Rhi[k] = left ? Rhi[k] : fmaxf(Rhi[k], bhi[b][k]);
        }
    }
}

// One lane decides and splits queued node q alone (items in LDS).  Returns
// false when q stays a leaf; else nl and the children's boxes are set and the
// node's items are partitioned in place (vira_blas.ipp:149-159).  The bins
// stay in registers (every bin updated under a compare with its compile-time
// index, no indexing into registers); each item's bin is kept in s_binof for
// the partition.
__device__ __forceinline__ bool lane_split(QEnt& q, uint8_t* s_perm, const float2* s_box, uint8_t* s_binof,
                                           const float root_lo[3], const float root_hi[3], int max_leaf,
                                           float trav_cost, uint32_t& nl, float Llo[3], float Lhi[3], float Rlo[3],
                                           float Rhi[3], bool& median) {
    const uint32_t n = q.n;
    const bool own = !q.ref.root && q.ref.b == q.b && q.ref.n == q.n;  // the box is the items' union
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
    for (int k = 0; k < 3; ++k) { q.lo[k] = INFINITY; q.hi[k] = -INFINITY; }
    for (uint32_t i = 0; i < n; ++i) {
        const Box6 x = box_at(s_box, s_perm[q.b + i]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float c = 0.5f * (x.lo[k] + x.hi[k]);
            clo[k] = fminf(clo[k], c);
            chi[k] = fmaxf(chi[k], c);
            if (own) {
                q.lo[k] = fminf(q.lo[k], x.lo[k]);
                q.hi[k] = fmaxf(q.hi[k], x.hi[k]);
            }
        }
    }
    if (q.ref.root) {
#pragma unroll
        for (int k = 0; k < 3; ++k) { q.lo[k] = root_lo[k]; q.hi[k] = root_hi[k]; }
    } else if (!own) {
        for (uint32_t i = 0; i < q.ref.n; ++i) {
            const Box6 x = box_at(s_box, s_perm[q.ref.b + i]);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                q.lo[k] = fminf(q.lo[k], x.lo[k]);
                q.hi[k] = fmaxf(q.hi[k], x.hi[k]);
            }
        }
    }
    median = false;
    const float ext[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
    int axis = (ext[1] > ext[0]) ? 1 : 0;
    if (ext[2] > pick3(ext, axis)) axis = 2;
    const float cmin = pick3(clo, axis);
    const float eaxis = pick3(ext, axis);
    const float scale = (float)kSmallBins / eaxis;
    int split = -1;
    if (!(eaxis > 0.f) || !isfinite(scale) || q.depth >= kMedianDepth) {
        if (n <= (uint32_t)max_leaf) return false;
    } else {
        uint32_t bcnt[kSmallBins];
        float blo[kSmallBins][3], bhi[kSmallBins][3];
#pragma unroll
        for (int b = 0; b < kSmallBins; ++b) {
            bcnt[b] = 0;
            blo[b][0] = blo[b][1] = blo[b][2] = INFINITY;
            bhi[b][0] = bhi[b][1] = bhi[b][2] = -INFINITY;
        }
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t j = s_perm[q.b + i];
            const Box6 x = box_at(s_box, j);
            const float c = 0.5f * (pick3(x.lo, axis) + pick3(x.hi, axis));
            const int bi = bin_of(c, cmin, scale, kSmallBins);
            s_binof[j] = (uint8_t)bi;
            const float* pl = x.lo;
            const float* ph = x.hi;
#pragma unroll
            for (int b = 0; b < kSmallBins; ++b) {
                const bool m = bi == b;
                bcnt[b] += m ? 1u : 0u;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    blo[b][k] = m ? fminf(blo[b][k], pl[k]) : blo[b][k];
                    bhi[b][k] = m ? fmaxf(bhi[b][k], ph[k]) : bhi[b][k];
                }
            }
        }
        float best;
        sweep8(bcnt, blo, bhi, split, best);
        const float A = half_area(q.lo, q.hi);
        const float leaf_cost = (float)n * A;
        if (split < 0) {
            if (n <= (uint32_t)max_leaf) return false;
        } else if (n <= (uint32_t)max_leaf && leaf_cost <= trav_cost * A + best) {
            return false;
        }
        if (split >= 0) {
            uint32_t i = q.b, j = q.b + n - 1;  // n >= 2: j >= i
            while (i <= j) {
                const uint8_t pi = s_perm[i];
                if ((int)s_binof[pi] <= split) {
                    i++;
                } else {
                    s_perm[i] = s_perm[j];
                    s_perm[j] = pi;
                    if (j == q.b) break;  // everything went right (cannot happen with a valid split)
                    j--;
                }
            }
            split_boxes8(bcnt, blo, bhi, split, nl, Llo, Lhi, Rlo, Rhi);
            return true;
        }
    }
    // object median; the children keep the parent box (conservative)
    nl = n / 2;
    median = true;
    for (int k = 0; k < 3; ++k) { Llo[k] = Rlo[k] = q.lo[k]; Lhi[k] = Rhi[k] = q.hi[k]; }
    return true;
}

// Eight lanes of a group (r = lane & 7 holds bin r) decide and split queued
// node q together: the same decision and children as lane_split, with the
// node's bins spread over the group instead of held by one lane, so every lane
// reads each item once instead of updating eight bins per item, and a stable
// partition ranked by the group's ballots.  act: the group has a node.  Every
// lane of a group returns the same values; the shuffles run on all 64 lanes.
// bad is set when the partition's count disagrees with the bins'.
__device__ __forceinline__ bool group_split(QEnt& q, bool act, uint32_t r, uint8_t* s_perm, uint8_t* s_tmp,
                                            const float2* s_box, uint8_t* s_binof, const float root_lo[3],
                                            const float root_hi[3], int max_leaf, float trav_cost, uint32_t& nl,
                                            float Llo[3], float Lhi[3], float Rlo[3], float Rhi[3], bool& median,
                                            bool& bad) {
    const uint32_t n = act ? q.n : 0u;
    // centroid bounds and the node's box: lane r takes items r, r + 8, ... (of the box range when it
    // is not the node's own item range)
    const bool own = !q.ref.root && q.ref.b == q.b && q.ref.n == q.n;
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = r; i < n; i += 8) {
        const Box6 x = box_at(s_box, s_perm[q.b + i]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float c = 0.5f * (x.lo[k] + x.hi[k]);
            clo[k] = fminf(clo[k], c);
            chi[k] = fmaxf(chi[k], c);
            if (own) {
                nlo[k] = fminf(nlo[k], x.lo[k]);
                nhi[k] = fmaxf(nhi[k], x.hi[k]);
            }
        }
    }
    if (act && !own && !q.ref.root) {
        for (uint32_t i = r; i < q.ref.n; i += 8) {
            const Box6 x = box_at(s_box, s_perm[q.ref.b + i]);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                nlo[k] = fminf(nlo[k], x.lo[k]);
                nhi[k] = fmaxf(nhi[k], x.hi[k]);
            }
        }
    }
#pragma unroll
    for (int off = 1; off < 8; off <<= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            clo[k] = fminf(clo[k], __shfl_xor(clo[k], off, 64));
            chi[k] = fmaxf(chi[k], __shfl_xor(chi[k], off, 64));
            nlo[k] = fminf(nlo[k], __shfl_xor(nlo[k], off, 64));
            nhi[k] = fmaxf(nhi[k], __shfl_xor(nhi[k], off, 64));
        }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        q.lo[k] = q.ref.root ? root_lo[k] : nlo[k];
        q.hi[k] = q.ref.root ? root_hi[k] : nhi[k];
    }
    median = false;
    const float ext[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
    int axis = (ext[1] > ext[0]) ? 1 : 0;
    if (ext[2] > pick3(ext, axis)) axis = 2;
    const float cmin = pick3(clo, axis);
    const float eaxis = pick3(ext, axis);
    const float scale = (float)kSmallBins / eaxis;
    const bool binned = act && eaxis > 0.f && isfinite(scale) && q.depth < kMedianDepth;
    // bin r: every lane of the group reads every item; item i's bin is stored by lane i & 7
    uint32_t cnt = 0;
    float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (binned) {
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t j = s_perm[q.b + i];
            const Box6 x = box_at(s_box, j);
            const float c = 0.5f * (pick3(x.lo, axis) + pick3(x.hi, axis));
            const int bi = bin_of(c, cmin, scale, kSmallBins);
            if ((i & 7u) == r) s_binof[j] = (uint8_t)bi;
            if (bi == (int)r) {
                cnt++;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    blo[k] = fminf(blo[k], x.lo[k]);
                    bhi[k] = fmaxf(bhi[k], x.hi[k]);
                }
            }
        }
    }
    wave_sync();  // the group's s_binof stores land before lane 0 of the group partitions
    // left of split r: bins 0..r (inclusive prefix over the group); right: bins r+1..7
    uint32_t lc = cnt, sc = cnt;
    float ll[3], lh[3], sl[3], sh[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { ll[k] = sl[k] = blo[k]; lh[k] = sh[k] = bhi[k]; }
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
        const uint32_t oc = __shfl_up(lc, off, 8), os = __shfl_down(sc, off, 8);
        const bool up = r >= (uint32_t)off, down = r + off < 8u;
        lc += up ? oc : 0u;
        sc += down ? os : 0u;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float a = __shfl_up(ll[k], off, 8), b = __shfl_up(lh[k], off, 8);
            const float c = __shfl_down(sl[k], off, 8), d = __shfl_down(sh[k], off, 8);
            ll[k] = up ? fminf(ll[k], a) : ll[k];
            lh[k] = up ? fmaxf(lh[k], b) : lh[k];
            sl[k] = down ? fminf(sl[k], c) : sl[k];
            sh[k] = down ? fmaxf(sh[k], d) : sh[k];
        }
    }
    uint32_t rc = __shfl_down(sc, 1, 8);
    float rl[3], rh[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        rl[k] = __shfl_down(sl[k], 1, 8);
        rh[k] = __shfl_down(sh[k], 1, 8);
    }
    if (r == 7u) rc = 0u;  // no split after the last bin
    // the cheapest split: lowest cost, then lowest bin (the sweep's first strict minimum)
    float cost = INFINITY;
    if (lc != 0u && rc != 0u) cost = (float)lc * half_area(ll, lh) + (float)rc * half_area(rl, rh);
    const bool cand = cost < INFINITY;
    float bc = cand ? cost : INFINITY;
    int bs = cand ? (int)r : kSmallBins;
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
        const float oc = __shfl_xor(bc, off, 64);
        const int os = __shfl_xor(bs, off, 64);
        if (oc < bc || (oc == bc && os < bs)) {
            bc = oc;
            bs = os;
        }
    }
    const int split = bs < kSmallBins ? bs : -1;
    const float best = bc;
    // the children of split `split` as lane `split` of the group holds them
    const uint32_t src = (threadIdx.x & ~7u) + (uint32_t)(split < 0 ? 0 : split);
    const uint32_t snl = __shfl(lc, (int)src, 64);
    float sLlo[3], sLhi[3], sRlo[3], sRhi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        sLlo[k] = __shfl(ll[k], (int)src, 64);
        sLhi[k] = __shfl(lh[k], (int)src, 64);
        sRlo[k] = __shfl(rl[k], (int)src, 64);
        sRhi[k] = __shfl(rh[k], (int)src, 64);
    }
    if (!act) return false;
    if (!binned) {
        if (n <= (uint32_t)max_leaf) return false;
    } else {
        const float A = half_area(q.lo, q.hi);
        const float leaf_cost = (float)n * A;
        if (split < 0) {
            if (n <= (uint32_t)max_leaf) return false;
        } else if (n <= (uint32_t)max_leaf && leaf_cost <= trav_cost * A + best) {
            return false;
        }
        if (split >= 0) {
            // stable partition through s_tmp: lane r takes items k0 + r, ranked by the group's ballots
            const uint32_t gshift = threadIdx.x & ~7u, below = (1u << r) - 1u;
            uint32_t lb = 0, rb = 0;
            for (uint32_t k0 = 0; k0 < n; k0 += 8) {
                const bool valid = k0 + r < n;
                uint32_t j = 0;
                bool left = false;
                if (valid) {
                    j = s_perm[q.b + k0 + r];
                    left = (int)s_binof[j] <= split;
                }
                const uint32_t lm = (uint32_t)(__ballot(valid && left) >> gshift) & 0xffu;
                const uint32_t rm = (uint32_t)(__ballot(valid && !left) >> gshift) & 0xffu;
                if (valid)
                    s_tmp[q.b + (left ? lb + __popc(lm & below) : snl + rb + __popc(rm & below))] = (uint8_t)j;
                lb += (uint32_t)__popc(lm);
                rb += (uint32_t)__popc(rm);
            }
            if (lb != snl) bad = true;
            wave_sync();
            for (uint32_t i = r; i < n; i += 8) s_perm[q.b + i] = s_tmp[q.b + i];
            nl = snl;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Llo[k] = sLlo[k];
                Lhi[k] = sLhi[k];
                Rlo[k] = sRlo[k];
                Rhi[k] = sRhi[k];
            }
            return true;
        }
    }
    // object median; the children keep the parent box (conservative)
    nl = n / 2;
    median = true;
    for (int k = 0; k < 3; ++k) { Llo[k] = Rlo[k] = q.lo[k]; Lhi[k] = Rhi[k] = q.hi[k]; }
    return true;
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_build_small(const SmallTask* tasks, const uint32_t* range,
                                                    uint32_t* prims, const uint32_t* alt, const PrimBox* pb, BNodeG* nodes, BNodeG* snodes,
                                                    uint32_t* used, uint32_t* max_depth, uint32_t* err, int max_leaf,
                                                    float trav_cost) {
    __shared__ float2 s_box[3 * kSmall];
    __shared__ uint8_t s_perm[kSmall], s_tmp[kSmall];  // the items in their current order
    __shared__ QEntL s_q[kQueueCap];     // queued nodes of 2..kGroupMin-1 triangles
    __shared__ QEntL s_bq[kBigCap];      // queued nodes of kGroupMin..kCoop triangles
    __shared__ CEnt s_fq[kFrontCap];     // nodes of more than kCoop triangles, breadth first
    __shared__ int s_bk[kFront * kSmallBins * kBinVals];  // wave-split bins (keys) of a round's nodes
    __shared__ int s_split[kFront];                       // their splits (-1: object median)
    __shared__ uint8_t s_binof[kSmall];                 // each item's bin in its current node
    const uint32_t lane = threadIdx.x;
    // the tasks range[0] .. range[0] + range[1] - 1 (the grid may be larger)
    if (blockIdx.x >= range[1]) return;
    const uint32_t ti = range[0] + blockIdx.x;
    const SmallTask task = tasks[ti];
    BNodeG* const local = snodes + task.base;
    const uint32_t n0 = task.count;
    const uint32_t* const src = task.buf ? alt : prims;  // the task's ids; its final order goes to prims
    // all ids, then all boxes: the gathers of a lane are in flight together; the ids stay in
    // registers for the final order
    constexpr int kPer = kSmall / 64;
    uint32_t ids[kPer];
    {
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const uint32_t i = lane + 64u * r;
            ids[r] = i < n0 ? src[task.begin + i] : 0u;
        }
        PrimBox p[kPer];
#pragma unroll
        for (int r = 0; r < kPer; ++r)
            if (lane + 64u * r < n0) p[r] = pb[ids[r]];
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const uint32_t i = lane + 64u * r;
            if (i < n0) {
                s_perm[i] = (uint8_t)i;
                s_box[3 * i] = make_float2(p[r].lo[0], p[r].lo[1]);
                s_box[3 * i + 1] = make_float2(p[r].lo[2], p[r].hi[0]);
                s_box[3 * i + 2] = make_float2(p[r].hi[1], p[r].hi[2]);
            }
        }
    }
    const BNodeG root = nodes[task.node];
    wave_sync();
    uint32_t next = 0;  // scratch nodes used
    uint32_t qtail = 0, btail = 0, md = task.depth;
    bool fail = false;
    auto slot_ptr = [&](uint32_t s) -> BNodeG* { return s == kTaskRoot ? &nodes[task.node] : &local[s]; };
    // a node that splits: its record (box written when it was created) gets its first child
    auto set_internal = [&](uint32_t s, uint32_t c) {
        BNodeG* nd = slot_ptr(s);
        nd->left = c;
        nd->count = 0u;
    };
    // the box of a whole-wave node's box range, by the whole wave
    auto wave_ref_box = [&](const BoxRef& r, float lo[3], float hi[3]) {
        float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = lane; i < r.n; i += 64) {
            const Box6 x = box_at(s_box, s_perm[r.b + i]);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                l[k] = fminf(l[k], x.lo[k]);
                h[k] = fmaxf(h[k], x.hi[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            for (int off = 32; off > 0; off >>= 1) {
                l[k] = fminf(l[k], __shfl_xor(l[k], off));
                h[k] = fmaxf(h[k], __shfl_xor(h[k], off));
            }
            lo[k] = r.root ? root.lo[k] : l[k];
            hi[k] = r.root ? root.hi[k] : h[k];
        }
    };
    // ------------------------------------------------ nodes of more than kCoop triangles: the whole wave,
    // up to kFront of them at a time, breadth first (they are disjoint, so at most kFrontCap - 1 exist)
    uint32_t fhead = 0, ftail = 0;
    if (n0 > kCoop) {
        CEnt e;
        e.slot = kTaskRoot;
        e.b = 0;
        e.n = n0;
        e.depth = task.depth;
        e.ref = ref_pack(BoxRef{0u, n0, true});
        wave_centroid_bounds(s_perm, s_box, 0, n0, lane, e.clo, e.chi);
        if (lane == 0) s_fq[0] = e;
        ftail = 1;
    } else if (n0 >= 2) {
        const QEntL r0 = q_pack(kTaskRoot, 0u, n0, task.depth, BoxRef{0u, n0, true});
        if (n0 >= kGroupMin) {
            if (lane == 0) s_bq[0] = r0;
            btail = 1;
        } else {
            if (lane == 0) s_q[0] = r0;
            qtail = 1;
        }
    }
    while (fhead < ftail && !fail) {
        const uint32_t nf = min((uint32_t)kFront, ftail - fhead);
        wave_sync();
        // bins as order-preserving integer keys, filled by LDS atomics (one visit per item)
        for (uint32_t t = lane; t < nf * (uint32_t)(kSmallBins * kBinVals); t += 64) {
            const int v = (int)(t % kBinVals);
            s_bk[t] = v == 0 ? 0 : (((v - 1) % 6) < 3 ? INT_MAX : INT_MIN);
        }
        wave_sync();
        for (uint32_t f = 0; f < nf; ++f) {
            const CEnt& ef = s_fq[(fhead + f) % kFrontCap];
            const uint32_t eb = ef.b, en = ef.n;
            const WaveBinning wb = wave_binning(ef.clo, ef.chi, ef.depth);
            if (!wb.binned) continue;  // the same for the whole wave
            int* bk = &s_bk[f * kSmallBins * kBinVals];
            for (uint32_t i = lane; i < en; i += 64) {
                const uint32_t j = s_perm[eb + i];
                const Box6 x = box_at(s_box, j);
                const float c[3] = {0.5f * (x.lo[0] + x.hi[0]), 0.5f * (x.lo[1] + x.hi[1]), 0.5f * (x.lo[2] + x.hi[2])};
                const int bi = bin_of(pick3(c, wb.axis), wb.cmin, wb.scale, kSmallBins);
                s_binof[j] = (uint8_t)bi;
                int* sb = &bk[bi * kBinVals];
                atomicAdd(&sb[0], 1);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    atomicMin(&sb[1 + k], f2key(x.lo[k]));
                    atomicMax(&sb[4 + k], f2key(x.hi[k]));
                    atomicMin(&sb[7 + k], f2key(c[k]));
                    atomicMax(&sb[10 + k], f2key(c[k]));
                }
            }
        }
        wave_sync();
        // lane f < nf: node f's SAH decision and its children (vira_blas.ipp:185-254 with bins)
        CEnt e{}, L{}, R{};
        float Llo[3] = {}, Lhi[3] = {}, Rlo[3] = {}, Rhi[3] = {};  // the children's boxes
        uint32_t nl = 0;
        int split = -1;
        if (lane < nf) {
            e = s_fq[(fhead + lane) % kFrontCap];
            md = max(md, e.depth);
            const WaveBinning wb = wave_binning(e.clo, e.chi, e.depth);
            if (wb.binned) {
                const int* bk = &s_bk[lane * kSmallBins * kBinVals];
                uint32_t bcnt[kSmallBins];
                float blo[kSmallBins][3], bhi[kSmallBins][3];
#pragma unroll
                for (int b = 0; b < kSmallBins; ++b) {
                    const int* sb = &bk[b * kBinVals];
                    bcnt[b] = (uint32_t)sb[0];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        blo[b][k] = bcnt[b] ? key2f(sb[1 + k]) : INFINITY;
                        bhi[b][k] = bcnt[b] ? key2f(sb[4 + k]) : -INFINITY;
                    }
                }
                float best;
                sweep8(bcnt, blo, bhi, split, best);
                if (split >= 0) {
                    split_boxes8(bcnt, blo, bhi, split, nl, Llo, Lhi, Rlo, Rhi);
                    // the children's centroid bounds: unions of the bins' (in key order, then converted)
                    int kl[3] = {INT_MAX, INT_MAX, INT_MAX}, kh[3] = {INT_MIN, INT_MIN, INT_MIN};
                    int rl[3] = {INT_MAX, INT_MAX, INT_MAX}, rh[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
                    for (int b = 0; b < kSmallBins; ++b) {
                        const bool left = b <= split;
                        const int* sb = &bk[b * kBinVals];
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            const int cl = sb[7 + k], ch = sb[10 + k];
                            kl[k] = left ? min(kl[k], cl) : kl[k];
                            kh[k] = left ? max(kh[k], ch) : kh[k];
                            rl[k] = left ? rl[k] : min(rl[k], cl);
                            rh[k] = left ? rh[k] : max(rh[k], ch);
                        }
                    }
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        L.clo[k] = key2f(kl[k]);
                        L.chi[k] = key2f(kh[k]);
                        R.clo[k] = key2f(rl[k]);
                        R.chi[k] = key2f(rh[k]);
                    }
                }
            }
            // object median: the children keep the parent box (conservative, below) and its box
            // range, and get exact centroid bounds (below)
            if (split < 0) nl = e.n / 2;
            L.ref = split < 0 ? e.ref : ref_pack(BoxRef{e.b, nl, false});
            R.ref = split < 0 ? e.ref : ref_pack(BoxRef{e.b + nl, e.n - nl, false});
            s_split[lane] = split;
        }
        wave_sync();
        // each node's stable partition through s_tmp, or its median halves' centroid bounds
        for (uint32_t f = 0; f < nf; ++f) {
            const int sf = s_split[f];
            const uint32_t eb = s_fq[(fhead + f) % kFrontCap].b, en = s_fq[(fhead + f) % kFrontCap].n;
            const uint32_t nlf = __shfl(nl, (int)f, 64);
            if (sf >= 0) {
                uint32_t lb = 0, rb = 0;
                for (uint32_t k0 = 0; k0 < en; k0 += 64) {
                    const bool valid = k0 + lane < en;
                    uint32_t j = 0;
                    bool left = false;
                    if (valid) {
                        j = s_perm[eb + k0 + lane];
                        left = (int)s_binof[j] <= sf;
                    }
                    const uint64_t lm = __ballot(valid && left), rm = __ballot(valid && !left);
                    if (valid) s_tmp[eb + (left ? lb + rank_below(lm) : nlf + rb + rank_below(rm))] = (uint8_t)j;
                    lb += (uint32_t)__popcll(lm);
                    rb += (uint32_t)__popcll(rm);
                }
                if (lb != nlf) fail = true;
                wave_sync();
                for (uint32_t i = lane; i < en; i += 64) s_perm[eb + i] = s_tmp[eb + i];
                wave_sync();
            } else {
                float a[3], bh[3], c[3], d[3], plo[3], phi[3];
                wave_centroid_bounds(s_perm, s_box, eb, nlf, lane, a, bh);
                wave_centroid_bounds(s_perm, s_box, eb + nlf, en - nlf, lane, c, d);
                const CEnt& ef = s_fq[(fhead + f) % kFrontCap];
                wave_ref_box(ref_unpack(ef.ref), plo, phi);
                if (lane == f)
                    for (int k = 0; k < 3; ++k) {
                        L.clo[k] = a[k];
                        L.chi[k] = bh[k];
                        R.clo[k] = c[k];
                        R.chi[k] = d[k];
                        Llo[k] = Rlo[k] = plo[k];
                        Lhi[k] = Rhi[k] = phi[k];
                    }
            }
        }
        wave_sync();  // every read of this round's entries is done before the children are queued
        fhead += nf;
        const bool act = lane < nf;
        const uint32_t cnode = next + 2 * lane;
        next += 2 * nf;  // every node here splits
        if (act) {
            L.slot = cnode;
            L.b = e.b;
            L.n = nl;
            R.slot = cnode + 1;
            R.b = e.b + nl;
            R.n = e.n - nl;
            L.depth = R.depth = e.depth + 1;
            local[cnode] = BNodeG{{Llo[0], Llo[1], Llo[2]}, {Lhi[0], Lhi[1], Lhi[2]}, task.begin + L.b, L.n};
            local[cnode + 1] = BNodeG{{Rlo[0], Rlo[1], Rlo[2]}, {Rhi[0], Rhi[1], Rhi[2]}, task.begin + R.b, R.n};
            set_internal(e.slot, cnode);
        }
        // children: more than kCoop triangles to the next rounds, kGroupMin.. to the groups' queue,
        // 2.. to the lanes' queue; single triangles are leaves as written
        const int cl = act ? (L.n > kCoop ? 0 : (L.n >= kGroupMin ? 1 : (L.n >= 2 ? 2 : 3))) : 4;
        const int cr = act ? (R.n > kCoop ? 0 : (R.n >= kGroupMin ? 1 : (R.n >= 2 ? 2 : 3))) : 4;
        {
            const uint64_t ml = __ballot(cl == 0), mr = __ballot(cr == 0);
            const uint32_t pos = ftail + rank_below(ml) + rank_below(mr);
            if (cl == 0) s_fq[pos % kFrontCap] = L;
            if (cr == 0) s_fq[(pos + (cl == 0 ? 1u : 0u)) % kFrontCap] = R;
            ftail += (uint32_t)(__popcll(ml) + __popcll(mr));
        }
        {
            const uint64_t ml = __ballot(cl == 1), mr = __ballot(cr == 1);
            const uint32_t pos = btail + rank_below(ml) + rank_below(mr);
            if (cl == 1) s_bq[pos % kBigCap] = q_pack(L.slot, L.b, L.n, L.depth, ref_unpack(L.ref));
            if (cr == 1)
                s_bq[(pos + (cl == 1 ? 1u : 0u)) % kBigCap] = q_pack(R.slot, R.b, R.n, R.depth, ref_unpack(R.ref));
            btail += (uint32_t)(__popcll(ml) + __popcll(mr));
        }
        {
            const uint64_t ml = __ballot(cl == 2), mr = __ballot(cr == 2);
            const uint32_t pos = qtail + rank_below(ml) + rank_below(mr);
            if (cl == 2) s_q[pos % kQueueCap] = q_pack(L.slot, L.b, L.n, L.depth, ref_unpack(L.ref));
            if (cr == 2)
                s_q[(pos + (cl == 2 ? 1u : 0u)) % kQueueCap] = q_pack(R.slot, R.b, R.n, R.depth, ref_unpack(R.ref));
            qtail += (uint32_t)(__popcll(ml) + __popcll(mr));
        }
        if (cl == 3) md = max(md, L.depth);
        if (cr == 3) md = max(md, R.depth);
        if (ftail - fhead > kFrontCap || btail > kBigCap || qtail > kQueueCap) fail = true;
    }
    if (qtail > kQueueCap || btail > kBigCap) fail = true;
    // ------------------------------------------------ nodes of kGroupMin..kCoop triangles: 8 lanes each
    {
        const uint32_t g = lane >> 3, r = lane & 7u;
        uint32_t bhead = 0;
        while (bhead < btail && !fail) {
            const uint32_t k = min(8u, btail - bhead);
            const bool act = g < k;
            QEnt q{};
            if (act) q = q_unpack(s_bq[(bhead + g) % kBigCap]);
            bhead += k;
            wave_sync();
            uint32_t nl = 0;
            float Llo[3], Lhi[3], Rlo[3], Rhi[3];
            bool median = false;
            if (act) md = max(md, q.depth);
            const bool split = group_split(q, act, r, s_perm, s_tmp, s_box, s_binof, root.lo, root.hi, max_leaf,
                                           trav_cost, nl, Llo, Lhi, Rlo, Rhi, median, fail);
            const bool lead = act && r == 0u;
            const uint64_t sm = __ballot(lead && split);
            const uint32_t c = next + 2 * rank_below(sm);
            next += 2 * (uint32_t)__popcll(sm);
            const uint32_t nr = q.n - nl;
            if (lead && split) {
                local[c] = BNodeG{{Llo[0], Llo[1], Llo[2]}, {Lhi[0], Lhi[1], Lhi[2]}, task.begin + q.b, nl};
                local[c + 1] = BNodeG{{Rlo[0], Rlo[1], Rlo[2]}, {Rhi[0], Rhi[1], Rhi[2]}, task.begin + q.b + nl, nr};
                set_internal(q.slot, c);
                if (nl == 1 || nr == 1) md = max(md, q.depth + 1);
            }
            // children: kGroupMin or more triangles to this queue, 2.. to the lanes' queue
            const bool lb = lead && split && nl >= kGroupMin, rbg = lead && split && nr >= kGroupMin;
            const bool ls = lead && split && nl >= 2 && nl < kGroupMin, rs = lead && split && nr >= 2 && nr < kGroupMin;
            const uint64_t mlb = __ballot(lb), mrb = __ballot(rbg), mls = __ballot(ls), mrs = __ballot(rs);
            const uint32_t pb2 = btail + rank_below(mlb) + rank_below(mrb);
            const uint32_t ps = qtail + rank_below(mls) + rank_below(mrs);
            const QEntL eL = q_pack(c, q.b, nl, q.depth + 1, pick_ref(median, q.ref, BoxRef{q.b, nl, false}));
            const QEntL eR =
                q_pack(c + 1, q.b + nl, nr, q.depth + 1, pick_ref(median, q.ref, BoxRef{q.b + nl, nr, false}));
            if (lb) s_bq[pb2 % kBigCap] = eL;
            if (rbg) s_bq[(pb2 + (lb ? 1u : 0u)) % kBigCap] = eR;
            if (ls) s_q[ps % kQueueCap] = eL;
            if (rs) s_q[(ps + (ls ? 1u : 0u)) % kQueueCap] = eR;
            btail += (uint32_t)(__popcll(mlb) + __popcll(mrb));
            qtail += (uint32_t)(__popcll(mls) + __popcll(mrs));
            if (btail - bhead > kBigCap || qtail > kQueueCap) fail = true;
            wave_sync();
        }
    }
    // ------------------------------------------------ nodes of 2..kGroupMin-1 triangles: one lane each
    uint32_t qhead = 0;
    while (qhead < qtail && !fail) {
        const uint32_t k = min(64u, qtail - qhead);
        const bool act = lane < k;
        QEnt q{};
        if (act) q = q_unpack(s_q[(qhead + lane) % kQueueCap]);
        qhead += k;
        wave_sync();
        uint32_t nl = 0;
        float Llo[3], Lhi[3], Rlo[3], Rhi[3];
        bool split = false, median = false;
        if (act) {
            md = max(md, q.depth);
            split = lane_split(q, s_perm, s_box, s_binof, root.lo, root.hi, max_leaf, trav_cost, nl, Llo, Lhi, Rlo,
                               Rhi, median);
        }
        const uint64_t sm = __ballot(split);
        const uint32_t c = next + 2 * rank_below(sm);
        next += 2 * (uint32_t)__popcll(sm);
        const uint32_t nr = q.n - nl;
        if (split) {
            local[c] = BNodeG{{Llo[0], Llo[1], Llo[2]}, {Lhi[0], Lhi[1], Lhi[2]}, task.begin + q.b, nl};
            local[c + 1] = BNodeG{{Rlo[0], Rlo[1], Rlo[2]}, {Rhi[0], Rhi[1], Rhi[2]}, task.begin + q.b + nl, nr};
            set_internal(q.slot, c);
            if (nl == 1 || nr == 1) md = max(md, q.depth + 1);
        }
        const bool pl = split && nl >= 2, pr = split && nr >= 2;
        const uint64_t ml = __ballot(pl), mr = __ballot(pr);
        const uint32_t pos = qtail + rank_below(ml) + rank_below(mr);
        if (pl) s_q[pos % kQueueCap] = q_pack(c, q.b, nl, q.depth + 1, pick_ref(median, q.ref, BoxRef{q.b, nl, false}));
        if (pr)
            s_q[(pos + (pl ? 1u : 0u)) % kQueueCap] =
                q_pack(c + 1, q.b + nl, nr, q.depth + 1, pick_ref(median, q.ref, BoxRef{q.b + nl, nr, false}));
        qtail += (uint32_t)(__popcll(ml) + __popcll(mr));
        if (qtail - qhead > kQueueCap) fail = true;
        wave_sync();
    }
    wave_sync();
    // the task's items in their final order
    {
        // the original ids through LDS (the boxes' space, no longer used)
        uint32_t* s_ids = reinterpret_cast<uint32_t*>(s_box);
#pragma unroll
        for (int r = 0; r < kPer; ++r)
            if (lane + 64u * r < n0) s_ids[lane + 64u * r] = ids[r];
        wave_sync();
        for (uint32_t i = lane; i < n0; i += 64) prims[task.begin + i] = s_ids[s_perm[i]];
    }
    for (int off = 32; off > 0; off >>= 1) md = max(md, (uint32_t)__shfl_xor((int)md, off));
    const bool any_fail = __ballot(fail) != 0;
    if (lane == 0) {
        used[ti] = next;
        atomicMax(max_depth, md);
        if (any_fail) atomicOr(err, 2u);
    }
}

// Moves each task's scratch nodes to their final place (first = the task's
// exclusive-scan offset behind the large-level nodes) and turns task-local
// child numbers into global ones, the task root's included.
__global__ void k_compact_small(const SmallTask* tasks, uint32_t ntasks, const uint32_t* used,
                                const uint32_t* first, uint32_t offset, const BNodeG* snodes, BNodeG* nodes) {
    const uint32_t ti = blockIdx.x;
    if (ti >= ntasks) return;
    const SmallTask task = tasks[ti];
    const uint32_t n = used[ti], dst = offset + first[ti];
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        BNodeG nd = snodes[task.base + i];
        if (nd.count == 0) nd.left += dst;
        nodes[dst + i] = nd;
    }
    if (threadIdx.x == 0 && nodes[task.node].count == 0) nodes[task.node].left += dst;
}

// Binary tree -> 64-byte child-box nodes.  internal_index[n] is the exclusive
// scan of "node n is internal".
__global__ void k_flag_internal(const BNodeG* nodes, uint32_t nn, uint32_t* flag) {
    uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    flag[n] = nodes[n].count == 0 ? 1u : 0u;
}

__device__ __forceinline__ uint32_t child_ref(const BNodeG* nodes, const uint32_t* iidx, uint32_t c,
                                              uint32_t* err) {
    const BNodeG& ch = nodes[c];
    if (ch.count == 0) return iidx[c];
    if (ch.count > 16) {  // not encodable: report instead of emitting a bad reference
        atomicOr(err, 1u);
        return kEmptyRef;
    }
    return make_leaf(ch.left, ch.count);
}

__global__ void k_emit_nodes(const BNodeG* nodes, uint32_t nn, const uint32_t* iidx, Node2* out, uint32_t* err) {
    uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const BNodeG& nd = nodes[n];
    if (nd.count != 0) return;
    const BNodeG& a = nodes[nd.left];
    const BNodeG& b = nodes[nd.left + 1];
    Node2 o;
    o.x = make_float4(a.lo[0], a.hi[0], b.lo[0], b.hi[0]);
    o.y = make_float4(a.lo[1], a.hi[1], b.lo[1], b.hi[1]);
    o.z = make_float4(a.lo[2], a.hi[2], b.lo[2], b.hi[2]);
    o.ref = make_uint4(child_ref(nodes, iidx, nd.left, err), child_ref(nodes, iidx, nd.left + 1, err), 0u, 0u);
    out[iidx[n]] = o;
}

// Root that is a leaf: one node with the leaf as child 0 and an empty child 1.
__global__ void k_emit_leaf_root(const BNodeG* nodes, Node2* out) {
    const BNodeG& a = nodes[0];
    Node2 o;
    o.x = make_float4(a.lo[0], a.hi[0], 0.f, 0.f);
    o.y = make_float4(a.lo[1], a.hi[1], 0.f, 0.f);
    o.z = make_float4(a.lo[2], a.hi[2], 0.f, 0.f);
    o.ref = make_uint4(make_leaf(a.left, a.count), kEmptyRef, 0u, 0u);
    out[0] = o;
}

// Binary tree -> four-wide nodes, one level of the four-wide tree per launch.
// Each item turns binary internal node `bnode` into Node4 `n4`: starting from
// its two children, the internal child of largest surface area is replaced by
// its own two children until there are four (or only leaves remain).
struct CItem {
    uint32_t bnode, n4;
};

// The up-to-4 descendants a Node4 takes from binary node bnode: its two
// children, then repeatedly the internal one of largest surface area opened
// into its two children.
__device__ int collapse_children(const BNodeG* __restrict__ nodes, uint32_t bnode, uint32_t ch[4]) {
    const BNodeG& bn = nodes[bnode];
    ch[0] = bn.left;
    ch[1] = bn.left + 1;
    int nc = 2;
    while (nc < 4) {
        int best = -1;
        float bestA = -1.0f;
        for (int k = 0; k < nc; ++k) {
            const BNodeG& c = nodes[ch[k]];
            if (c.count != 0) continue;
            float A = half_area(c.lo, c.hi);
            if (A > bestA) {
                bestA = A;
                best = k;
            }
        }
        if (best < 0) break;
        uint32_t l = nodes[ch[best]].left;
        ch[best] = l;
        ch[nc++] = l + 1;
    }
    return nc;
}

// Binary tree -> Node4, one level of Node4s per pair of launches: the count
// pass gives each item its number of internal children, an exclusive scan
// numbers them, and the emit pass writes the item's Node4 and queues those
// children (Node4 index = first index of the next level + scan offset), so the
// layout is the same on every run and each level's Node4s are consecutive.
__global__ void k_collapse4_count(const BNodeG* __restrict__ nodes, const CItem* __restrict__ in, uint32_t nin,
                                  uint32_t* __restrict__ cnt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nin) return;
    uint32_t ch[4];
    const int nc = collapse_children(nodes, in[i].bnode, ch);
    uint32_t n = 0;
    for (int k = 0; k < nc; ++k) n += nodes[ch[k]].count == 0 ? 1u : 0u;
    cnt[i] = n;
}

// The Node4 of item it (children ch[0..nc)) and its internal children queued at
// qout[q..], numbered next_base + their queue position.
__device__ __forceinline__ void collapse4_write(const BNodeG* __restrict__ nodes, const CItem it, const uint32_t ch[4],
                                                int nc, uint32_t q, uint32_t next_base, Node4* __restrict__ out,
                                                CItem* __restrict__ qout, uint32_t* __restrict__ err) {
    float lo[3][4], hi[3][4];
    uint32_t ref[4];
    for (int k = 0; k < 4; ++k) {
        if (k >= nc) {
            for (int a = 0; a < 3; ++a) lo[a][k] = hi[a][k] = 0.0f;
            ref[k] = kEmptyRef;
            continue;
        }
        const BNodeG& c = nodes[ch[k]];
        for (int a = 0; a < 3; ++a) {
            lo[a][k] = c.lo[a];
            hi[a][k] = c.hi[a];
        }
        if (c.count != 0) {
            if (c.count > 16) {
                atomicOr(err, 1u);
                ref[k] = kEmptyRef;
            } else {
                ref[k] = make_leaf(c.left, c.count);
            }
        } else {
            const uint32_t n = next_base + q;
            qout[q++] = CItem{ch[k], n};
            ref[k] = n;
        }
    }
    Node4 o;
    o.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    o.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    o.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    o.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    o.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    o.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    o.ref = make_uint4(ref[0], ref[1], ref[2], ref[3]);
    o.pad = make_uint4(0u, 0u, 0u, 0u);
    out[it.n4] = o;
}

__global__ void k_collapse4_emit(const BNodeG* __restrict__ nodes, const CItem* __restrict__ in, uint32_t nin,
                                 const uint32_t* __restrict__ off, uint32_t next_base, Node4* __restrict__ out,
                                 CItem* __restrict__ qout, uint32_t* __restrict__ err) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nin) return;
    const CItem it = in[i];
    uint32_t ch[4];
    const int nc = collapse_children(nodes, it.bnode, ch);
    collapse4_write(nodes, it, ch, nc, off[i], next_base, out, qout, err);
}

// The first levels of the collapse in one workgroup (one item per thread, while
// a level has at most kTopItems items), with the same numbering as the
// count / scan / emit passes.  info: [0] items of the next level (left to those
// passes; 0 = done), [1] Node4s numbered (cap + 1: more than cap were needed),
// [2] levels done, [3] queue of the next level (0: qa, 1: qb), [4 + l] Node4s
// numbered after level l.
constexpr int kTopItems = 1024;
constexpr int kTopMaxLevels = 12;
__global__ void __launch_bounds__(kTopItems) k_collapse4_top(const BNodeG* __restrict__ nodes, CItem* qa, CItem* qb,
                                                             uint32_t cap, Node4* __restrict__ out, uint32_t* info,
                                                             uint32_t* __restrict__ err) {
    __shared__ uint32_t s_wsum[kTopItems / 64];
    const uint32_t i = threadIdx.x, lane = i & 63, w = i >> 6;
    uint32_t nin = 1, total = 1, lv = 0;
    CItem* qin = qa;
    CItem* qout = qb;
    while (nin > 0 && nin <= (uint32_t)kTopItems && lv < (uint32_t)kTopMaxLevels) {
        CItem it{0u, 0u};
        uint32_t ch[4];
        int nc = 0;
        uint32_t n = 0;
        if (i < nin) {
            it = qin[i];
            nc = collapse_children(nodes, it.bnode, ch);
            for (int k = 0; k < nc; ++k) n += nodes[ch[k]].count == 0 ? 1u : 0u;
        }
        // exclusive scan of n over the workgroup: within the wave, then over the waves' sums
        uint32_t x = n;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= (uint32_t)off) x += y;
        }
        if (lane == 63) s_wsum[w] = x;
        block_sync();
        uint32_t before = 0, nnext = 0;
        for (uint32_t q = 0; q < kTopItems / 64; ++q) {
            const uint32_t v = s_wsum[q];
            before += q < w ? v : 0u;
            nnext += v;
        }
        if (total + nnext > cap) {  // the same for every thread; reported as total = cap + 1
            total = cap + 1;
            break;
        }
        if (i < nin) collapse4_write(nodes, it, ch, nc, before + x - n, total, out, qout, err);
        total += nnext;
        if (i == 0) info[4 + lv] = total;
        lv++;
        nin = nnext;
        CItem* t = qin;
        qin = qout;
        qout = t;
        block_sync_all();  // the next level's queue is complete (global memory) and s_wsum is read
    }
    if (i == 0) {
        info[0] = nin;
        info[1] = total;
        info[2] = lv;
        info[3] = qin == qa ? 0u : 1u;
    }
}

// Refit (the tree kept, boxes recomputed from the current vertices): the
// leaf-order triangles are rewritten from their ids, then each level of
// Node4s, deepest first, takes every child's box from its triangles (leaf)
// or from the union of the child Node4's boxes.
__global__ void k_refit_tris(size_t nt, const uint32_t* __restrict__ idx, const GVertex* __restrict__ V,
                             float4* __restrict__ tris) {
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nt) return;
    const float4 t0 = tris[3 * s];
    const size_t p = __float_as_uint(t0.w);
    const GVertex& a = V[idx[3 * p]];
    const GVertex& b = V[idx[3 * p + 1]];
    const GVertex& c = V[idx[3 * p + 2]];
    tris[3 * s] = make_float4(a.px, a.py, a.pz, t0.w);
    tris[3 * s + 1] = make_float4(b.px, b.py, b.pz, 0.f);
    tris[3 * s + 2] = make_float4(c.px, c.py, c.pz, 0.f);
}

__device__ __forceinline__ float f4_get(const float4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }
__device__ __forceinline__ uint32_t u4_get(const uint4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }

// One level of the four-wide refit, one lane per child (four consecutive lanes
// per node): a leaf child's box from its triangles, loaded up to four at a time
// before the first min/max; an internal child's box as the union of its own
// children's boxes.  Each lane writes only its child's six bounds, so the
// references are not rewritten.  An empty child keeps a zero box.
constexpr int kRefitThreads = 256;
__global__ __launch_bounds__(kRefitThreads) void k_refit4(Node4* __restrict__ nodes, uint32_t first, uint32_t count,
                                                         const float4* __restrict__ tris) {
    const uint32_t g = blockIdx.x * kRefitThreads + threadIdx.x;
    const uint32_t i = g >> 2, k = g & 3u;
    if (i >= count) return;
    Node4* nd = nodes + first + i;
    const uint32_t r = reinterpret_cast<const uint32_t*>(&nd->ref)[k];
    float l0 = INFINITY, l1 = INFINITY, l2 = INFINITY, h0 = -INFINITY, h1 = -INFINITY, h2 = -INFINITY;
    if (r == kEmptyRef) {
        l0 = l1 = l2 = h0 = h1 = h2 = 0.0f;
    } else if (ref_is_leaf(r)) {
        const uint32_t f = leaf_first(r), n = leaf_count(r);
        for (uint32_t t0 = 0; t0 < n; t0 += 4) {
            float4 p[12];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const size_t t = f + min(t0 + (uint32_t)q, n - 1u);  // past the leaf: its last triangle again
                p[3 * q] = tris[3 * t];
                p[3 * q + 1] = tris[3 * t + 1];
                p[3 * q + 2] = tris[3 * t + 2];
            }
#pragma unroll
            for (int v = 0; v < 12; ++v) {
                l0 = fminf(l0, p[v].x); l1 = fminf(l1, p[v].y); l2 = fminf(l2, p[v].z);
                h0 = fmaxf(h0, p[v].x); h1 = fmaxf(h1, p[v].y); h2 = fmaxf(h2, p[v].z);
            }
        }
    } else {
        const Node4 cn = nodes[r];
        for (int q = 0; q < 4; ++q) {
            if (u4_get(cn.ref, q) == kEmptyRef) continue;
            l0 = fminf(l0, f4_get(cn.lox, q)); h0 = fmaxf(h0, f4_get(cn.hix, q));
            l1 = fminf(l1, f4_get(cn.loy, q)); h1 = fmaxf(h1, f4_get(cn.hiy, q));
            l2 = fminf(l2, f4_get(cn.loz, q)); h2 = fmaxf(h2, f4_get(cn.hiz, q));
        }
    }
    reinterpret_cast<float*>(&nd->lox)[k] = l0;
    reinterpret_cast<float*>(&nd->loy)[k] = l1;
    reinterpret_cast<float*>(&nd->loz)[k] = l2;
    reinterpret_cast<float*>(&nd->hix)[k] = h0;
    reinterpret_cast<float*>(&nd->hiy)[k] = h1;
    reinterpret_cast<float*>(&nd->hiz)[k] = h2;
}

// The root's box (the union of its children) for the BLAS box kept on the host.
__global__ void k_root_box4(const Node4* __restrict__ nodes, float* __restrict__ box) {
    const Node4 nd = nodes[0];
    float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int q = 0; q < 4; ++q) {
        if (u4_get(nd.ref, q) == kEmptyRef) continue;
        l[0] = fminf(l[0], f4_get(nd.lox, q)); h[0] = fmaxf(h[0], f4_get(nd.hix, q));
        l[1] = fminf(l[1], f4_get(nd.loy, q)); h[1] = fmaxf(h[1], f4_get(nd.hiy, q));
        l[2] = fminf(l[2], f4_get(nd.loz, q)); h[2] = fmaxf(h[2], f4_get(nd.hiz, q));
    }
    for (int a = 0; a < 3; ++a) { box[a] = l[a]; box[3 + a] = h[a]; }
}

__global__ void k_emit_leaf_root4(const BNodeG* nodes, Node4* out) {
    const BNodeG& a = nodes[0];
    Node4 o;
    o.lox = make_float4(a.lo[0], 0.f, 0.f, 0.f);
    o.hix = make_float4(a.hi[0], 0.f, 0.f, 0.f);
    o.loy = make_float4(a.lo[1], 0.f, 0.f, 0.f);
    o.hiy = make_float4(a.hi[1], 0.f, 0.f, 0.f);
    o.loz = make_float4(a.lo[2], 0.f, 0.f, 0.f);
    o.hiz = make_float4(a.hi[2], 0.f, 0.f, 0.f);
    o.ref = make_uint4(make_leaf(a.left, a.count), kEmptyRef, kEmptyRef, kEmptyRef);
    o.pad = make_uint4(0u, 0u, 0u, 0u);
    out[0] = o;
}

// Triangles in leaf order, 48 B each (three float4 of vertex positions, the
// first carrying the triangle id).  A block stages its kEmitThreads triangles
// in LDS and writes them as one contiguous run, so each store instruction
// covers 1 KB of consecutive addresses instead of 16 B pieces 48 B apart.
constexpr int kEmitThreads = 256;
__global__ __launch_bounds__(kEmitThreads) void k_emit_tris(size_t nt, const uint32_t* prims, const uint32_t* idx,
                                                           const GVertex* V, float4* tris) {
    __shared__ float4 st[3 * kEmitThreads];
    const size_t base = (size_t)blockIdx.x * kEmitThreads;
    const size_t s = base + threadIdx.x;
    if (s < nt) {
        const uint32_t p = prims[s];
        const GVertex& a = V[idx[3 * (size_t)p]];
        const GVertex& b = V[idx[3 * (size_t)p + 1]];
        const GVertex& c = V[idx[3 * (size_t)p + 2]];
        st[3 * threadIdx.x] = make_float4(a.px, a.py, a.pz, __uint_as_float(p));
        st[3 * threadIdx.x + 1] = make_float4(b.px, b.py, b.pz, 0.f);
        st[3 * threadIdx.x + 2] = make_float4(c.px, c.py, c.pz, 0.f);
    }
    block_sync();
    const uint32_t n = 3u * (uint32_t)min((size_t)kEmitThreads, nt - base);
    float4* out = tris + 3 * base;
    for (uint32_t k = threadIdx.x; k < n; k += kEmitThreads) out[k] = st[k];
}

// Tree statistics (internal / leaf counts, largest leaf, SAH cost): a fixed
// grid strides over the nodes and reduces in LDS to one partial per block,
// then one block adds the partials in a fixed order (no contended atomics,
// the same sum on every run).
constexpr int kStatsBlocks = 1024;
struct StatPart {
    unsigned long long internal, leaves;
    uint32_t max_leaf, pad;
    double sah;
};

__device__ void stats_reduce(StatPart* sp, StatPart v) {
    const uint32_t t = threadIdx.x;
    sp[t] = v;
    __syncthreads();
    for (uint32_t off = blockDim.x / 2; off > 0; off >>= 1) {
        if (t < off) {
            sp[t].internal += sp[t + off].internal;
            sp[t].leaves += sp[t + off].leaves;
            sp[t].max_leaf = max(sp[t].max_leaf, sp[t + off].max_leaf);
            sp[t].sah += sp[t + off].sah;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_tree_stats(const BNodeG* nodes, uint32_t nn, float inv_root_area,
                                                    StatPart* parts) {
    __shared__ StatPart sp[256];
    StatPart v{0ull, 0ull, 0u, 0u, 0.0};
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += gridDim.x * blockDim.x) {
        const BNodeG& nd = nodes[n];
        const double a = (double)(half_area(nd.lo, nd.hi) * inv_root_area);
        if (nd.count == 0) {
            v.internal++;
            v.sah += a;
        } else {
            v.leaves++;
            v.max_leaf = max(v.max_leaf, nd.count);
            v.sah += a * (double)nd.count;
        }
    }
    stats_reduce(sp, v);
    if (threadIdx.x == 0) parts[blockIdx.x] = sp[0];
}

__global__ void __launch_bounds__(256) k_tree_stats_sum(const StatPart* parts, uint32_t nparts, StatPart* out) {
    __shared__ StatPart sp[256];
    StatPart v{0ull, 0ull, 0u, 0u, 0.0};
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) {
        const StatPart p = parts[i];
        v.internal += p.internal;
        v.leaves += p.leaves;
        v.max_leaf = max(v.max_leaf, p.max_leaf);
        v.sah += p.sah;
    }
    stats_reduce(sp, v);
    if (threadIdx.x == 0) *out = sp[0];
}

// Statistics of a four-wide tree: every Node4's box (the union of its
// children) and every leaf's box weighted by its triangle count, over the
// root's area.  The same measure for either build.
__global__ void __launch_bounds__(256) k_tree_stats4(const Node4* nodes, uint32_t nn, StatPart* parts) {
    __shared__ StatPart sp[256];
    __shared__ float s_inv;
    if (threadIdx.x == 0) {
        float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
        const Node4 r = nodes[0];
        for (int q = 0; q < 4; ++q) {
            if (u4_get(r.ref, q) == kEmptyRef) continue;
            l[0] = fminf(l[0], f4_get(r.lox, q)); h[0] = fmaxf(h[0], f4_get(r.hix, q));
            l[1] = fminf(l[1], f4_get(r.loy, q)); h[1] = fmaxf(h[1], f4_get(r.hiy, q));
            l[2] = fminf(l[2], f4_get(r.loz, q)); h[2] = fmaxf(h[2], f4_get(r.hiz, q));
        }
        const float a = half_area(l, h);
        s_inv = a > 0.f ? 1.0f / a : 0.f;
    }
    __syncthreads();
    const float inv = s_inv;
    StatPart v{0ull, 0ull, 0u, 0u, 0.0};
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += gridDim.x * blockDim.x) {
        const Node4 nd = nodes[n];
        float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int q = 0; q < 4; ++q) {
            const uint32_t r = u4_get(nd.ref, q);
            if (r == kEmptyRef) continue;
            const float cl[3] = {f4_get(nd.lox, q), f4_get(nd.loy, q), f4_get(nd.loz, q)};
            const float ch[3] = {f4_get(nd.hix, q), f4_get(nd.hiy, q), f4_get(nd.hiz, q)};
            for (int k = 0; k < 3; ++k) {
                l[k] = fminf(l[k], cl[k]);
                h[k] = fmaxf(h[k], ch[k]);
            }
            if (ref_is_leaf(r)) {
                v.leaves++;
                v.max_leaf = max(v.max_leaf, leaf_count(r));
                v.sah += (double)(half_area(cl, ch) * inv) * (double)leaf_count(r);
            }
        }
        v.internal++;
        v.sah += (double)(half_area(l, h) * inv);
    }
    stats_reduce(sp, v);
    if (threadIdx.x == 0) parts[blockIdx.x] = sp[0];
}

float host_area(const float lo[3], const float hi[3]) {
    float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    return ex * ey + ey * ez + ez * ex;
}
float key2f_h(int k) {
    int i = k >= 0 ? k : (k ^ 0x7fffffff);
    float f;
    std::memcpy(&f, &i, 4);
    return f;
}


}  // namespace

void blas_refit(Ctx& c, MeshG& m) {
    Blas& B = m.blas;
    VG_CHECK(B.built, VIRA_GPU_ERR_INVALID, "blas_refit: the mesh has no built BLAS");
    if (m.nt == 0) return;
    if (B.width != 4 || B.level_start.size() < 2) {
        blas_build(c, m, B.opts);
        return;
    }
    EventPair ev;  // released on every path
    hipEvent_t e0 = ev.a, e1 = ev.b;
    VG_HIP(hipEventRecord(e0, c.stream));
    k_refit_tris<<<blocks_for(m.nt, 256), 256, 0, c.stream>>>(m.nt, m.idx.as<uint32_t>(), m.verts.as<GVertex>(),
                                                              B.tris.as<float4>());
    VG_HIP(hipGetLastError());
    for (size_t l = B.level_start.size() - 1; l-- > 0;) {
        const uint32_t first = B.level_start[l], count = B.level_start[l + 1] - first;
        if (count == 0) continue;
        k_refit4<<<blocks_for(4 * (size_t)count, kRefitThreads), kRefitThreads, 0, c.stream>>>(B.nodes.as<Node4>(), first, count,
                                                               B.tris.as<float4>());
        VG_HIP(hipGetLastError());
    }
    DBuf box;
    box.alloc(6 * sizeof(float));
    k_root_box4<<<1, 1, 0, c.stream>>>(B.nodes.as<Node4>(), box.as<float>());
    VG_HIP(hipGetLastError());
    VG_HIP(hipEventRecord(e1, c.stream));
    float hb[6];
    VG_HIP(hipMemcpyAsync(hb, box.p, sizeof(hb), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < 3; ++k) { B.lo[k] = hb[k]; B.hi[k] = hb[3 + k]; }
    float ms = 0.f;
    VG_HIP(hipEventElapsedTime(&ms, e0, e1));
    B.stats.build_ms = ms;
    if (B.compressed) compress_blas(c, B);  // the quantised copy follows the new boxes
}

// ---------------------------------------------------------------- grid build
// A mesh made from a DEM grid has its structure in the index buffer: triangle
// A of cell (i, j) is (i, j+1), (i+1, j), (i, j) and triangle B is (i+1, j),
// (i, j+1), (i+1, j+1) (image_utils.ipp:364-412), so its cell and half follow
// from two of its vertex ids.  The grid build orders the triangles by a
// space-filling key of their cells (Morton over the low bits of both axes, the
// longer axis's remaining bits above): each triangle fills its key's slot and
// one stream compaction of the slots gives the order.  The sequence is cut into
// leaves of max_leaf triangles under a complete four-wide tree: node k of a
// level holds items 4k..4k+3 of the level below, so on a square power-of-two
// grid every node is a 2x2 block of its children.  One pass writes the
// leaf-order triangles with the bottom level's boxes; the levels above are
// reduced by the refit kernels.  Every step is one pass over the triangles,
// the slots or the nodes; nothing returns to the host between levels.
struct GridKey {
    uint32_t W, H;    // samples
    int kx, ky, kmin; // ceil(log2) of the cell counts, and their minimum
};

__device__ __forceinline__ uint32_t spread_bits(uint32_t x) {  // bit b to bit 2b (16 bits)
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

__device__ __forceinline__ uint64_t grid_cell_key(const GridKey& g, uint32_t i, uint32_t j) {
    const uint32_t lm = (1u << g.kmin) - 1u;
    const uint64_t low = (uint64_t)spread_bits(i & lm) | ((uint64_t)spread_bits(j & lm) << 1);
    const uint64_t high = g.kx > g.ky ? (i >> g.kmin) : (j >> g.kmin);
    return (high << (2 * g.kmin)) | low;
}

// slot[2 * key(cell) + half] = triangle; err set where a triangle is not a grid triangle
__global__ void k_grid_slots(size_t nt, const uint32_t* __restrict__ idx, GridKey g, uint32_t* __restrict__ slot,
                             uint64_t nslots, uint32_t* __restrict__ err) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint32_t v0 = idx[3 * t], v1 = idx[3 * t + 1], v2 = idx[3 * t + 2];
    uint32_t i, j, half;
    if (v0 == v2 + g.W && v1 == v2 + 1u) {  // A: v2 = (i, j)
        i = v2 % g.W;
        j = v2 / g.W;
        half = 0u;
    } else if (v2 == v0 + g.W && v1 + 1u == v2) {  // B: v0 = (i + 1, j)
        i = v0 % g.W - 1u;
        j = v0 / g.W;
        half = 1u;
    } else {
        atomicOr(err, 1u);
        return;
    }
    if (i + 1u >= g.W || j + 1u >= g.H) {
        atomicOr(err, 1u);
        return;
    }
    const uint64_t p = 2u * grid_cell_key(g, i, j) + half;
    if (p >= nslots) {
        atomicOr(err, 1u);
        return;
    }
    slot[p] = (uint32_t)t;
}

// Levels of the complete four-wide tree, root level first: level l holds
// count[l] nodes from first[l]; the level below the last holds the leaves.
struct GridLevels {
    int n;
    uint32_t first[24], count[24];
};

// Child references of every node (boxes are left to the refit).
__global__ void k_grid_topo(GridLevels L, uint32_t total, uint32_t nt, uint32_t leaf, uint32_t nleaves,
                            Node4* __restrict__ nodes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int l = 0;
    while (l + 1 < L.n && i >= L.first[l + 1]) ++l;
    const uint32_t k = i - L.first[l];
    uint32_t r[4];
    for (int q = 0; q < 4; ++q) {
        const uint32_t child = 4u * k + (uint32_t)q;
        if (l + 1 < L.n) {
            r[q] = child < L.count[l + 1] ? L.first[l + 1] + child : kEmptyRef;
        } else if (child < nleaves) {
            const uint32_t f = child * leaf;
            r[q] = make_leaf(f, min(leaf, nt - f));
        } else {
            r[q] = kEmptyRef;
        }
    }
    Node4 o;
    o.lox = o.hix = o.loy = o.hiy = o.loz = o.hiz = make_float4(0.f, 0.f, 0.f, 0.f);
    o.ref = make_uint4(r[0], r[1], r[2], r[3]);
    o.pad = make_uint4(0u, 0u, 0u, 0u);
    nodes[i] = o;
}

// The grid build's leaf-order triangles and its bottom level in one pass.  Leaf
// c holds triangles [c * leaf, min((c + 1) * leaf, nt)) and bottom node k its
// leaves 4k .. 4k + 3 (leaf divides kEmitThreads / 4, so a block covers whole
// nodes).  The block stages its triangles in LDS as k_emit_tris does, writes
// them as one run, and one thread per leaf then reduces the leaf's box from
// LDS and writes that child of its node: bounds, reference (empty past the
// last leaf, with a zero box, as k_grid_topo and k_refit4 leave one).
__global__ __launch_bounds__(kEmitThreads) void k_grid_emit(size_t nt, const uint32_t* prims, const uint32_t* idx,
                                                           const GVertex* V, float4* tris, Node4* __restrict__ bottom,
                                                           uint32_t nbottom, uint32_t leaf, uint32_t nleaves) {
    __shared__ float4 st[3 * kEmitThreads];
    const size_t base = (size_t)blockIdx.x * kEmitThreads;
    const size_t s = base + threadIdx.x;
    if (s < nt) {
        const uint32_t p = prims[s];
        const GVertex& a = V[idx[3 * (size_t)p]];
        const GVertex& b = V[idx[3 * (size_t)p + 1]];
        const GVertex& c = V[idx[3 * (size_t)p + 2]];
        st[3 * threadIdx.x] = make_float4(a.px, a.py, a.pz, __uint_as_float(p));
        st[3 * threadIdx.x + 1] = make_float4(b.px, b.py, b.pz, 0.f);
        st[3 * threadIdx.x + 2] = make_float4(c.px, c.py, c.pz, 0.f);
    }
    block_sync();
    const uint32_t nblk = (uint32_t)min((size_t)kEmitThreads, nt - base);
    float4* out = tris + 3 * base;
    for (uint32_t k = threadIdx.x; k < 3u * nblk; k += kEmitThreads) out[k] = st[k];
    if (threadIdx.x >= kEmitThreads / leaf) return;
    const uint32_t c = (uint32_t)(base / leaf) + threadIdx.x;  // this thread's leaf
    const uint32_t node = c >> 2, slot = c & 3u;
    if (node >= nbottom) return;
    Node4* nd = bottom + node;
    float l0 = 0.f, l1 = 0.f, l2 = 0.f, h0 = 0.f, h1 = 0.f, h2 = 0.f;
    uint32_t r = kEmptyRef;
    if (c < nleaves) {
        const uint32_t f = c * leaf, n = (uint32_t)min((size_t)leaf, nt - f);
        l0 = l1 = l2 = INFINITY;
        h0 = h1 = h2 = -INFINITY;
        const float4* q = st + 3 * (f - (uint32_t)base);
        for (uint32_t v = 0; v < 3u * n; ++v) {
            const float4 x = q[v];
            l0 = fminf(l0, x.x); l1 = fminf(l1, x.y); l2 = fminf(l2, x.z);
            h0 = fmaxf(h0, x.x); h1 = fmaxf(h1, x.y); h2 = fmaxf(h2, x.z);
        }
        r = make_leaf(f, n);
    }
    reinterpret_cast<float*>(&nd->lox)[slot] = l0;
    reinterpret_cast<float*>(&nd->loy)[slot] = l1;
    reinterpret_cast<float*>(&nd->loz)[slot] = l2;
    reinterpret_cast<float*>(&nd->hix)[slot] = h0;
    reinterpret_cast<float*>(&nd->hiy)[slot] = h1;
    reinterpret_cast<float*>(&nd->hiz)[slot] = h2;
    reinterpret_cast<uint32_t*>(&nd->ref)[slot] = r;
    reinterpret_cast<uint32_t*>(&nd->pad)[slot] = 0u;
}

// ---------------------------------------------------------------- quantised copy
// Node4 -> Node4Q.  Per node and axis the origin is the lowest child bound and
// the scale the smallest power of two with which 255 steps reach the highest
// (checked with the decoding fmaf itself); each child's lower byte is the
// largest whose decoded value is at most its bound, the upper byte the
// smallest whose decoded value is at least its bound.  err is set if a node
// cannot be encoded (never for finite boxes).
__global__ void k_compress4(const Node4* __restrict__ in, uint32_t n, Node4Q* __restrict__ out,
                            uint32_t* __restrict__ err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Node4 nd = in[i];
    Node4Q q;
    q.ref = nd.ref;
    q.pad[0] = q.pad[1] = 0u;
    q.ex = 0u;
    const float4 L[3] = {nd.lox, nd.loy, nd.loz}, Hh[3] = {nd.hix, nd.hiy, nd.hiz};
    float org[3];
    for (int a = 0; a < 3; ++a) {
        float mn = INFINITY, mx = -INFINITY;
        for (int c = 0; c < 4; ++c) {
            if (u4_get(nd.ref, c) == kEmptyRef) continue;
            mn = fminf(mn, f4_get(L[a], c));
            mx = fmaxf(mx, f4_get(Hh[a], c));
        }
        uint32_t qlo = 0u, qhi = 0u, ex = 127u;
        if (!(mn <= mx)) {  // no child: nothing to bound
            org[a] = 0.0f;
        } else {
            org[a] = mn;
            const double ext = (double)mx - (double)mn;
            int e = 1;
            if (ext > 0.0) {
                int k = 0;
                (void)frexp(ext / 255.0, &k);  // 2^k >= ext / 255
                e = max(1, k + 127);
            }
            while (e < 254 && __builtin_fmaf(255.0f, __uint_as_float((uint32_t)e << 23), mn) < mx) ++e;
            if (__builtin_fmaf(255.0f, __uint_as_float((uint32_t)e << 23), mn) < mx) atomicOr(err, 1u);
            ex = (uint32_t)e;
            const float sc = __uint_as_float(ex << 23);
            for (int c = 0; c < 4; ++c) {
                if (u4_get(nd.ref, c) == kEmptyRef) continue;
                const float lo = f4_get(L[a], c), hi = f4_get(Hh[a], c);
                int ql = (int)fmin(255.0, fmax(0.0, floor(((double)lo - (double)mn) / (double)sc)));
                while (ql > 0 && __builtin_fmaf((float)ql, sc, mn) > lo) --ql;
                int qh = (int)fmin(255.0, fmax(0.0, ceil(((double)hi - (double)mn) / (double)sc)));
                while (qh < 255 && __builtin_fmaf((float)qh, sc, mn) < hi) ++qh;
                if (__builtin_fmaf((float)ql, sc, mn) > lo || __builtin_fmaf((float)qh, sc, mn) < hi) atomicOr(err, 1u);
                qlo |= (uint32_t)ql << (8 * c);
                qhi |= (uint32_t)qh << (8 * c);
            }
        }
        q.qlo[a] = qlo;
        q.qhi[a] = qhi;
        q.ex |= ex << (8 * a);
    }
    q.ox = org[0];
    q.oy = org[1];
    q.oz = org[2];
    out[i] = q;
}

// The quantised copy of a built four-wide tree (kept beside the Node4s).
void compress_blas(Ctx& c, Blas& B) {
    VG_CHECK(B.width == 4, VIRA_GPU_ERR_INVALID, "BVH compression needs a four-wide tree");
    const uint32_t n = (uint32_t)B.num_nodes;
    B.qnodes.alloc((size_t)n * sizeof(Node4Q));
    DBuf err;
    err.alloc(4);
    VG_HIP(hipMemsetAsync(err.p, 0, 4, c.stream));
    k_compress4<<<blocks_for(n, 256), 256, 0, c.stream>>>(B.nodes.as<Node4>(), n, B.qnodes.as<Node4Q>(),
                                                          err.as<uint32_t>());
    VG_HIP(hipGetLastError());
    uint32_t herr = 0;
    VG_HIP(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    VG_CHECK(herr == 0, VIRA_GPU_ERR_HIP, "BVH compression: a node's boxes could not be encoded");
    B.compressed = true;
    B.stats.compressed_bytes = (uint64_t)n * sizeof(Node4Q);
}

// Statistics of a built Node4 tree (wide SAH, leaves, largest leaf).
StatPart wide_stats(Ctx& c, const Blas& B) {
    const uint32_t nn = (uint32_t)B.num_nodes;
    const uint32_t sblocks = std::min<uint32_t>(kStatsBlocks, blocks_for(nn, 256));
    DBuf sbuf;
    sbuf.alloc((sblocks + 1) * sizeof(StatPart));
    StatPart* parts = sbuf.as<StatPart>();
    k_tree_stats4<<<sblocks, 256, 0, c.stream>>>(B.nodes.as<Node4>(), nn, parts);
    VG_HIP(hipGetLastError());
    k_tree_stats_sum<<<1, 256, 0, c.stream>>>(parts, sblocks, parts + sblocks);
    VG_HIP(hipGetLastError());
    StatPart hs;
    VG_HIP(hipMemcpyAsync(&hs, parts + sblocks, sizeof(hs), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    return hs;
}

int ceil_log2(uint32_t n) {
    int k = 0;
    while ((1u << k) < n) ++k;
    return k;
}

// The grid build (see the kernels above); the mesh has grid provenance and
// at least one triangle.  A leaf holds the largest power of two of triangles
// up to max_leaf and 8 (whole cells, so Morton-aligned blocks: 1, 2, 4 or 8;
// 3 or 6 would straddle cells and measured 2.7x slower to render).
void blas_build_grid(Ctx& c, MeshG& m, int max_leaf) {
    const size_t N = m.nt;
    Blas& B = m.blas;
    EventPair ev;
    VG_HIP(hipEventRecord(ev.a, c.stream));
    GridKey g{(uint32_t)m.grid_w, (uint32_t)m.grid_h, ceil_log2((uint32_t)m.grid_w - 1u),
              ceil_log2((uint32_t)m.grid_h - 1u), 0};
    g.kmin = std::min(g.kx, g.ky);
    VG_CHECK(g.kx <= 16 && g.ky <= 16, VIRA_GPU_ERR_INVALID, "grid BVH: more than 65536 cells on an axis");
    const uint64_t nslots = 2ull << (g.kx + g.ky);
    VG_CHECK(nslots < (1ull << 31), VIRA_GPU_ERR_INVALID, "grid BVH: key space too large");
    DBuf slot, prims, err;
    slot.alloc(nslots * 4);
    prims.alloc(N * 4);
    err.alloc(4);
    VG_HIP(hipMemsetAsync(slot.p, 0xff, nslots * 4, c.stream));
    VG_HIP(hipMemsetAsync(err.p, 0, 4, c.stream));
    k_grid_slots<<<blocks_for(N, 256), 256, 0, c.stream>>>(N, m.idx.as<uint32_t>(), g, slot.as<uint32_t>(), nslots,
                                                           err.as<uint32_t>());
    VG_HIP(hipGetLastError());
    // the placed triangles in key order (each triangle fills at most one slot, so at most N are placed)
    const uint32_t placed = compact_u32(c, slot.as<uint32_t>(), prims.as<uint32_t>(), nslots);
    uint32_t herr = 0;
    VG_HIP(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    VG_CHECK(herr == 0 && placed == N, VIRA_GPU_ERR_INVALID, "grid BVH: the mesh's triangles are not its grid's cells");
    // levels of the complete four-wide tree, bottom up, then numbered from the root
    uint32_t leaf = 1;
    while (leaf * 2u <= (uint32_t)max_leaf && leaf < 8u) leaf *= 2u;
    const uint32_t nleaves = (uint32_t)((N + leaf - 1) / leaf);
    std::vector<uint32_t> counts;
    for (uint32_t n = nleaves; ; ) {
        n = (n + 3) / 4;
        counts.push_back(n);
        if (n == 1) break;
    }
    GridLevels L{};
    L.n = (int)counts.size();
    VG_CHECK(L.n <= 24, VIRA_GPU_ERR_INVALID, "grid BVH: too many levels");
    uint32_t total = 0;
    B.level_start.clear();
    for (int l = 0; l < L.n; ++l) {
        L.first[l] = total;
        L.count[l] = counts[L.n - 1 - l];
        B.level_start.push_back(total);
        total += L.count[l];
    }
    B.level_start.push_back(total);
    B.width = 4;
    B.num_nodes = total;
    B.nodes.alloc((size_t)total * sizeof(Node4));
    // the levels above the bottom get their references here; the bottom level, its leaves'
    // boxes and the leaf-order triangles come from one pass over the triangles
    const uint32_t upper = L.first[L.n - 1];
    if (upper > 0) {
        k_grid_topo<<<blocks_for(upper, 256), 256, 0, c.stream>>>(L, upper, (uint32_t)N, leaf, nleaves,
                                                                  B.nodes.as<Node4>());
        VG_HIP(hipGetLastError());
    }
    B.tris.alloc(N * 3 * sizeof(float4));
    k_grid_emit<<<blocks_for(N, kEmitThreads), kEmitThreads, 0, c.stream>>>(
        N, prims.as<uint32_t>(), m.idx.as<uint32_t>(), m.verts.as<GVertex>(), B.tris.as<float4>(),
        B.nodes.as<Node4>() + upper, L.count[L.n - 1], leaf, nleaves);
    VG_HIP(hipGetLastError());
    for (int l = L.n - 2; l >= 0; --l) {
        k_refit4<<<blocks_for(4 * (size_t)L.count[l], kRefitThreads), kRefitThreads, 0, c.stream>>>(B.nodes.as<Node4>(), L.first[l], L.count[l],
                                                                    B.tris.as<float4>());
        VG_HIP(hipGetLastError());
    }
    DBuf box;
    box.alloc(6 * sizeof(float));
    k_root_box4<<<1, 1, 0, c.stream>>>(B.nodes.as<Node4>(), box.as<float>());
    VG_HIP(hipGetLastError());
    VG_HIP(hipEventRecord(ev.b, c.stream));
    float hb[6];
    VG_HIP(hipMemcpyAsync(hb, box.p, sizeof(hb), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < 3; ++k) { B.lo[k] = hb[k]; B.hi[k] = hb[3 + k]; }
    float ms = 0.f;
    VG_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
    const StatPart hs = wide_stats(c, B);
    B.stats = vira_gpu_bvh_stats{};
    B.stats.num_nodes = nleaves - 1;  // a binary tree over the same leaves
    B.stats.num_wide_nodes = total;
    B.stats.width = 4;
    B.stats.wide_depth = (uint32_t)L.n;
    B.stats.num_leaves = hs.leaves;
    B.stats.max_depth = (uint32_t)L.n + 1;
    B.stats.max_leaf_size = hs.max_leaf;
    B.stats.sah_cost = 0.0;
    B.stats.wide_sah_cost = hs.sah;
    B.stats.method = 2;
    B.stats.build_ms = ms;
    B.stats.bytes = (uint64_t)total * sizeof(Node4) + N * 3 * sizeof(float4);
    B.built = true;
}

// After any build: the quantised copy when asked for, else none.
void finish_compression(Ctx& c, Blas& B, const vira_gpu_bvh_opts& opt) {
    B.compressed = false;
    B.qnodes.release();
    B.stats.compressed_bytes = 0;
    if (opt.compress && B.built && B.width == 4) compress_blas(c, B);
}

void blas_build(Ctx& c, MeshG& m, const vira_gpu_bvh_opts& opt) {
    const int max_leaf = std::min(15, std::max(1, opt.max_leaf > 0 ? opt.max_leaf : 4));
    m.blas.compressed = false;
    m.blas.qnodes.release();
    const bool grid_ok = m.grid_w >= 2 && m.grid_h >= 2 && opt.width != 2 && m.nt > 0;
    VG_CHECK(opt.method >= 0 && opt.method <= 2, VIRA_GPU_ERR_INVALID, "blas_build: method must be 0, 1 or 2");
    VG_CHECK(opt.method != 2 || grid_ok, VIRA_GPU_ERR_INVALID,
             "blas_build: the grid build takes a mesh made from a DEM grid, width 4");
    VG_CHECK(!opt.compress || opt.width != 2, VIRA_GPU_ERR_INVALID, "blas_build: compression needs width 4");
    if (opt.method == 2 || (opt.method == 0 && grid_ok && kGridByDefault)) {
        m.blas.built = false;
        blas_build_grid(c, m, max_leaf);
        finish_compression(c, m.blas, opt);
        return;
    }
    const int nb = std::min(kMaxBins, std::max(4, opt.bins > 0 ? opt.bins : 16));
    const size_t N = m.nt;
    Blas& B = m.blas;
    B.built = false;
    B.level_start.clear();
    VG_CHECK(N < (1u << 27), VIRA_GPU_ERR_INVALID, "BLAS limited to 2^27 triangles");
    EventPair ev;  // released on every path
    hipEvent_t e0 = ev.a, e1 = ev.b;
    VG_HIP(hipEventRecord(e0, c.stream));
    if (N == 0) {
        // one root with only empty children, valid as a Node2 (refs at word 12) and as a Node4 (refs at word 24)
        B.width = opt.width == 2 ? 2 : 4;
        B.nodes.alloc(sizeof(Node4));
        uint32_t empty[32] = {};
        empty[12] = empty[13] = kEmptyRef;
        for (int k = 24; k < 28; ++k) empty[k] = kEmptyRef;
        VG_HIP(hipMemcpyAsync(B.nodes.p, empty, sizeof(empty), hipMemcpyHostToDevice, c.stream));
        B.tris.alloc(16);
        B.num_nodes = 1;
        VG_HIP(hipStreamSynchronize(c.stream));
        B.stats = vira_gpu_bvh_stats{};
        B.stats.num_nodes = 1;
        for (int k = 0; k < 3; ++k) { B.lo[k] = 1.f; B.hi[k] = -1.f; }  // empty box: never entered
        B.built = true;
        return;
    }
    // The build's scratch, in the context's arena.  Level buffers are sized once for
    // the largest level: a large node holds more than kSmall triangles, and its chunks
    // number at most count / kChunk + 1.  Small subtrees: scratch nodes (region
    // 2 begin of each), nodes used, first node, and counter words 0 (unused), 1 (max
    // depth), 2 (builder error flag; word 1 later becomes the emit error flag).
    constexpr int kMaxLevels = 256;
    const size_t tasks_cap = N / (kSmall + 1) + 1, chunks_cap = N / kChunk + tasks_cap + 1;
    const size_t ssum_words = 2 * scan_sums_words(std::max(tasks_cap, chunks_cap));  // room for a pair of scans
    Span pbuf, prims, prims2, nodes, keys, slot_bin;  // prims2: the second id buffer of the large levels
    Span d_task[2], d_params, d_nchunk, d_chunk0, d_chunks, d_bins, d_counts, d_bases, d_left, d_lscan, d_sweep,
        d_nlarge, d_nsmall, d_loff, d_soff, d_small, d_cnt, d_ssum;
    Span snodes, d_used, d_first, d_counter, d_rec, d_cinfo;
    {
        ScratchPlan plan;
        plan.add(pbuf, N * sizeof(PrimBox));
        plan.add(slot_bin, N);
        plan.add(prims, N * 4);
        plan.add(prims2, N * 4);
        plan.add(nodes, 2 * N * sizeof(BNodeG));
        plan.add(keys, 12 * 4);
        for (auto& t : d_task) plan.add(t, tasks_cap * sizeof(LTask));
        plan.add(d_params, tasks_cap * sizeof(TaskParam));
        plan.add(d_nchunk, tasks_cap * 4);
        plan.add(d_chunk0, tasks_cap * 4);
        plan.add(d_chunks, chunks_cap * sizeof(Chunk));
        plan.add(d_bins, tasks_cap * nb * kBinVals * 4);
        plan.add(d_counts, chunks_cap * nb * 4);
        plan.add(d_bases, chunks_cap * sizeof(ChunkBase));
        plan.add(d_left, chunks_cap * 4);
        plan.add(d_lscan, chunks_cap * 4);
        plan.add(d_sweep, tasks_cap * sizeof(SweepOut));
        plan.add(d_nlarge, tasks_cap * 4);
        plan.add(d_nsmall, tasks_cap * 4);
        plan.add(d_loff, tasks_cap * 4);
        plan.add(d_soff, tasks_cap * 4);
        plan.add(d_small, N * sizeof(SmallTask));  // disjoint subtrees of at least one triangle each
        plan.add(d_cnt, (LC_WORDS + 1) * 4);        // + one word for scan totals nobody reads
        plan.add(d_ssum, ssum_words * 4);
        plan.add(snodes, std::max<size_t>(2, 2 * N) * sizeof(BNodeG));
        plan.add(d_used, N * 4);
        plan.add(d_first, N * 4);
        plan.add(d_counter, 12);
        plan.add(d_rec, (2 * kMaxLevels + 2) * 4);  // per level: first small subtree, count; last pair: the root's
        plan.add(d_cinfo, (4 + kTopMaxLevels) * 4);  // the collapse's first levels (k_collapse4_top)
        plan.place(c.bvh_arena);
    }
    k_prim_boxes<<<blocks_for(N, 256), 256, 0, c.stream>>>(N, m.idx.as<uint32_t>(), m.verts.as<GVertex>(),
                                                            pbuf.as<PrimBox>(), prims.as<uint32_t>());
    VG_HIP(hipGetLastError());
    int hk[12];
    for (int k = 0; k < 12; ++k) hk[k] = (k % 6) < 3 ? INT_MAX : INT_MIN;
    VG_HIP(hipMemcpyAsync(keys.p, hk, sizeof(hk), hipMemcpyHostToDevice, c.stream));
    k_reduce_bounds<<<std::min<unsigned>(1024, blocks_for(N, 256)), 256, 0, c.stream>>>(N, pbuf.as<PrimBox>(),
                                                                                      keys.as<int>());
    VG_HIP(hipGetLastError());

    uint32_t* cnt = d_cnt.as<uint32_t>();
    uint32_t* ssum = d_ssum.as<uint32_t>();
    VG_HIP(hipMemsetAsync(d_counter.p, 0, 12, c.stream));
    k_lv_root<<<1, 1, 0, c.stream>>>(keys.as<int>(), (uint32_t)N, nodes.as<BNodeG>(), d_task[0].as<LTask>(),
                                     d_small.as<SmallTask>(), cnt);
    VG_HIP(hipGetLastError());
    const float trav_cost = opt.traversal_cost > 0.0f ? opt.traversal_cost : 1.0f;
    uint32_t* rec = d_rec.as<uint32_t>();
    auto build_small = [&](hipStream_t st, const uint32_t* range, uint32_t grid) {
        if (grid == 0) return;
        k_build_small<<<grid, 64, 0, st>>>(d_small.as<SmallTask>(), range, prims.as<uint32_t>(),
                                           prims2.as<uint32_t>(), pbuf.as<PrimBox>(), nodes.as<BNodeG>(),
                                           snodes.as<BNodeG>(), d_used.as<uint32_t>(),
                                           d_counter.as<uint32_t>() + 1, d_counter.as<uint32_t>() + 2, max_leaf,
                                           trav_cost);
        VG_HIP(hipGetLastError());
    };

    // The level loop is queued without reading anything back, except each
    // level's count of large nodes for the next one, copied to pinned memory:
    // the host waits for level L - 1's count after queueing level L, so the
    // GPU always has the next level queued, and at most one empty level runs.
    // The small subtrees a level emits are built on the second stream while
    // the next levels run (their triangles and nodes are disjoint from those
    // of every later level); a level of n large nodes emits at most 2 n.
    c.bvh_counts.alloc((kMaxLevels + LC_WORDS + 12 + 4 + kTopMaxLevels) * 4);
    uint32_t* pin = c.bvh_counts.as<uint32_t>();
    if (!c.aux_stream) VG_HIP(hipStreamCreateWithFlags(&c.aux_stream, hipStreamNonBlocking));
    struct Events {  // destroyed on every path out of the build
        std::vector<hipEvent_t> e;
        hipEvent_t make() {
            hipEvent_t x;
            VG_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
            e.push_back(x);
            return x;
        }
        ~Events() { for (hipEvent_t x : e) (void)hipEventDestroy(x); }
    } evs;
    struct Drain {  // on every path out, both streams are idle before the buffers above are freed
        hipStream_t a, b;
        ~Drain() {
            (void)hipStreamSynchronize(b);
            (void)hipStreamSynchronize(a);
        }
    } drain{c.stream, c.aux_stream};
    std::vector<hipEvent_t> level_ev;
    const unsigned task_blocks = (unsigned)std::min<size_t>(kLevelBlocks, blocks_for(tasks_cap, 256));
    const unsigned chunk_blocks = (unsigned)std::min<size_t>(4 * kLevelBlocks, chunks_cap);
    const unsigned bin_blocks = (unsigned)std::min<size_t>(kLevelBlocks, chunks_cap);  // all resident at once
    const unsigned chunk_tblocks = (unsigned)std::min<size_t>(kLevelBlocks, blocks_for(chunks_cap, 256));
    const unsigned sweep_blocks = (unsigned)std::min<size_t>(4 * kLevelBlocks, blocks_for(tasks_cap, 64));
    if (N <= kSmall) {
        const uint32_t root_range[2] = {0u, 1u};
        VG_HIP(hipMemcpyAsync(rec + 2 * kMaxLevels, root_range, sizeof(root_range), hipMemcpyHostToDevice, c.stream));
        build_small(c.stream, rec + 2 * kMaxLevels, 1);
    }
    for (int L = 0; N > kSmall; ++L) {
        VG_CHECK(L < kMaxLevels, VIRA_GPU_ERR_HIP, "BVH build: too many large-node levels");
        const int par = L & 1;
        const uint32_t* d_nt = cnt + (par ? LC_NT1 : LC_NT0);
        uint32_t* d_ntn = cnt + (par ? LC_NT0 : LC_NT1);
        const LTask* cur = d_task[par].as<LTask>();
        LTask* nxt = d_task[par ^ 1].as<LTask>();
        k_lv_plan<<<task_blocks, 256, 0, c.stream>>>(cur, d_nt, nb, d_params.as<TaskParam>(),
                                                     d_nchunk.as<uint32_t>(), cnt);
        VG_HIP(hipGetLastError());
        scan_exclusive_u32_dev(c, d_nchunk.as<uint32_t>(), d_chunk0.as<uint32_t>(), d_nt, tasks_cap, ssum,
                               ssum_words, cnt + LC_NCH);
        k_lv_chunks<<<kLevelBlocks, 256, 0, c.stream>>>(cur, d_nt, d_params.as<TaskParam>(), d_chunk0.as<uint32_t>(),
                                                        d_chunks.as<Chunk>());
        VG_HIP(hipGetLastError());
        k_init_bins<<<kLevelBlocks, 256, 0, c.stream>>>(d_bins.as<int>(), d_nt, nb);
        VG_HIP(hipGetLastError());
        k_bin_large<<<bin_blocks, kChunkThreads, 0, c.stream>>>(
            d_chunks.as<Chunk>(), cnt + LC_NCH, d_params.as<TaskParam>(), nb, prims.as<uint32_t>(),
            prims2.as<uint32_t>(), pbuf.as<PrimBox>(), d_bins.as<int>(), d_counts.as<uint32_t>(),
            slot_bin.as<uint8_t>());
        VG_HIP(hipGetLastError());
        k_sweep_large<<<sweep_blocks, 64, 0, c.stream>>>(d_bins.as<int>(), nb, cur, d_nt, nodes.as<BNodeG>(),
                                                         d_params.as<TaskParam>(), d_sweep.as<SweepOut>(),
                                                         d_nlarge.as<uint32_t>(), d_nsmall.as<uint32_t>(), cnt);
        VG_HIP(hipGetLastError());
        k_chunk_left<<<chunk_tblocks, 256, 0, c.stream>>>(d_chunks.as<Chunk>(), cnt + LC_NCH,
                                                          d_params.as<TaskParam>(), d_counts.as<uint32_t>(), nb,
                                                          d_left.as<uint32_t>());
        VG_HIP(hipGetLastError());
        scan_exclusive_u32_dev(c, d_left.as<uint32_t>(), d_lscan.as<uint32_t>(), cnt + LC_NCH, chunks_cap, ssum,
                               ssum_words, cnt + LC_WORDS);
        k_chunk_bases<<<chunk_tblocks, 256, 0, c.stream>>>(d_chunks.as<Chunk>(), cnt + LC_NCH,
                                                           d_params.as<TaskParam>(), d_sweep.as<SweepOut>(),
                                                           d_lscan.as<uint32_t>(), d_bases.as<ChunkBase>());
        VG_HIP(hipGetLastError());
        // a node without a valid split leaves its triangles where they are (split -1: all go right at
        // offset nleft = 0, in order); the build is rejected below before they are used
        k_partition_large<<<chunk_blocks, kChunkThreads, 0, c.stream>>>(
            d_chunks.as<Chunk>(), cnt + LC_NCH, d_params.as<TaskParam>(), d_bases.as<ChunkBase>(), prims.as<uint32_t>(),
            prims2.as<uint32_t>(), slot_bin.as<uint8_t>());
        VG_HIP(hipGetLastError());
        scan_exclusive_u32x2_dev(c, d_nlarge.as<uint32_t>(), d_loff.as<uint32_t>(), d_ntn, d_nsmall.as<uint32_t>(),
                                 d_soff.as<uint32_t>(), cnt + LC_LSMALL, d_nt, tasks_cap, ssum, ssum_words);
        k_lv_emit<<<task_blocks, 256, 0, c.stream>>>(cur, d_nt, d_params.as<TaskParam>(), d_sweep.as<SweepOut>(),
                                                     d_nlarge.as<uint32_t>(), d_nsmall.as<uint32_t>(),
                                                     d_loff.as<uint32_t>(),
                                                     d_soff.as<uint32_t>(), nodes.as<BNodeG>(), nxt,
                                                     d_small.as<SmallTask>(), cnt);
        VG_HIP(hipGetLastError());
        k_lv_advance<<<1, 1, 0, c.stream>>>(cnt, par, rec + 2 * L);
        VG_HIP(hipGetLastError());
        VG_HIP(hipMemcpyAsync(pin + L, d_ntn, 4, hipMemcpyDeviceToHost, c.stream));
        level_ev.push_back(evs.make());
        VG_HIP(hipEventRecord(level_ev[L], c.stream));
        // level L holds 1 node (the root) or pin[L - 1]; its small subtrees are built beside the next levels
        uint32_t nodes_l = 1;
        if (L >= 1) {
            VG_HIP(hipEventSynchronize(level_ev[L - 1]));
            nodes_l = pin[L - 1];
        }
        if (nodes_l > 0) {
            VG_HIP(hipStreamWaitEvent(c.aux_stream, level_ev[L], 0));
            build_small(c.aux_stream, rec + 2 * L, 2 * nodes_l);
        }
        if (nodes_l == 0) break;  // level L (queued) has no nodes
    }
    // the main stream waits for the small subtrees, then the loop's results are read:
    // node and subtree counts, depth, error flags, root bounds
    {
        hipEvent_t done = evs.make();
        VG_HIP(hipEventRecord(done, c.aux_stream));
        VG_HIP(hipStreamWaitEvent(c.stream, done, 0));
    }
    // the triangles in leaf order, on the second stream beside the rest of the layout
    // (queued once the main stream has reached the point where this is called)
    hipEvent_t tris_done = evs.make();
    auto emit_tris_aside = [&]() {
        hipEvent_t ready = evs.make();
        VG_HIP(hipEventRecord(ready, c.stream));
        VG_HIP(hipStreamWaitEvent(c.aux_stream, ready, 0));
        B.tris.alloc(N * 3 * sizeof(float4));
        k_emit_tris<<<blocks_for(N, kEmitThreads), kEmitThreads, 0, c.aux_stream>>>(N, prims.as<uint32_t>(), m.idx.as<uint32_t>(),
                                                                  m.verts.as<GVertex>(), B.tris.as<float4>());
        VG_HIP(hipGetLastError());
        VG_HIP(hipEventRecord(tris_done, c.aux_stream));
    };
    uint32_t* hcnt = pin + kMaxLevels;
    int* hkeys = reinterpret_cast<int*>(pin + kMaxLevels + LC_WORDS);
    uint32_t* hsmall = pin + kMaxLevels + LC_WORDS + 6;  // counter words 1 and 2 of the small builder
    VG_HIP(hipMemcpyAsync(hcnt, cnt, LC_WORDS * 4, hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipMemcpyAsync(hkeys, keys.p, 6 * 4, hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipMemcpyAsync(hsmall, d_counter.as<uint32_t>() + 1, 8, hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    VG_CHECK(hcnt[LC_ERR] == 0, VIRA_GPU_ERR_HIP, "BVH build: no valid split for a binned node");
    VG_CHECK(hsmall[1] == 0, VIRA_GPU_ERR_HIP, "BVH build: small-subtree builder failed (stack, queue or partition)");
    for (int k = 0; k < 3; ++k) {
        B.lo[k] = key2f_h(hkeys[k]);
        B.hi[k] = key2f_h(hkeys[3 + k]);
    }
    const float root_area = host_area(B.lo, B.hi);
    uint32_t nn = hcnt[LC_NODE];
    const uint32_t max_depth = std::max(hcnt[LC_DEPTH], hsmall[0]);
    const uint32_t ntasks = hcnt[LC_SMALL];
    VG_CHECK(ntasks >= 1 && ntasks <= N && nn <= 2 * N, VIRA_GPU_ERR_HIP, "BVH build: level loop overflow");
    // the subtrees' nodes behind the large-level nodes, in subtree order
    const uint32_t small_nodes = scan_exclusive_u32(c, d_used.as<uint32_t>(), d_first.as<uint32_t>(), ntasks);
    VG_CHECK((size_t)nn + small_nodes <= 2 * N, VIRA_GPU_ERR_HIP, "BVH node count overflow");
    k_compact_small<<<ntasks, 64, 0, c.stream>>>(d_small.as<SmallTask>(), ntasks, d_used.as<uint32_t>(),
                                                 d_first.as<uint32_t>(), nn, snodes.as<BNodeG>(), nodes.as<BNodeG>());
    VG_HIP(hipGetLastError());
    nn += small_nodes;
    VG_CHECK(max_depth < 63, VIRA_GPU_ERR_HIP, "BVH deeper than the traversal stack");
    VG_CHECK(nn <= 2 * N, VIRA_GPU_ERR_HIP, "BVH node count overflow");

    // layout: internal nodes -> Node2 / Node4, triangles -> leaf order (counter word 1 becomes the error flag)
    VG_HIP(hipMemsetAsync(d_counter.as<char>() + 4, 0, 4, c.stream));
    uint32_t* d_err = reinterpret_cast<uint32_t*>(d_counter.as<char>() + 4);
    B.width = opt.width == 2 ? 2 : 4;
    uint32_t wide_depth = 0;
    if (B.width == 2) {
        emit_tris_aside();
        // flags and internal-node indices in the boxes' space (2 nn words, nn < 2 N)
        VG_CHECK((size_t)nn * 8 <= pbuf.bytes, VIRA_GPU_ERR_HIP, "BVH2 emission scratch");
        uint32_t* flag = pbuf.as<uint32_t>();
        uint32_t* iidx = flag + nn;
        k_flag_internal<<<blocks_for(nn, 256), 256, 0, c.stream>>>(nodes.as<BNodeG>(), nn, flag);
        VG_HIP(hipGetLastError());
        uint32_t ninternal = scan_exclusive_u32(c, flag, iidx, nn);
        B.num_nodes = std::max<uint32_t>(1, ninternal);
        B.nodes.alloc(B.num_nodes * sizeof(Node2));
        if (ninternal == 0) {
            k_emit_leaf_root<<<1, 1, 0, c.stream>>>(nodes.as<BNodeG>(), B.nodes.as<Node2>());
        } else {
            k_emit_nodes<<<blocks_for(nn, 256), 256, 0, c.stream>>>(nodes.as<BNodeG>(), nn, iidx, B.nodes.as<Node2>(),
                                                                    d_err);
        }
        VG_HIP(hipGetLastError());
        wide_depth = max_depth;
    } else {
        // a full binary tree of nn nodes has (nn - 1) / 2 internal nodes, each at most one Node4
        const size_t cap = std::max<size_t>(1, (nn - 1) / 2);
        B.nodes.alloc(cap * sizeof(Node4));
        if (nn == 1) {
            emit_tris_aside();
            k_emit_leaf_root4<<<1, 1, 0, c.stream>>>(nodes.as<BNodeG>(), B.nodes.as<Node4>());
            VG_HIP(hipGetLastError());
            B.num_nodes = 1;
            B.level_start = {0u, 1u};
        } else {
            // the level queues, counts and offsets in the boxes' space (24 cap bytes, cap < N)
            static_assert(2 * sizeof(CItem) + 8 <= sizeof(PrimBox), "collapse queues fit in the boxes' space");
            VG_CHECK(cap * (2 * sizeof(CItem) + 8) <= pbuf.bytes, VIRA_GPU_ERR_HIP, "BVH4 collapse scratch");
            CItem* qa = pbuf.as<CItem>();
            CItem* qb = qa + cap;
            uint32_t* qcnt = reinterpret_cast<uint32_t*>(qb + cap);
            uint32_t* qoff = qcnt + cap;
            CItem root{0u, 0u};
            VG_HIP(hipMemcpyAsync(qa, &root, sizeof(root), hipMemcpyHostToDevice, c.stream));
            // the first levels in one workgroup, then one count / scan / emit round per level
            k_collapse4_top<<<1, kTopItems, 0, c.stream>>>(nodes.as<BNodeG>(), qa, qb, (uint32_t)cap,
                                                           B.nodes.as<Node4>(), d_cinfo.as<uint32_t>(), d_err);
            VG_HIP(hipGetLastError());
            emit_tris_aside();  // after the latency-bound first levels, not beside them
            uint32_t* hinfo = pin + kMaxLevels + LC_WORDS + 12;
            VG_HIP(hipMemcpyAsync(hinfo, d_cinfo.p, (4 + kTopMaxLevels) * 4, hipMemcpyDeviceToHost, c.stream));
            VG_HIP(hipStreamSynchronize(c.stream));
            VG_CHECK(hinfo[1] <= cap && hinfo[2] <= (uint32_t)kTopMaxLevels, VIRA_GPU_ERR_HIP, "BVH4 collapse overflow");
            B.level_start = {0u, 1u};  // level 0 = the root; each level appends its end
            for (uint32_t l = 0; l < hinfo[2]; ++l) B.level_start.push_back(hinfo[4 + l]);
            wide_depth += hinfo[2];
            uint32_t total = hinfo[1], nin = hinfo[0];
            CItem* qin = hinfo[3] ? qb : qa;
            CItem* qout = hinfo[3] ? qa : qb;
            while (nin > 0) {
                k_collapse4_count<<<blocks_for(nin, 256), 256, 0, c.stream>>>(nodes.as<BNodeG>(), qin, nin,
                                                                              qcnt);
                VG_HIP(hipGetLastError());
                const uint32_t nnext = scan_exclusive_u32(c, qcnt, qoff, nin);
                VG_CHECK((size_t)total + nnext <= cap, VIRA_GPU_ERR_HIP, "BVH4 collapse overflow");
                k_collapse4_emit<<<blocks_for(nin, 256), 256, 0, c.stream>>>(nodes.as<BNodeG>(), qin, nin,
                                                                             qoff, total,
                                                                             B.nodes.as<Node4>(), qout, d_err);
                VG_HIP(hipGetLastError());
                total += nnext;
                B.level_start.push_back(total);
                nin = nnext;
                std::swap(qin, qout);
                wide_depth++;
            }
            B.num_nodes = total;
        }
    }
    VG_HIP(hipStreamWaitEvent(c.stream, tris_done, 0));
    VG_HIP(hipEventRecord(e1, c.stream));
    // statistics (outside the timed region)
    const uint32_t sblocks = std::min<uint32_t>(kStatsBlocks, blocks_for(nn, 256));
    DBuf sbuf;
    sbuf.alloc((sblocks + 1) * sizeof(StatPart));
    StatPart* parts = sbuf.as<StatPart>();
    k_tree_stats<<<sblocks, 256, 0, c.stream>>>(nodes.as<BNodeG>(), nn, root_area > 0.f ? 1.0f / root_area : 0.f,
                                               parts);
    VG_HIP(hipGetLastError());
    k_tree_stats_sum<<<1, 256, 0, c.stream>>>(parts, sblocks, parts + sblocks);
    VG_HIP(hipGetLastError());
    StatPart hs;
    VG_HIP(hipMemcpyAsync(&hs, parts + sblocks, sizeof(hs), hipMemcpyDeviceToHost, c.stream));
    uint32_t emit_err = 0;
    VG_HIP(hipMemcpyAsync(&emit_err, d_counter.as<char>() + 4, 4, hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    VG_CHECK(emit_err == 0, VIRA_GPU_ERR_HIP, "BVH build produced a leaf with more than 16 triangles");
    float ms = 0.f;
    VG_HIP(hipEventElapsedTime(&ms, e0, e1));
    B.stats = vira_gpu_bvh_stats{};
    B.stats.num_nodes = hs.internal;
    B.stats.num_wide_nodes = B.num_nodes;
    B.stats.width = (uint32_t)B.width;
    B.stats.wide_depth = wide_depth;
    B.stats.num_leaves = hs.leaves;
    B.stats.max_depth = max_depth + 1;
    B.stats.max_leaf_size = hs.max_leaf;
    B.stats.sah_cost = hs.sah;
    B.stats.build_ms = ms;
    B.stats.bytes = B.num_nodes * (B.width == 2 ? sizeof(Node2) : sizeof(Node4)) + N * 3 * sizeof(float4);
    B.stats.method = 1;
    if (B.width == 4) B.stats.wide_sah_cost = wide_stats(c, B).sah;
    B.built = true;
    finish_compression(c, B, opt);
}

}  // namespace vg

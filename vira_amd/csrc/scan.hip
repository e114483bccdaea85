// scan.hip -- device-wide exclusive prefix sum of uint32 (reduce, scan the
// block sums in one block, rescan and add).  Used by tessellation (triangle
// ids with holes), normal adjacency and the BVH build; the _dev form takes its
// element count from device memory and leaves the total there, so a caller can
// queue it without reading anything back.
#include "context.hpp"

namespace vg {
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;  // per thread (the apply pass loads them as two uint4)
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
    // wave-level inclusive scan
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < kScanBlock / 64; ++w) {
            uint32_t t = lds[w];
            lds[w] = s;
            s += t;
        }
        lds[kScanBlock / 64] = s;
    }
    __syncthreads();
    uint32_t excl = x - v + lds[wid];
    total = lds[kScanBlock / 64];
    __syncthreads();
    return excl;
}

// One or two independent scans per launch: array y of a launch is blockIdx.y
// (reduce, apply) or blockIdx.x (sums), and its block sums start at
// sums + y * sstride.
struct ScanArgs {
    const uint32_t* in[2];
    uint32_t* out[2];
    uint32_t* total[2];
};

constexpr uint32_t kNone = 0xffffffffu;  // COMPACT: the value that is dropped

// dn (optional): the element count in device memory, at most n (the grid is sized for n).
// COMPACT: the scanned value is 1 for an element other than kNone, else 0.
template <bool COMPACT>
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(ScanArgs a, size_t n, uint32_t* sums, size_t sstride,
                                                            const uint32_t* dn) {
    __shared__ uint32_t lds[kScanBlock / 64 + 1];
    if (dn) n = *dn;
    const uint32_t* in = a.in[blockIdx.y];
    size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + (size_t)k * kScanBlock + threadIdx.x;
        if (i < n) s += COMPACT ? (in[i] != kNone ? 1u : 0u) : in[i];
    }
    uint32_t total;
    (void)block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.y * sstride + blockIdx.x] = total;
}

// One block per array: exclusive scan of its block sums in place; writes its grand total.
__global__ void __launch_bounds__(kScanBlock) k_scan_sums(ScanArgs a, uint32_t* sums, size_t nb, size_t sstride) {
    __shared__ uint32_t lds[kScanBlock / 64 + 1];
    sums += blockIdx.x * sstride;
    uint32_t carry = 0;
    for (size_t base = 0; base < nb; base += kScanTile) {
        const size_t i0 = base + (size_t)threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        uint32_t s = 0;
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = i0 + k < nb ? sums[i0 + k] : 0;
            s += v[k];
        }
        uint32_t total;
        uint32_t e = block_exclusive_scan(s, lds, total) + carry;
        for (int k = 0; k < kScanItems; ++k) {
            if (i0 + k < nb) sums[i0 + k] = e;
            e += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) *a.total[blockIdx.x] = carry;
}

// COMPACT: each element other than kNone is written to out[its rank] instead of
// its prefix being written to out[i].
template <bool COMPACT>
__global__ void __launch_bounds__(kScanBlock) k_scan_apply(ScanArgs a, size_t n, const uint32_t* sums,
                                                           size_t sstride, const uint32_t* dn) {
    __shared__ uint32_t lds[kScanBlock / 64 + 1];
    if (dn) n = *dn;
    const uint32_t* in = a.in[blockIdx.y];
    uint32_t* out = a.out[blockIdx.y];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    // a thread's kScanItems consecutive elements: two 16-byte loads and stores where
    // the arrays are 16-byte aligned and the run is whole
    const bool vec = base + kScanItems <= n && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    if (vec) {
        const uint4 x = reinterpret_cast<const uint4*>(in + base)[0], y = reinterpret_cast<const uint4*>(in + base)[1];
        v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w; v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
    } else {
        for (int k = 0; k < kScanItems; ++k) v[k] = base + k < n ? in[base + k] : 0;
    }
    uint32_t f[kScanItems];  // the scanned values
    for (int k = 0; k < kScanItems; ++k) {
        f[k] = COMPACT ? (base + k < n && v[k] != kNone ? 1u : 0u) : v[k];
        s += f[k];
    }
    uint32_t total;
    uint32_t e = block_exclusive_scan(s, lds, total) + sums[blockIdx.y * sstride + blockIdx.x];
    if (COMPACT) {
        for (int k = 0; k < kScanItems; ++k) {
            if (f[k]) out[e] = v[k];
            e += f[k];
        }
        return;
    }
    uint32_t o[kScanItems];
    for (int k = 0; k < kScanItems; ++k) {
        o[k] = e;
        e += v[k];
    }
    if (vec) {
        reinterpret_cast<uint4*>(out + base)[0] = make_uint4(o[0], o[1], o[2], o[3]);
        reinterpret_cast<uint4*>(out + base)[1] = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < n) out[base + k] = o[k];
    }
}

// Queues the reduce / sums / apply launches for `arrays` scans of n elements.
template <bool COMPACT = false>
void scan_launch(Ctx& c, const ScanArgs& a, int arrays, size_t n, uint32_t* sums, size_t sstride,
                 const uint32_t* dn) {
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    k_scan_reduce<COMPACT><<<dim3((unsigned)nb, (unsigned)arrays), kScanBlock, 0, c.stream>>>(a, n, sums, sstride, dn);
    VG_HIP(hipGetLastError());
    k_scan_sums<<<(unsigned)arrays, kScanBlock, 0, c.stream>>>(a, sums, nb, sstride);
    VG_HIP(hipGetLastError());
    k_scan_apply<COMPACT><<<dim3((unsigned)nb, (unsigned)arrays), kScanBlock, 0, c.stream>>>(a, n, sums, sstride, dn);
    VG_HIP(hipGetLastError());
}

}  // namespace

uint32_t scan_exclusive_u32(Ctx& c, const uint32_t* in, uint32_t* out, size_t n, bool total_wanted) {
    if (n == 0) return 0;
    size_t nb = (n + kScanTile - 1) / kScanTile;
    c.scan_sums.alloc((nb + 1) * sizeof(uint32_t));
    c.scan_total.alloc(sizeof(uint32_t));
    uint32_t* s = c.scan_sums.as<uint32_t>();
    // k_scan_reduce sums each tile's elements in a strided order and
    // k_scan_apply scans them contiguously per thread: both cover the same
    // kScanTile elements of the tile, so the block totals agree.
    const ScanArgs a{{in, nullptr}, {out, nullptr}, {s + nb, nullptr}};
    scan_launch(c, a, 1, n, s, nb + 1, nullptr);
    if (!total_wanted) return 0;  // the scan stays queued on the stream
    uint32_t* total = c.scan_total.as<uint32_t>();
    VG_HIP(hipMemcpyAsync(total, s + nb, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    return *total;
}

uint32_t compact_u32(Ctx& c, const uint32_t* in, uint32_t* out, size_t n) {
    if (n == 0) return 0;
    size_t nb = (n + kScanTile - 1) / kScanTile;
    c.scan_sums.alloc((nb + 1) * sizeof(uint32_t));
    c.scan_total.alloc(sizeof(uint32_t));
    uint32_t* s = c.scan_sums.as<uint32_t>();
    const ScanArgs a{{in, nullptr}, {out, nullptr}, {s + nb, nullptr}};
    scan_launch<true>(c, a, 1, n, s, nb + 1, nullptr);
    uint32_t* total = c.scan_total.as<uint32_t>();
    VG_HIP(hipMemcpyAsync(total, s + nb, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    return *total;
}

size_t scan_sums_words(size_t n_max) { return (n_max + kScanTile - 1) / kScanTile + 1; }

void scan_exclusive_u32_dev(Ctx& c, const uint32_t* in, uint32_t* out, const uint32_t* dn, size_t n_max,
                            uint32_t* sums, size_t sums_words, uint32_t* total) {
    if (n_max == 0) return;
    const size_t nb = (n_max + kScanTile - 1) / kScanTile;
    VG_CHECK(sums_words >= nb + 1, VIRA_GPU_ERR_INVALID, "scan: block-sum scratch smaller than the scan needs");
    const ScanArgs a{{in, nullptr}, {out, nullptr}, {total, nullptr}};
    scan_launch(c, a, 1, n_max, sums, nb + 1, dn);
}

void scan_exclusive_u32x2_dev(Ctx& c, const uint32_t* in0, uint32_t* out0, uint32_t* total0, const uint32_t* in1,
                              uint32_t* out1, uint32_t* total1, const uint32_t* dn, size_t n_max, uint32_t* sums,
                              size_t sums_words) {
    if (n_max == 0) return;
    const size_t nb = (n_max + kScanTile - 1) / kScanTile;
    VG_CHECK(sums_words >= 2 * (nb + 1), VIRA_GPU_ERR_INVALID,
             "scan: block-sum scratch smaller than the two scans need");
    const ScanArgs a{{in0, in1}, {out0, out1}, {total0, total1}};
    scan_launch(c, a, 2, n_max, sums, nb + 1, dn);
}

}  // namespace vg

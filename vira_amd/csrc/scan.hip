// scan.hip -- device-wide exclusive prefix sum of uint32 (reduce, scan the
// block sums in one block, rescan and add).  Used by tessellation (triangle
// ids with holes), normal adjacency and the BVH build; the _dev form takes its
// element count from device memory and leaves the total there, so a caller can
// queue it without reading anything back.
#include "context.hpp"

namespace vg {
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;  // per thread
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

// dn (optional): the element count in device memory, at most n (the grid is sized for n)
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const uint32_t* in, size_t n, uint32_t* sums,
                                                            const uint32_t* dn) {
    __shared__ uint32_t lds[kScanBlock / 64 + 1];
    if (dn) n = *dn;
    size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + (size_t)k * kScanBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t total;
    (void)block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the block sums in place; writes grand total.
__global__ void __launch_bounds__(kScanBlock) k_scan_sums(uint32_t* sums, size_t nb, uint32_t* grand) {
    __shared__ uint32_t lds[kScanBlock / 64 + 1];
    uint32_t carry = 0;
    for (size_t base = 0; base < nb; base += kScanBlock) {
        size_t i = base + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0;
        uint32_t total;
        uint32_t e = block_exclusive_scan(v, lds, total);
        if (i < nb) sums[i] = e + carry;
        carry += total;
    }
    if (threadIdx.x == 0) *grand = carry;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_apply(const uint32_t* in, uint32_t* out, size_t n,
                                                           const uint32_t* sums, const uint32_t* dn) {
    __shared__ uint32_t lds[kScanBlock / 64 + 1];
    if (dn) n = *dn;
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    uint32_t total;
    uint32_t e = block_exclusive_scan(s, lds, total) + sums[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + k;
        if (i < n) out[i] = e;
        e += v[k];
    }
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
    k_scan_reduce<<<(unsigned)nb, kScanBlock, 0, c.stream>>>(in, n, s, nullptr);
    VG_HIP(hipGetLastError());
    k_scan_sums<<<1, kScanBlock, 0, c.stream>>>(s, nb, s + nb);
    VG_HIP(hipGetLastError());
    k_scan_apply<<<(unsigned)nb, kScanBlock, 0, c.stream>>>(in, out, n, s, nullptr);
    VG_HIP(hipGetLastError());
    if (!total_wanted) return 0;  // the scan stays queued on the stream
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
    k_scan_reduce<<<(unsigned)nb, kScanBlock, 0, c.stream>>>(in, n_max, sums, dn);
    VG_HIP(hipGetLastError());
    k_scan_sums<<<1, kScanBlock, 0, c.stream>>>(sums, nb, total);
    VG_HIP(hipGetLastError());
    k_scan_apply<<<(unsigned)nb, kScanBlock, 0, c.stream>>>(in, out, n_max, sums, dn);
    VG_HIP(hipGetLastError());
}

}  // namespace vg

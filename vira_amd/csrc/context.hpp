// context.hpp -- host-side state behind the vira_gpu C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/vira_gpu.h"
#include "device_math.hpp"

// An open quipu DEM pyramid: its header and level table (the levels are read
// on demand).
struct vira_gpu_qld {
    std::string path;
    int32_t levels = 0;
    bool has_albedos = false;
    double M[16] = {};
    float normal[3] = {};
    float cone = 0.0f;
    std::vector<double> gsd;
    std::vector<uint64_t> offset;
    uint64_t file_size = 0;
};

namespace vg {

// Error raised inside the library and turned into a status at the C-ABI.
struct Error : std::runtime_error {
    vira_gpu_status code;
    Error(vira_gpu_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define VG_HIP(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw ::vg::Error(e_ == hipErrorOutOfMemory ? VIRA_GPU_ERR_OOM : VIRA_GPU_ERR_HIP,        \
                              std::string(#call) + ": " + hipGetErrorString(e_));                    \
    } while (0)

#define VG_CHECK(cond, code, msg)                          \
    do {                                                   \
        if (!(cond)) throw ::vg::Error((code), (msg));     \
    } while (0)

// Owning device buffer.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        if (n <= bytes && p) return;
        release();
        if (n == 0) return;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(e == hipErrorOutOfMemory ? VIRA_GPU_ERR_OOM : VIRA_GPU_ERR_HIP,
                        "hipMalloc(" + std::to_string(n) + " B): " + hipGetErrorString(e));
        }
        bytes = n;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// A part of a scratch arena (not owning): placed by ScratchPlan.
struct Span {
    void* p = nullptr;
    size_t bytes = 0;
    template <class T> T* as() const { return static_cast<T*>(p); }
};
// Lays out scratch buffers of one call in one device allocation kept by the
// context: the arena grows to the largest call seen and is reused after that,
// so repeated calls neither allocate nor free (hipFree would also wait for
// the device).  Parts are 256-byte aligned.
struct ScratchPlan {
    std::vector<std::pair<Span*, size_t>> parts;
    void add(Span& s, size_t n) { parts.emplace_back(&s, n); }
    void place(DBuf& arena) {
        size_t total = 0;
        for (auto& pn : parts) total += (pn.second + 255) & ~size_t(255);
        arena.alloc(std::max<size_t>(total, 256));
        size_t off = 0;
        for (auto& pn : parts) {
            pn.first->p = arena.as<char>() + off;
            pn.first->bytes = pn.second;
            off += (pn.second + 255) & ~size_t(255);
        }
    }
};

// Owning page-locked host buffer (hipHostMalloc), for transfers that must not
// stage through pageable memory.
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { release(); }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        if (n <= bytes && p) return;
        release();
        if (n == 0) return;
        hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(e == hipErrorOutOfMemory ? VIRA_GPU_ERR_OOM : VIRA_GPU_ERR_HIP,
                        "hipHostMalloc(" + std::to_string(n) + " B): " + hipGetErrorString(e));
        }
        bytes = n;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Two HIP events owned for the duration of one call (released on every path).
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        VG_HIP(hipEventCreate(&a));
        hipError_t e = hipEventCreate(&b);
        if (e != hipSuccess) {
            (void)hipEventDestroy(a);
            a = nullptr;
            VG_HIP(e);
        }
    }
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

struct Blas {
    bool built = false;
    int width = 2;
    DBuf nodes;   // Node2[num_nodes] or Node4[num_nodes]
    DBuf tris;    // float4[3 * nt] in leaf order: v0.xyz+id, v1.xyz, v2.xyz
    uint64_t num_nodes = 0;
    vira_gpu_bvh_stats stats{};
    vira_gpu_bvh_opts opts{4, 16, 4};  // options of the last build (dem_update rebuilds with them)
    float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {0.f, 0.f, 0.f};  // box of all triangles (empty: lo > hi)
    std::vector<uint32_t> level_start;  // Node4 levels: level l is [level_start[l], level_start[l + 1])
    DBuf qnodes;              // Node4Q[num_nodes]: quantised copy (opts.compress), same indices
    bool compressed = false;
};

struct MeshG {
    bool alive = false;
    size_t nv = 0, nt = 0;
    DBuf verts;   // GVertex[nv]
    DBuf idx;     // uint32[3 * nt]
    DBuf fnrm;    // float4[nt]: unit face normal from the FP64 positions, w = 0 where skipped
    Blas blas;
    vira_gpu_material mat{};
    // grid provenance (dem_mesh, quipu levels): vertex v is sample (v % grid_w, v / grid_w) and the
    // triangles are the grid's cells in imageToIndexBuffer order, holes left out
    int grid_w = 0, grid_h = 0;
    // DEM provenance (dem_mesh): node v is grid sample (v % dem_w, v / dem_w)
    bool is_dem = false;
    int dem_w = 0, dem_h = 0;
    vira_gpu_projection dem_proj{};
    // vertex -> incident triangles, sorted by id (CSR), built on first use
    bool adj_built = false;
    DBuf adj_offs, adj_count, adj;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // second stream of the BVH build (small subtrees beside the large levels),
    // made on first use, default priority; it only ever waits on events of
    // `stream`, and `stream` waits on it before the build returns
    hipStream_t aux_stream = nullptr;
    std::vector<std::unique_ptr<MeshG>> meshes;
    int32_t scene_mesh = -1;  // the (first) instance's mesh of the current TLAS
    std::vector<vira_gpu_instance> tlas;
    // TLAS topology over the instances' placed boxes: (instance or -1, skip) per
    // node in depth-first order, on the device; boxes are made per render
    int32_t tlas_nodes = 0;
    DBuf tlas_topo;
    bool instanced = false;   // TLAS other than one identity instance: rendered camera-relative
    std::vector<vira_gpu_light> lights;
    // Scene::getAmbient / hasAmbient and the background image (getBackgroundEmission)
    float env_ambient[VIRA_GPU_NUM_BINS] = {};
    bool env_has_ambient = false;
    const float* env_bg = nullptr;
    int env_bw = 0, env_bh = 0;
    // render scratch, grown on demand
    DBuf r_ray_o, r_ray_d, r_hit, r_queue[2], r_beta, r_path_rad, r_path_dir, r_srays_o, r_srays_d,
        r_scontrib, r_svis, r_sinfo, r_acc, r_counters, r_lights, r_inst, r_inst_m, r_light_pos, r_ext, r_tnodes;
    // timing events of vira_gpu_render_noalloc, created with the context
    hipEvent_t na_ev[2] = {};
    // bumped by every call that changes geometry, BVHs, lights or materials
    uint64_t scene_gen = 1;
    // fixed-visibility direct illumination: hit records of the last forward pass
    DBuf di_rec, di_gpos, di_gnrm, di_gs;
    bool di_rec_valid = false;
    uint64_t di_gen = 0;
    vira_gpu_camera di_cam{};
    int32_t di_spp = 0;
    uint32_t di_seed = 0;
    // batched rendering: descriptors, queue state and the captured launch
    DBuf b_reqs, b_firsts, b_state, b_acc;
    hipGraphExec_t b_exec = nullptr;
    std::vector<uint64_t> b_key;
    bool b_tuned = false;
    vira_gpu_batch_config b_cfg{4, 1, 16, 1, 0};
    // unresolved sources: per-source records, per-tile counts/offsets/fill, tile lists
    DBuf st_rec, st_count, st_start, st_fill, st_list;
    // denoiser work images: two per pass (direct, indirect) in ping-pong
    DBuf dn_work[4];
    // scan scratch (block sums) and its pinned total; BVH build: each level's
    // node count and the level loop's counters, read back in pinned memory
    DBuf scan_sums;
    PinnedBuf scan_total, bvh_counts;
    // binned-SAH build scratch (about 190 B per triangle), kept for the next build
    DBuf bvh_arena;
    // quipu reads: page-locked staging of a level's decoded heights and albedo images
    PinnedBuf qld_heights, qld_albedo;
    // pixel solid angles of the last camera geometry (width, height, z sign, K^-1)
    DBuf rp_omega;
    std::vector<float> rp_key;
    // FFT convolution: twiddle tables, row/column spectra, the kernel's spectrum,
    // the kernel values (on the device) and padded sizes it was computed for,
    // and the device flag saying whether it was recomputed
    DBuf fc_tw, fc_spec, fc_kspec, fc_kept, fc_flag;
    int fc_K = 0, fc_logLx = 0, fc_logLy = 0;
    bool fc_kvalid = false, fc_attrs = false;
    ~Ctx() {
        if (b_exec) (void)hipGraphExecDestroy(b_exec);
        for (hipEvent_t& e : na_ev)
            if (e) (void)hipEventDestroy(e);
    }

    MeshG& mesh(int32_t id) {
        VG_CHECK(id >= 0 && (size_t)id < meshes.size() && meshes[id] && meshes[id]->alive, VIRA_GPU_ERR_INVALID,
                 "invalid mesh id " + std::to_string(id));
        return *meshes[id];
    }
};

// Launch helpers ------------------------------------------------------------
inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// Kernel entry points implemented in the .hip translation units.
void launch_synth_fbm(hipStream_t s, int W, int H, uint32_t seed, double amp, int oct, double base, float* out);
void launch_synth_albedo(hipStream_t s, size_t nv, uint32_t seed, float* out);
void dem_build_mesh(Ctx& c, MeshG& m, const float* heights, int W, int H, const vira_gpu_projection& p,
                    const float* grey);
// The DEM grid's triangles (imageToIndexBuffer order, holes skipped), face
// normals and vertex normals around vertices that write_vertices fills
// (GVertex[W*H], and the FP64 positions, 3 per vertex, +inf x for holes;
// normals are overwritten).
void dem_grid_mesh(Ctx& c, MeshG& m, const float* heights, int W, int H,
                   const std::function<void(GVertex*, double*)>& write_vertices);
void mesh_compute_normals(Ctx& c, MeshG& m);
void mesh_adjacency(Ctx& c, MeshG& m);
void dem_update(Ctx& c, MeshG& m, const float* heights);
void di_render(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_di_opts& o, float* radiance,
               vira_gpu_di_stats* st);
void di_adjoint(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_di_opts& o, const float* grad_image,
                float* grad_heights, float* grad_texture, vira_gpu_di_stats* st);
void received_power(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_sensor& s, const float* radiance,
                    float* power, float* omega);
void render_stars(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_sensor& sen, const vira_gpu_star_psf& psf,
                  const float* dirs, const float* irr, size_t n, float* img);
void psf_kernel(Ctx& c, const vira_gpu_psf& P, int K, float* out);
void star_light(Ctx& c, const vira_gpu_star* stars, size_t n, double et, float* dirs, float* irr);
void qsc_read(const char* path, vira_gpu_star* out, size_t capacity, size_t* count);
vira_gpu_qld* qld_open(const char* path, vira_gpu_qld_info* info);
int qld_select(const vira_gpu_qld& q, double required_gsd);
void qld_read(Ctx& c, const vira_gpu_qld& q, int level, const vira_gpu_qld_read_opts& o, float* heights_dev,
              float* albedo_dev, int32_t* mesh_id_out, vira_gpu_qld_level* info,
              const std::function<int32_t()>& new_mesh);
void denoise(Ctx& c, const vira_gpu_denoise_opts& o, int W, int H, float* direct, float* indirect,
             const float* albedo, const float* depth, const float* normal, float* total);
void convolve(Ctx& c, const float* in, int W, int H, const float* kern, int K, float* out);
// FFT correlation for kernels of 256 taps or more; fits = both padded sizes <= 8192
bool fft_convolve_fits(int W, int H, int K);
void fft_convolve(Ctx& c, const float* in, int W, int H, const float* kern, int K, float* out);
void simulate_sensor(Ctx& c, const vira_gpu_sensor_sim& cfg, const float* power, int W, int H, float* photons,
                     float* image);
void batch_submit(Ctx& c, const vira_gpu_request* reqs, size_t n, const vira_gpu_batch_config& cfg, int stride,
                  vira_gpu_batch_stats* st);
void batch_autotune(Ctx& c, const vira_gpu_request* reqs, size_t n, int stride, vira_gpu_batch_config* best,
                    double* best_ms);
void blas_build(Ctx& c, MeshG& m, const vira_gpu_bvh_opts& o);
// New boxes for a built BLAS after its mesh's vertices moved (same triangles);
// 4-wide BLASes are refitted in place, 2-wide ones rebuilt.
void blas_refit(Ctx& c, MeshG& m);
void trace_rays(Ctx& c, MeshG& m, const float* rays, size_t n, int any_hit, float* t, uint32_t* id, float* b1,
                float* b2, double* ms);
void render(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_render_opts& o, const vira_gpu_outputs& out,
            vira_gpu_render_stats* st);
// vira_gpu_render_noalloc: the per-pixel render with every scratch buffer in
// the library's fixed per-device array; noalloc_init finds that array's
// device address once per device (with the device current).
void render_noalloc(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_render_opts& o, const vira_gpu_outputs& out,
                    vira_gpu_render_stats* st);
void noalloc_init(int device);
size_t noalloc_max_tiles();

// Exclusive scan of n uint32 values (in place allowed); returns the total after
// a stream synchronisation, or 0 without waiting when total_wanted is false.
uint32_t scan_exclusive_u32(Ctx& c, const uint32_t* in, uint32_t* out, size_t n, bool total_wanted = true);
// Stream compaction: the values of in[0..n) other than 0xffffffff, in order, to
// out (not in place); returns their count after a stream synchronisation.
uint32_t compact_u32(Ctx& c, const uint32_t* in, uint32_t* out, size_t n);
// The same with the element count in device memory (*dn <= n_max) and the total
// written to device memory; sums: scratch of sums_words >= scan_sums_words(n_max)
// words (checked).
size_t scan_sums_words(size_t n_max);
void scan_exclusive_u32_dev(Ctx& c, const uint32_t* in, uint32_t* out, const uint32_t* dn, size_t n_max,
                            uint32_t* sums, size_t sums_words, uint32_t* total);
// Two such scans over the same device count in one set of launches; sums:
// sums_words >= 2 * scan_sums_words(n_max) (checked).
void scan_exclusive_u32x2_dev(Ctx& c, const uint32_t* in0, uint32_t* out0, uint32_t* total0, const uint32_t* in1,
                              uint32_t* out1, uint32_t* total1, const uint32_t* dn, size_t n_max, uint32_t* sums,
                              size_t sums_words);

}  // namespace vg

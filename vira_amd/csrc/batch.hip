// batch.hip -- many render requests in one persistent, hipGraph-captured launch.
//
// Each request is CPUPathTracer::render (cpu_path_tracer.ipp:29-161) of one
// mesh with one camera, options and up to VIRA_GPU_MAX_LIGHTS lights, in the
// context's light environment (ambient light, background).  The work unit is the per-pixel
// renderer's 8x8 tile (fused.hpp: render_tile), so a request's passes are the
// ones the per-request launch writes, bit for bit.  All requests' tiles form
// one queue, ordered by request (longest rays-per-tile first by default); a
// fixed grid of persistent waves takes tiles from it with one atomic per grab
// until it is empty, so small images and high-spp requests share the GPU
// instead of each leaving most compute units idle at its tail.
//
// The queue counter and ray counters are cleared by a kernel inside the
// captured graph (no memset nodes); the descriptors are copied in before the
// graph launch, so an identical batch layout replays the same graph.
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <numeric>

#include "context.hpp"
#include "fused.hpp"
#include "scene_params.hpp"

namespace vg {
namespace {

struct BatchReq {
    TileJob job;
    JobExt ext;
    LightF lights[VIRA_GPU_MAX_LIGHTS];
    uint32_t first;        // first queue item of this request
    uint32_t tile_stride;  // queue item i of the request renders tile i * tile_stride
};

// state[0] = queue head, state[2..7] = ray counters (64-bit)
__global__ void k_batch_reset(unsigned long long* state) {
    if (threadIdx.x < 8) state[threadIdx.x] = 0ull;
}

template <int W, int WPB, bool MULTI, bool FEAT>
__global__ void __launch_bounds__(kFusedBS* WPB, kFusedWavesPerSimd) k_render_batch(const BatchReq* __restrict__ reqs, int nreq,
                                                                uint32_t nitems, uint32_t grab,
                                                                float* __restrict__ acc,
                                                                unsigned long long* __restrict__ state) {
    constexpr int kL = fused_lds_floats<MULTI>();
    __shared__ uint32_t s_ref[WPB * kStackLds * kFusedBS];
    __shared__ float s_dist[WPB * kStackLds * kFusedBS];
    __shared__ float s_contrib[WPB * kL * kFusedBS];
    const uint32_t wave = threadIdx.x / kFusedBS, lane = threadIdx.x % kFusedBS;
    const uint32_t gw = blockIdx.x * WPB + wave;
    float* __restrict__ a = acc + (size_t)gw * (A_FIELDS * 64);
    uint32_t* s_ref_w = s_ref + wave * (kStackLds * kFusedBS);
    float* s_dist_w = s_dist + wave * (kStackLds * kFusedBS);
    float* s_contrib_w = s_contrib + wave * (kL * kFusedBS);
    uint32_t n_prim = 0, n_shadow = 0, n_bounce = 0, n_hit = 0;
    uint32_t* head = reinterpret_cast<uint32_t*>(state);
    while (true) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(head, grab);
        base = __builtin_amdgcn_readfirstlane(__shfl(base, 0));
        if (base >= nitems) break;  // every wave reaches this once the queue is drained
        const uint32_t end = min(base + grab, nitems);
        for (uint32_t k = base; k < end; ++k) {
            // the request owning item k: the last one whose first item is <= k
            int lo = 0, hi = nreq - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (reqs[mid].first <= k) lo = mid;
                else hi = mid - 1;
            }
            const BatchReq& R = reqs[lo];
            const uint32_t tile = (k - R.first) * R.tile_stride;
            render_tile<W, false, MULTI, FEAT>(R.job, tile, a, s_ref_w, s_dist_w, s_contrib_w, lane, n_prim, n_shadow,
                                         n_bounce, n_hit);
        }
    }
    wave_add(&state[2], n_prim);
    wave_add(&state[3], n_shadow);
    wave_add(&state[4], n_bounce);
    wave_add(&state[5], n_hit);
}

using BatchKernel = void (*)(const BatchReq*, int, uint32_t, uint32_t, float*, unsigned long long*);

template <bool MULTI, bool FEAT>
BatchKernel pick_kernel_t(int width, int wpb) {
    if (width == 4) {
        if (wpb == 1) return k_render_batch<4, 1, MULTI, FEAT>;
        if (wpb == 2) return k_render_batch<4, 2, MULTI, FEAT>;
        return k_render_batch<4, 4, MULTI, FEAT>;
    }
    if (wpb == 1) return k_render_batch<2, 1, MULTI, FEAT>;
    if (wpb == 2) return k_render_batch<2, 2, MULTI, FEAT>;
    return k_render_batch<2, 4, MULTI, FEAT>;
}
// Several lights or ambient light in any request: the variant that keeps the
// vertex in LDS (it gives single-light requests the same numbers); a shown
// background, adaptive sampling or a normal/spectral map in any request: the
// feature variant (likewise the same numbers for the others).
BatchKernel pick_kernel(int width, int wpb, bool multi, bool feat) {
    if (multi) return feat ? pick_kernel_t<true, true>(width, wpb) : pick_kernel_t<true, false>(width, wpb);
    return feat ? pick_kernel_t<false, true>(width, wpb) : pick_kernel_t<false, false>(width, wpb);
}

void check_config(const vira_gpu_batch_config& g) {
    VG_CHECK(g.bvh_width == 2 || g.bvh_width == 4, VIRA_GPU_ERR_INVALID, "batch: bvh_width must be 2 or 4");
    VG_CHECK(g.waves_per_block == 1 || g.waves_per_block == 2 || g.waves_per_block == 4, VIRA_GPU_ERR_INVALID,
             "batch: waves_per_block must be 1, 2 or 4");
    VG_CHECK(g.waves_per_cu >= 1 && g.waves_per_cu <= 32, VIRA_GPU_ERR_INVALID, "batch: waves_per_cu 1..32");
    VG_CHECK(g.tiles_per_grab >= 1 && g.tiles_per_grab <= 16, VIRA_GPU_ERR_INVALID, "batch: tiles_per_grab 1..16");
    VG_CHECK(g.order == 0 || g.order == 1, VIRA_GPU_ERR_INVALID, "batch: order must be 0 or 1");
}

int num_cus(Ctx& c) {
    int n = 0;
    VG_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, c.device));
    return std::max(1, n);
}

}  // namespace

void batch_submit(Ctx& c, const vira_gpu_request* reqs, size_t n, const vira_gpu_batch_config& cfg, int stride,
                  vira_gpu_batch_stats* st) {
    check_config(cfg);
    VG_CHECK(n >= 1 && n <= (1u << 20), VIRA_GPU_ERR_INVALID, "batch: 1..2^20 requests");
    VG_CHECK(stride >= 1, VIRA_GPU_ERR_INVALID, "batch: sample stride must be positive");
    // descriptors, in queue order
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::vector<double> cost(n);
    bool multi = c.env_has_ambient, feat = false;
    for (size_t i = 0; i < n; ++i) {
        const vira_gpu_request& r = reqs[i];
        MeshG& m = c.mesh(r.mesh_id);
        VG_CHECK(m.blas.built, VIRA_GPU_ERR_STATE, "batch: request " + std::to_string(i) + ": BLAS not built");
        VG_CHECK(m.blas.width == cfg.bvh_width, VIRA_GPU_ERR_STATE,
                 "batch: request " + std::to_string(i) + ": mesh BVH width " + std::to_string(m.blas.width) +
                     " differs from the configured " + std::to_string(cfg.bvh_width));
        VG_CHECK(r.camera.width > 0 && r.camera.height > 0 &&
                     (size_t)r.camera.width * r.camera.height < (1ull << 30),
                 VIRA_GPU_ERR_INVALID, "batch: request " + std::to_string(i) + ": bad image size");
        VG_CHECK(r.opts.samples > 0 && r.opts.bounces >= 0 && r.opts.bounces < 64, VIRA_GPU_ERR_INVALID,
                 "batch: request " + std::to_string(i) + ": bad samples/bounces");
        VG_CHECK(!r.opts.count_visits, VIRA_GPU_ERR_UNSUPPORTED,
                 "batch: request " + std::to_string(i) + ": visit counting runs in vira_gpu_render only");
        VG_CHECK(r.num_lights >= 0 && r.num_lights <= VIRA_GPU_MAX_LIGHTS, VIRA_GPU_ERR_INVALID,
                 "batch: request " + std::to_string(i) + ": 0.." + std::to_string(VIRA_GPU_MAX_LIGHTS) + " lights");
        for (int l = 0; l < r.num_lights; ++l) {
            VG_CHECK(r.lights[l].kind == 0 || r.lights[l].kind == 1, VIRA_GPU_ERR_INVALID,
                     "batch: request " + std::to_string(i) + ": unknown light kind");
            VG_CHECK(r.lights[l].kind == 0 || r.lights[l].radius > 0, VIRA_GPU_ERR_INVALID,
                     "batch: request " + std::to_string(i) + ": SphereLight radius must be positive");
        }
        if (r.num_lights > 1) multi = true;
        if ((r.opts.show_background && c.env_bg) || r.opts.adaptive_sampling || material_feat(m.mat)) feat = true;
        cost[i] = (double)r.opts.samples * (1.0 + r.opts.bounces);
    }
    if (cfg.order == 0)
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
    std::vector<BatchReq> hr(n);
    uint64_t items = 0, tiles_total = 0;
    const BatchReq* dev_base = nullptr;
    c.b_reqs.alloc(n * sizeof(BatchReq));
    dev_base = c.b_reqs.as<BatchReq>();
    size_t max_tiles = 0;
    for (size_t qi = 0; qi < n; ++qi) {
        const vira_gpu_request& r = reqs[order[qi]];
        MeshG& m = c.mesh(r.mesh_id);
        const TileGrid g(r.camera);
        BatchReq& d = hr[qi];
        for (int l = 0; l < VIRA_GPU_MAX_LIGHTS; ++l) d.lights[l] = l < r.num_lights ? light_params(r.lights[l]) : LightF{};
        d.job.nodes = m.blas.nodes.p;
        d.job.tris = m.blas.tris.as<float4>();
        d.job.verts = m.verts.as<GVertex>();
        d.job.idx = m.idx.as<uint32_t>();
        d.job.mat = mat_params(m);
        // the lights live in this descriptor's device copy
        d.job.lights = Lights{r.num_lights, reinterpret_cast<const LightF*>(reinterpret_cast<const char*>(dev_base + qi) +
                                                                            offsetof(BatchReq, lights))};
        d.job.cam = cam_params(r.camera);
        d.job.seed = r.opts.seed;
        d.job.spp = r.opts.samples;
        d.job.max_bounces = r.opts.bounces;
        d.job.rr = r.opts.russian_roulette;
        d.job.rr_start = r.opts.rr_start;
        d.job.tiles_x = g.tiles_x;
        d.job.inst = nullptr;
        d.job.tlas = nullptr;
        d.job.ntlas = 0;
        d.ext.out = out_params(r.out);
        d.ext.env = env_params(c, r.opts);
        d.ext.adapt = adaptive_params(r.opts);
        d.ext.pass = pass_params(r.camera, r.mesh_id, m);
        d.job.ext = reinterpret_cast<const JobExt*>(reinterpret_cast<const char*>(dev_base + qi) +
                                                    offsetof(BatchReq, ext));
        d.first = (uint32_t)items;
        d.tile_stride = (uint32_t)stride;
        const uint64_t nt = (g.tiles + (uint32_t)stride - 1) / (uint32_t)stride;
        items += nt;
        tiles_total += nt;
        max_tiles = std::max<size_t>(max_tiles, g.tiles);
        VG_CHECK(items < (1ull << 32), VIRA_GPU_ERR_INVALID, "batch: more than 2^32 tiles");
    }
    const int cus = num_cus(c);
    const int wpb = cfg.waves_per_block;
    const uint32_t blocks = (uint32_t)std::max(1, cus * cfg.waves_per_cu / wpb);
    const size_t waves = (size_t)blocks * wpb;
    c.b_acc.alloc(waves * A_FIELDS * 64 * sizeof(float));
    c.b_state.alloc(64 * 8);
    VG_HIP(hipMemcpyAsync(c.b_reqs.p, hr.data(), n * sizeof(BatchReq), hipMemcpyHostToDevice, c.stream));

    const std::vector<uint64_t> key = {(uint64_t)(uintptr_t)c.b_reqs.p, (uint64_t)(uintptr_t)c.b_acc.p,
                                       (uint64_t)(uintptr_t)c.b_state.p, (uint64_t)n, items,
                                       (uint64_t)cfg.bvh_width, (uint64_t)cfg.waves_per_block,
                                       (uint64_t)cfg.waves_per_cu, (uint64_t)cfg.tiles_per_grab, (uint64_t)blocks,
                                       (uint64_t)multi, (uint64_t)feat};
    bool reused = c.b_exec && c.b_key == key;
    if (!reused) {
        if (c.b_exec) {
            (void)hipGraphExecDestroy(c.b_exec);
            c.b_exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        VG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        hipError_t e1 = hipSuccess, e2 = hipSuccess;
        k_batch_reset<<<1, 64, 0, c.stream>>>(c.b_state.as<unsigned long long>());
        e1 = hipGetLastError();
        BatchKernel kern = pick_kernel(cfg.bvh_width, wpb, multi, feat);
        kern<<<blocks, kFusedBS * wpb, 0, c.stream>>>(c.b_reqs.as<BatchReq>(), (int)n, (uint32_t)items,
                                                      (uint32_t)cfg.tiles_per_grab, c.b_acc.as<float>(),
                                                      c.b_state.as<unsigned long long>());
        e2 = hipGetLastError();
        hipError_t ec = hipStreamEndCapture(c.stream, &graph);
        VG_HIP(e1);
        VG_HIP(e2);
        VG_HIP(ec);
        hipError_t ei = hipGraphInstantiate(&c.b_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        VG_HIP(ei);
        c.b_key = key;
    }
    EventPair ev;
    VG_HIP(hipEventRecord(ev.a, c.stream));
    VG_HIP(hipGraphLaunch(c.b_exec, c.stream));
    VG_HIP(hipEventRecord(ev.b, c.stream));
    unsigned long long hc[8];
    VG_HIP(hipMemcpyAsync(hc, c.b_state.p, sizeof(hc), hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    float ms = 0.f;
    VG_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
    if (st) {
        *st = vira_gpu_batch_stats{};
        st->primary_rays = hc[2];
        st->shadow_rays = hc[3];
        st->bounce_rays = hc[4];
        st->closest_hits = hc[5];
        st->tiles = tiles_total;
        st->ms = ms;
        st->graph_reused = reused ? 1 : 0;
    }
}

void batch_autotune(Ctx& c, const vira_gpu_request* reqs, size_t n, int stride, vira_gpu_batch_config* best,
                    double* best_ms) {
    VG_CHECK(n >= 1, VIRA_GPU_ERR_INVALID, "batch_autotune: no requests");
    VG_CHECK(stride >= 1, VIRA_GPU_ERR_INVALID, "batch_autotune: sample stride must be positive");
    std::vector<int32_t> meshes;
    for (size_t i = 0; i < n; ++i) {
        MeshG& m = c.mesh(reqs[i].mesh_id);
        VG_CHECK(m.blas.built, VIRA_GPU_ERR_STATE, "batch_autotune: BLAS not built");
        if (std::find(meshes.begin(), meshes.end(), reqs[i].mesh_id) == meshes.end())
            meshes.push_back(reqs[i].mesh_id);
    }
    auto set_width = [&](int w) {
        for (int32_t id : meshes) {
            MeshG& m = c.mesh(id);
            if (m.blas.width != w) {
                vira_gpu_bvh_opts o = m.blas.opts;
                o.width = w;
                blas_build(c, m, o);
                m.blas.opts = o;
                c.scene_gen++;
            }
        }
    };
    // Stage 1: every candidate on the sampled tiles (every stride-th tile of
    // each request).  A sample changes the queue's tail, so stage 2 re-times
    // the three fastest on the whole workload and keeps the best of those.
    const int w0 = c.mesh(meshes[0]).blas.width;
    struct Cand {
        vira_gpu_batch_config g;
        double ms;
    };
    std::vector<Cand> cands;
    for (int w : {w0, w0 == 4 ? 2 : 4}) {
        set_width(w);
        for (int wpb : {1, 2})
            for (int wpcu : {8, 12, 16})
                for (int grab : {1, 4}) {
                    vira_gpu_batch_config g{w, wpb, wpcu, grab, 0};
                    vira_gpu_batch_stats s{};
                    batch_submit(c, reqs, n, g, stride, &s);  // capture + warm
                    batch_submit(c, reqs, n, g, stride, &s);
                    double ms = s.ms;
                    batch_submit(c, reqs, n, g, stride, &s);
                    cands.push_back({g, std::min(ms, s.ms)});
                }
    }
    std::stable_sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.ms < b.ms; });
    const size_t top = std::min<size_t>(stride > 1 ? 3 : 1, cands.size());
    // the BVH width in place now is the last tried: visit the finalists of that width first
    std::stable_sort(cands.begin(), cands.begin() + top, [&](const Cand& a, const Cand& b) {
        return (a.g.bvh_width == c.mesh(meshes[0]).blas.width) > (b.g.bvh_width == c.mesh(meshes[0]).blas.width);
    });
    vira_gpu_batch_config bc = cands[0].g;
    double bms = cands[0].ms;
    if (stride > 1) {
        bms = 1e300;
        for (size_t i = 0; i < top; ++i) {
            set_width(cands[i].g.bvh_width);
            vira_gpu_batch_stats s{};
            batch_submit(c, reqs, n, cands[i].g, 1, &s);  // capture + warm
            batch_submit(c, reqs, n, cands[i].g, 1, &s);
            if (s.ms < bms) {
                bms = s.ms;
                bc = cands[i].g;
            }
        }
    }
    set_width(bc.bvh_width);
    c.b_cfg = bc;
    c.b_tuned = true;
    if (best) *best = bc;
    if (best_ms) *best_ms = bms;
}

}  // namespace vg

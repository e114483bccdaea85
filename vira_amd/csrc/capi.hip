// capi.hip -- the extern "C" entry points of include/vira_gpu.h.
// Every entry point catches library errors and returns a status; the message
// is kept per thread for vira_gpu_last_error().
#include <algorithm>
#include <cmath>
#include <functional>
#include <cstring>
#include <string>

#include "context.hpp"
#include "lz4_block.hpp"
#include "instance.hpp"

namespace vg {
void mesh_upload_pack(Ctx& c, MeshG& m, const float* pos, const float* alb, const float* uv, const float* nrm);
}

namespace {
thread_local std::string g_last_error;

template <class F>
vira_gpu_status guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return VIRA_GPU_OK;
    } catch (const vg::Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return VIRA_GPU_ERR_OOM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return VIRA_GPU_ERR_HIP;
    }
}

vg::Ctx& C(vira_gpu_ctx* c) {
    VG_CHECK(c != nullptr, VIRA_GPU_ERR_INVALID, "null context");
    return *reinterpret_cast<vg::Ctx*>(c);
}

int32_t new_mesh(vg::Ctx& c) {
    c.meshes.push_back(std::make_unique<vg::MeshG>());
    c.meshes.back()->alive = true;
    vira_gpu_material& m = c.meshes.back()->mat;
    for (int k = 0; k < VIRA_GPU_NUM_BINS; ++k) m.albedo[k] = 1.0f;
    m.texture_dev = nullptr;
    m.tex_width = m.tex_height = 0;
    m.smooth_shading = 0;
    m.tex_channels = 1;
    m.normal_map_dev = nullptr;
    m.nmap_width = m.nmap_height = 0;
    m.material_id = 0;
    return (int32_t)c.meshes.size() - 1;
}

void set_device(vg::Ctx& c) { VG_HIP(hipSetDevice(c.device)); }
}  // namespace

extern "C" {

const char* vira_gpu_last_error(void) { return g_last_error.c_str(); }
const char* vira_gpu_version(void) { return "vira_gpu 0.1 (gfx950)"; }

vira_gpu_status vira_gpu_create(vira_gpu_ctx** out, const vira_gpu_opts* opts) {
    return guarded([&] {
        VG_CHECK(out != nullptr, VIRA_GPU_ERR_INVALID, "null output pointer");
        auto c = std::make_unique<vg::Ctx>();
        c->device = opts ? opts->device : 0;
        int n = 0;
        VG_HIP(hipGetDeviceCount(&n));
        VG_CHECK(c->device >= 0 && c->device < n, VIRA_GPU_ERR_INVALID, "invalid device ordinal");
        VG_HIP(hipSetDevice(c->device));
        // the allocation-free render's scratch address on this device and its timing events
        vg::noalloc_init(c->device);
        VG_HIP(hipEventCreate(&c->na_ev[0]));
        VG_HIP(hipEventCreate(&c->na_ev[1]));
        if (opts && opts->stream) {
            c->stream = static_cast<hipStream_t>(opts->stream);
        } else {
            VG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            c->own_stream = true;
        }
        *out = reinterpret_cast<vira_gpu_ctx*>(c.release());
    });
}

vira_gpu_status vira_gpu_destroy(vira_gpu_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        vg::Ctx* c = reinterpret_cast<vg::Ctx*>(ctx);
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        if (c->aux_stream) {
            (void)hipStreamSynchronize(c->aux_stream);
            (void)hipStreamDestroy(c->aux_stream);
        }
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
    });
}

vira_gpu_status vira_gpu_set_stream(vira_gpu_ctx* ctx, void* stream) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_HIP(hipStreamSynchronize(c.stream));
        if (c.own_stream) VG_HIP(hipStreamDestroy(c.stream));
        if (stream) {
            c.stream = static_cast<hipStream_t>(stream);
            c.own_stream = false;
        } else {
            VG_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
            c.own_stream = true;
        }
    });
}

vira_gpu_status vira_gpu_synchronize(vira_gpu_ctx* ctx) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_HIP(hipStreamSynchronize(c.stream));
    });
}

vira_gpu_status vira_gpu_synth_fbm(vira_gpu_ctx* ctx, int width, int height, uint32_t seed, double amplitude,
                                   int octaves, double base_cells, float* heights_dev) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(width > 0 && height > 0 && heights_dev, VIRA_GPU_ERR_INVALID, "synth_fbm: bad arguments");
        VG_CHECK(octaves >= 1 && octaves <= 30 && base_cells > 0, VIRA_GPU_ERR_INVALID, "synth_fbm: bad octaves/base");
        vg::launch_synth_fbm(c.stream, width, height, seed, amplitude, octaves, base_cells, heights_dev);
        VG_HIP(hipStreamSynchronize(c.stream));
    });
}

vira_gpu_status vira_gpu_synth_vertex_albedo(vira_gpu_ctx* ctx, size_t nv, uint32_t seed, float* albedo_dev) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(albedo_dev || nv == 0, VIRA_GPU_ERR_INVALID, "synth_vertex_albedo: null output");
        if (nv == 0) return;
        vg::launch_synth_albedo(c.stream, nv, seed, albedo_dev);
        VG_HIP(hipStreamSynchronize(c.stream));
    });
}

vira_gpu_status vira_gpu_dem_mesh(vira_gpu_ctx* ctx, const float* heights_dev, int width, int height,
                                  const vira_gpu_projection* proj, const float* grey_albedo_dev, int32_t* mesh_id) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(heights_dev && proj && mesh_id, VIRA_GPU_ERR_INVALID, "dem_mesh: null argument");
        VG_CHECK(proj->kind == 0 || proj->kind == 1, VIRA_GPU_ERR_INVALID, "dem_mesh: unknown projection kind");
        VG_CHECK(proj->spacing > 0 && (proj->kind == 0 || proj->radius > 0), VIRA_GPU_ERR_INVALID,
                 "dem_mesh: spacing/radius must be positive");
        int32_t id = new_mesh(c);
        try {
            vg::dem_build_mesh(c, *c.meshes[id], heights_dev, width, height, *proj, grey_albedo_dev);
            c.scene_gen++;
        } catch (...) {
            c.meshes[id].reset();
            throw;
        }
        *mesh_id = id;
    });
}

vira_gpu_status vira_gpu_mesh_upload(vira_gpu_ctx* ctx, const float* pos, size_t nv, const uint32_t* idx, size_t nt,
                                     const float* alb, const float* uv, const float* nrm, int32_t* mesh_id) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(mesh_id && (pos || nv == 0) && (idx || nt == 0), VIRA_GPU_ERR_INVALID, "mesh_upload: null argument");
        VG_CHECK(nv < (1ull << 31) && nt < (1ull << 27), VIRA_GPU_ERR_INVALID, "mesh_upload: mesh too large");
        for (size_t i = 0; i < 3 * nt; ++i)
            VG_CHECK(idx[i] < nv, VIRA_GPU_ERR_INVALID, "mesh_upload: index out of range");
        int32_t id = new_mesh(c);
        vg::MeshG& m = *c.meshes[id];
        try {
            m.nv = nv;
            m.nt = nt;
            m.verts.alloc(std::max<size_t>(1, nv) * sizeof(vg::GVertex));
            m.idx.alloc(std::max<size_t>(1, nt) * 12);
            vg::DBuf dpos, dalb, duv, dnrm;
            dpos.alloc(std::max<size_t>(1, nv) * 12);
            if (nv) VG_HIP(hipMemcpyAsync(dpos.p, pos, nv * 12, hipMemcpyHostToDevice, c.stream));
            if (alb) { dalb.alloc(nv * 32); VG_HIP(hipMemcpyAsync(dalb.p, alb, nv * 32, hipMemcpyHostToDevice, c.stream)); }
            if (uv) { duv.alloc(nv * 8); VG_HIP(hipMemcpyAsync(duv.p, uv, nv * 8, hipMemcpyHostToDevice, c.stream)); }
            if (nrm) { dnrm.alloc(nv * 12); VG_HIP(hipMemcpyAsync(dnrm.p, nrm, nv * 12, hipMemcpyHostToDevice, c.stream)); }
            if (nt) VG_HIP(hipMemcpyAsync(m.idx.p, idx, nt * 12, hipMemcpyHostToDevice, c.stream));
            if (nv) vg::mesh_upload_pack(c, m, dpos.as<float>(), dalb.as<float>(), duv.as<float>(), dnrm.as<float>());
            VG_HIP(hipStreamSynchronize(c.stream));
            // Mesh::init computes normals when the first vertex normal is zero (mesh.ipp:62-64)
            bool need = !nrm || (nv > 0 && nrm[0] == 0.f && nrm[1] == 0.f && nrm[2] == 0.f);
            if (need) vg::mesh_compute_normals(c, m);
            c.scene_gen++;
        } catch (...) {
            c.meshes[id].reset();
            throw;
        }
        *mesh_id = id;
    });
}

vira_gpu_status vira_gpu_mesh_info(vira_gpu_ctx* ctx, int32_t mesh_id, size_t* nv, size_t* nt) {
    return guarded([&] {
        vg::MeshG& m = C(ctx).mesh(mesh_id);
        if (nv) *nv = m.nv;
        if (nt) *nt = m.nt;
    });
}

vira_gpu_status vira_gpu_mesh_download(vira_gpu_ctx* ctx, int32_t mesh_id, float* pos, uint32_t* idx, float* nrm) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        vg::MeshG& m = c.mesh(mesh_id);
        VG_HIP(hipStreamSynchronize(c.stream));
        if (idx && m.nt) VG_HIP(hipMemcpy(idx, m.idx.p, m.nt * 12, hipMemcpyDeviceToHost));
        if ((pos || nrm) && m.nv) {
            std::vector<vg::GVertex> v(m.nv);
            VG_HIP(hipMemcpy(v.data(), m.verts.p, m.nv * sizeof(vg::GVertex), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < m.nv; ++i) {
                if (pos) { pos[3 * i] = v[i].px; pos[3 * i + 1] = v[i].py; pos[3 * i + 2] = v[i].pz; }
                if (nrm) { nrm[3 * i] = v[i].nx; nrm[3 * i + 1] = v[i].ny; nrm[3 * i + 2] = v[i].nz; }
            }
        }
    });
}

vira_gpu_status vira_gpu_mesh_release(vira_gpu_ctx* ctx, int32_t mesh_id) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        c.mesh(mesh_id);
        VG_HIP(hipStreamSynchronize(c.stream));
        c.meshes[mesh_id].reset();
        if (c.scene_mesh == mesh_id) c.scene_mesh = -1;
        c.scene_gen++;
    });
}

vira_gpu_status vira_gpu_blas_build(vira_gpu_ctx* ctx, int32_t mesh_id, const vira_gpu_bvh_opts* opts) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        vira_gpu_bvh_opts o{4, 16, 4};
        if (opts) o = *opts;
        if (o.width == 0) o.width = 4;
        VG_CHECK(o.max_leaf >= 0 && o.max_leaf <= 15 && o.bins >= 0 && o.bins <= 32, VIRA_GPU_ERR_INVALID,
                 "blas_build: max_leaf must be 1..15 and bins 4..32");
        VG_CHECK(o.width == 2 || o.width == 4, VIRA_GPU_ERR_INVALID, "blas_build: width must be 2 or 4");
        VG_CHECK(o.traversal_cost >= 0.0f && o.traversal_cost <= 64.0f, VIRA_GPU_ERR_INVALID,
                 "blas_build: traversal_cost must be 0..64");
        VG_CHECK(o.method >= 0 && o.method <= 2, VIRA_GPU_ERR_INVALID, "blas_build: method must be 0, 1 or 2");
        vg::MeshG& m = c.mesh(mesh_id);
        c.scene_gen++;
        vg::blas_build(c, m, o);
        m.blas.opts = o;
    });
}

vira_gpu_status vira_gpu_blas_stats(vira_gpu_ctx* ctx, int32_t mesh_id, vira_gpu_bvh_stats* stats) {
    return guarded([&] {
        vg::MeshG& m = C(ctx).mesh(mesh_id);
        VG_CHECK(stats, VIRA_GPU_ERR_INVALID, "blas_stats: null output");
        VG_CHECK(m.blas.built, VIRA_GPU_ERR_STATE, "blas_stats: BLAS not built");
        *stats = m.blas.stats;
    });
}

vira_gpu_status vira_gpu_tlas_build(vira_gpu_ctx* ctx, const vira_gpu_instance* inst, size_t n) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        VG_CHECK(inst || n == 0, VIRA_GPU_ERR_INVALID, "tlas_build: null instances");
        VG_CHECK(n != 0, VIRA_GPU_ERR_INVALID, "tlas_build: empty TLAS");  // vira_tlas.ipp:33
        VG_CHECK(n <= (size_t)vg::kMaxInstances, VIRA_GPU_ERR_UNSUPPORTED, "tlas_build: at most 65536 instances");
        static const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        int width = 0;
        for (size_t i = 0; i < n; ++i) {
            vg::MeshG& m = c.mesh(inst[i].mesh_id);
            VG_CHECK(m.blas.built, VIRA_GPU_ERR_STATE, "tlas_build: BLAS not built");
            VG_CHECK(width == 0 || m.blas.width == width, VIRA_GPU_ERR_INVALID,
                     "tlas_build: all instances need BLASes of one BVH width");
            width = m.blas.width;
            const double* M = inst[i].transform;
            VG_CHECK(M[3] == 0.0 && M[7] == 0.0 && M[11] == 0.0 && M[15] == 1.0, VIRA_GPU_ERR_INVALID,
                     "tlas_build: model matrices must be affine (last row 0 0 0 1)");
            for (int k = 0; k < 16; ++k)
                VG_CHECK(std::isfinite(M[k]), VIRA_GPU_ERR_INVALID, "tlas_build: non-finite model matrix");
            const double det = M[0] * (M[5] * M[10] - M[9] * M[6]) - M[4] * (M[1] * M[10] - M[9] * M[2]) +
                               M[8] * (M[1] * M[6] - M[5] * M[2]);
            VG_CHECK(det != 0.0 && std::isfinite(det), VIRA_GPU_ERR_INVALID, "tlas_build: singular model matrix");
        }
        c.tlas.assign(inst, inst + n);
        c.instanced = !(n == 1 && std::memcmp(inst[0].transform, I, sizeof(I)) == 0);
        // TLAS topology (EmbreeTLAS::build, embree_tlas.ipp:77-115; ViraTLAS clusters, vira_tlas.ipp:29-79):
        // a binary tree over the instances' placed boxes, split at the median centroid of the longest axis
        std::vector<double> box(6 * n), cen(3 * n);
        for (size_t i = 0; i < n; ++i) {
            const vg::MeshG& m = c.mesh(inst[i].mesh_id);
            const double* M = inst[i].transform;
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            if (m.blas.lo[0] <= m.blas.hi[0]) {
                for (int q = 0; q < 8; ++q) {
                    const double p[3] = {(q & 1) ? m.blas.hi[0] : m.blas.lo[0], (q & 2) ? m.blas.hi[1] : m.blas.lo[1],
                                         (q & 4) ? m.blas.hi[2] : m.blas.lo[2]};
                    for (int r = 0; r < 3; ++r) {
                        const double w = M[r] * p[0] + M[4 + r] * p[1] + M[8 + r] * p[2] + M[12 + r];
                        lo[r] = std::min(lo[r], w);
                        hi[r] = std::max(hi[r], w);
                    }
                }
            } else {
                for (int r = 0; r < 3; ++r) lo[r] = hi[r] = M[12 + r];
            }
            for (int r = 0; r < 3; ++r) {
                box[6 * i + r] = lo[r];
                box[6 * i + 3 + r] = hi[r];
                cen[3 * i + r] = 0.5 * (lo[r] + hi[r]);
            }
        }
        std::vector<uint32_t> ids(n);
        for (size_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
        std::vector<int32_t> topo;  // (inst, skip) pairs
        topo.reserve(4 * n);
        std::function<void(size_t, size_t)> build = [&](size_t b, size_t e) {
            const size_t node = topo.size() / 2;
            topo.push_back(-1);
            topo.push_back(0);
            if (e - b == 1) {
                topo[2 * node] = (int32_t)ids[b];
                topo[2 * node + 1] = (int32_t)node + 1;
                return;
            }
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (size_t k = b; k < e; ++k)
                for (int r = 0; r < 3; ++r) {
                    lo[r] = std::min(lo[r], cen[3 * ids[k] + r]);
                    hi[r] = std::max(hi[r], cen[3 * ids[k] + r]);
                }
            int ax = 0;
            for (int r = 1; r < 3; ++r)
                if (hi[r] - lo[r] > hi[ax] - lo[ax]) ax = r;
            const size_t mid = b + (e - b) / 2;
            std::nth_element(ids.begin() + b, ids.begin() + mid, ids.begin() + e, [&](uint32_t x, uint32_t y) {
                return cen[3 * x + ax] < cen[3 * y + ax] || (cen[3 * x + ax] == cen[3 * y + ax] && x < y);
            });
            build(b, mid);
            build(mid, e);
            topo[2 * node + 1] = (int32_t)(topo.size() / 2);
        };
        build(0, n);
        c.tlas_nodes = (int32_t)(topo.size() / 2);
        c.tlas_topo.alloc(topo.size() * sizeof(int32_t));
        VG_HIP(hipMemcpy(c.tlas_topo.p, topo.data(), topo.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        c.scene_mesh = inst[0].mesh_id;
        c.scene_gen++;
    });
}

vira_gpu_status vira_gpu_lights_set(vira_gpu_ctx* ctx, const vira_gpu_light* lights, int n) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        VG_CHECK(n >= 0 && n <= VIRA_GPU_MAX_LIGHTS, VIRA_GPU_ERR_INVALID, "lights_set: 0..8 lights");
        VG_CHECK(lights || n == 0, VIRA_GPU_ERR_INVALID, "lights_set: null lights");
        for (int l = 0; l < n; ++l) {
            VG_CHECK(lights[l].kind == 0 || lights[l].kind == 1, VIRA_GPU_ERR_INVALID, "lights_set: unknown kind");
            VG_CHECK(lights[l].kind == 0 || lights[l].radius > 0, VIRA_GPU_ERR_INVALID,
                     "lights_set: SphereLight Radius must be positive");  // sphere_light.ipp:20
        }
        c.lights.assign(lights, lights + n);
        c.scene_gen++;
    });
}

vira_gpu_status vira_gpu_material_set(vira_gpu_ctx* ctx, int32_t mesh_id, const vira_gpu_material* mat) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        vg::MeshG& m = c.mesh(mesh_id);
        c.scene_gen++;
        VG_CHECK(mat, VIRA_GPU_ERR_INVALID, "material_set: null material");
        VG_CHECK(!mat->texture_dev || (mat->tex_width > 0 && mat->tex_height > 0), VIRA_GPU_ERR_INVALID,
                 "material_set: texture needs positive size");
        VG_CHECK(mat->tex_channels == 0 || mat->tex_channels == 1 || mat->tex_channels == VIRA_GPU_NUM_BINS,
                 VIRA_GPU_ERR_INVALID, "material_set: tex_channels must be 1 (grey) or 8 (spectral)");
        VG_CHECK(!mat->normal_map_dev || (mat->nmap_width > 0 && mat->nmap_height > 0), VIRA_GPU_ERR_INVALID,
                 "material_set: normal map needs positive size");
        m.mat = *mat;
        if (m.mat.tex_channels == 0) m.mat.tex_channels = 1;
    });
}

vira_gpu_status vira_gpu_environment_set(vira_gpu_ctx* ctx, const float* ambient, const float* background,
                                         int bg_width, int bg_height) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        VG_CHECK(!background || (bg_width > 0 && bg_height > 0), VIRA_GPU_ERR_INVALID,
                 "environment_set: background needs positive size");
        c.scene_gen++;
        bool any = false;
        for (int k = 0; k < VIRA_GPU_NUM_BINS; ++k) {
            c.env_ambient[k] = ambient ? ambient[k] : 0.0f;
            any = any || c.env_ambient[k] != 0.0f;
        }
        c.env_has_ambient = any;  // Scene::setAmbient: hasAmbient when the magnitude is not 0
        c.env_bg = background;
        c.env_bw = background ? bg_width : 0;
        c.env_bh = background ? bg_height : 0;
    });
}

vira_gpu_status vira_gpu_render(vira_gpu_ctx* ctx, const vira_gpu_camera* cam, const vira_gpu_render_opts* opts,
                                const vira_gpu_outputs* out, vira_gpu_render_stats* stats) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(cam && opts && out, VIRA_GPU_ERR_INVALID, "render: null argument");
        vg::render(c, *cam, *opts, *out, stats);
    });
}

vira_gpu_status vira_gpu_render_noalloc(vira_gpu_ctx* ctx, const vira_gpu_camera* cam,
                                        const vira_gpu_render_opts* opts, const vira_gpu_outputs* out,
                                        vira_gpu_render_stats* stats) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(cam && opts && out, VIRA_GPU_ERR_INVALID, "render_noalloc: null argument");
        vg::render_noalloc(c, *cam, *opts, *out, stats);
    });
}

vira_gpu_status vira_gpu_trace_rays(vira_gpu_ctx* ctx, int32_t mesh_id, const float* rays, size_t n, int any_hit,
                                    float* t, uint32_t* id, float* b1, float* b2, double* ms) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(rays || n == 0, VIRA_GPU_ERR_INVALID, "trace_rays: null rays");
        vg::trace_rays(c, c.mesh(mesh_id), rays, n, any_hit, t, id, b1, b2, ms);
    });
}

vira_gpu_status vira_gpu_dem_update(vira_gpu_ctx* ctx, int32_t mesh_id, const float* heights_dev) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(heights_dev, VIRA_GPU_ERR_INVALID, "dem_update: null heights");
        vg::MeshG& m = c.mesh(mesh_id);
        c.scene_gen++;
        vg::dem_update(c, m, heights_dev);
        if (m.blas.built) vg::blas_build(c, m, m.blas.opts);
    });
}

vira_gpu_status vira_gpu_dem_update_refit(vira_gpu_ctx* ctx, int32_t mesh_id, const float* heights_dev) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(heights_dev, VIRA_GPU_ERR_INVALID, "dem_update_refit: null heights");
        vg::MeshG& m = c.mesh(mesh_id);
        c.scene_gen++;
        vg::dem_update(c, m, heights_dev);
        if (m.blas.built) vg::blas_refit(c, m);
    });
}

vira_gpu_status vira_gpu_blas_refit(vira_gpu_ctx* ctx, int32_t mesh_id) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        vg::MeshG& m = c.mesh(mesh_id);
        c.scene_gen++;
        vg::blas_refit(c, m);
    });
}

vira_gpu_status vira_gpu_di_render(vira_gpu_ctx* ctx, const vira_gpu_camera* cam, const vira_gpu_di_opts* opts,
                                   float* radiance_dev, vira_gpu_di_stats* stats) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(cam && opts, VIRA_GPU_ERR_INVALID, "di_render: null argument");
        vg::di_render(c, *cam, *opts, radiance_dev, stats);
    });
}

vira_gpu_status vira_gpu_adjoint(vira_gpu_ctx* ctx, const vira_gpu_camera* cam, const vira_gpu_di_opts* opts,
                                 const float* dL_dimage_dev, float* dL_dheights_dev, float* dL_dtexture_dev,
                                 vira_gpu_di_stats* stats) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(cam && opts, VIRA_GPU_ERR_INVALID, "adjoint: null argument");
        vg::di_adjoint(c, *cam, *opts, dL_dimage_dev, dL_dheights_dev, dL_dtexture_dev, stats);
    });
}

vira_gpu_status vira_gpu_batch_submit(vira_gpu_ctx* ctx, const vira_gpu_request* requests, size_t n,
                                      const vira_gpu_batch_config* cfg, vira_gpu_batch_stats* stats) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(requests || n == 0, VIRA_GPU_ERR_INVALID, "batch_submit: null requests");
        if (n == 0) {
            if (stats) *stats = vira_gpu_batch_stats{};
            return;
        }
        vira_gpu_batch_config g = c.b_cfg;  // the last autotuned configuration, or the default
        if (cfg) g = *cfg;
        else g.bvh_width = c.mesh(requests[0].mesh_id).blas.width;
        vg::batch_submit(c, requests, n, g, 1, stats);
    });
}

vira_gpu_status vira_gpu_batch_autotune(vira_gpu_ctx* ctx, const vira_gpu_request* requests, size_t n,
                                        int sample_stride, vira_gpu_batch_config* best, double* best_ms) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(requests, VIRA_GPU_ERR_INVALID, "batch_autotune: null requests");
        vg::batch_autotune(c, requests, n, sample_stride, best, best_ms);
    });
}

vira_gpu_status vira_gpu_received_power(vira_gpu_ctx* ctx, const vira_gpu_camera* cam, const vira_gpu_sensor* sensor,
                                        const float* radiance_dev, float* power_dev, float* solid_angle_dev) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(cam && sensor, VIRA_GPU_ERR_INVALID, "received_power: null argument");
        vg::received_power(c, *cam, *sensor, radiance_dev, power_dev, solid_angle_dev);
    });
}

vira_gpu_status vira_gpu_psf_kernel(vira_gpu_ctx* ctx, const vira_gpu_psf* psf, int size, float* kernel_dev) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(psf, VIRA_GPU_ERR_INVALID, "psf_kernel: null psf");
        vg::psf_kernel(c, *psf, size, kernel_dev);
    });
}

vira_gpu_status vira_gpu_convolve(vira_gpu_ctx* ctx, const float* image_dev, int width, int height,
                                  const float* kernel_dev, int ksize, float* out_dev) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        vg::convolve(c, image_dev, width, height, kernel_dev, ksize, out_dev);
    });
}

vira_gpu_status vira_gpu_simulate_sensor(vira_gpu_ctx* ctx, const vira_gpu_sensor_sim* cfg, const float* power_dev,
                                         int width, int height, float* photons_dev, float* image_dev) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(cfg, VIRA_GPU_ERR_INVALID, "simulate_sensor: null config");
        vg::simulate_sensor(c, *cfg, power_dev, width, height, photons_dev, image_dev);
    });
}

vira_gpu_status vira_gpu_render_stars(vira_gpu_ctx* ctx, const vira_gpu_camera* cam, const vira_gpu_sensor* sensor,
                                      const vira_gpu_star_psf* psf, const float* directions_dev,
                                      const float* irradiance_dev, size_t n, float* power_dev) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(cam && sensor && psf, VIRA_GPU_ERR_INVALID, "render_stars: null argument");
        vg::render_stars(c, *cam, *sensor, *psf, directions_dev, irradiance_dev, n, power_dev);
    });
}

vira_gpu_status vira_gpu_star_light(vira_gpu_ctx* ctx, const vira_gpu_star* stars_dev, size_t n, double et,
                                    float* directions_dev, float* irradiance_dev) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        vg::star_light(c, stars_dev, n, et, directions_dev, irradiance_dev);
    });
}

vira_gpu_status vira_gpu_qsc_read(const char* path, vira_gpu_star* stars_host, size_t capacity, size_t* count) {
    return guarded([&] { vg::qsc_read(path, stars_host, capacity, count); });
}

int64_t vira_gpu_lz4_decode_block(const void* src, size_t src_size, void* dst, size_t dst_capacity) {
    if (!src || (!dst && dst_capacity)) return -1;
    return vg::lz4_decode_block(static_cast<const uint8_t*>(src), src_size, static_cast<uint8_t*>(dst), dst_capacity);
}

vira_gpu_status vira_gpu_qld_open(const char* path, vira_gpu_qld** qld, vira_gpu_qld_info* info) {
    return guarded([&] {
        VG_CHECK(qld, VIRA_GPU_ERR_INVALID, "qld_open: null handle pointer");
        *qld = vg::qld_open(path, info);
    });
}

vira_gpu_status vira_gpu_qld_close(vira_gpu_qld* qld) {
    return guarded([&] { delete qld; });
}

vira_gpu_status vira_gpu_qld_level_gsd(const vira_gpu_qld* qld, int32_t level, double* gsd) {
    return guarded([&] {
        VG_CHECK(qld && gsd, VIRA_GPU_ERR_INVALID, "qld_level_gsd: null argument");
        VG_CHECK(level >= 0 && level < qld->levels, VIRA_GPU_ERR_INVALID, "qld_level_gsd: level out of range");
        *gsd = qld->gsd[level];
    });
}

vira_gpu_status vira_gpu_qld_select(const vira_gpu_qld* qld, double required_gsd, int32_t* level) {
    return guarded([&] {
        VG_CHECK(qld && level, VIRA_GPU_ERR_INVALID, "qld_select: null argument");
        *level = vg::qld_select(*qld, required_gsd);
    });
}

vira_gpu_status vira_gpu_qld_read(vira_gpu_ctx* ctx, vira_gpu_qld* qld, int32_t level,
                                  const vira_gpu_qld_read_opts* opts, int32_t* mesh_id, float* heights_dev,
                                  float* albedo_dev, vira_gpu_qld_level* info) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(qld && opts, VIRA_GPU_ERR_INVALID, "qld_read: null argument");
        int32_t made = -1;
        try {
            vg::qld_read(c, *qld, level, *opts, heights_dev, albedo_dev, mesh_id ? &made : nullptr, info, [&] {
                made = new_mesh(c);
                return made;
            });
        } catch (...) {
            if (made >= 0) c.meshes[made].reset();
            throw;
        }
        if (mesh_id) {
            c.scene_gen++;
            *mesh_id = made;
        }
    });
}

vira_gpu_status vira_gpu_denoise(vira_gpu_ctx* ctx, const vira_gpu_denoise_opts* opts, int width, int height,
                                 float* direct_dev, float* indirect_dev, const float* albedo_dev,
                                 const float* depth_dev, const float* normal_dev, float* total_dev) {
    return guarded([&] {
        vg::Ctx& c = C(ctx);
        set_device(c);
        VG_CHECK(opts, VIRA_GPU_ERR_INVALID, "denoise: null options");
        vg::denoise(c, *opts, width, height, direct_dev, indirect_dev, albedo_dev, depth_dev, normal_dev, total_dev);
    });
}

}  // extern "C"

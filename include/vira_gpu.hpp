// vira_gpu.hpp -- C++ wrapper of the vira_gpu C-ABI (include/vira_gpu.h).
//
// RAII ownership of the context, and every non-OK status rethrown as
// vira_gpu::Error (a std::runtime_error), as the reference throws
// (vira_tlas.ipp:33, cpu_path_tracer.ipp:121).  PathTracer mirrors the public
// surface of vira::rendering::CPUPathTracer (cpu_path_tracer.hpp:46-71):
// `options`, `renderPasses` and `render(camera)` over the context's scene.
// No HIP types appear here: device buffers are plain pointers the caller owns.
#ifndef VIRA_GPU_HPP
#define VIRA_GPU_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vira_gpu.h"

namespace vira_gpu {

class Error : public std::runtime_error {
public:
    Error(vira_gpu_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
    vira_gpu_status status;
};

inline void check(vira_gpu_status s) {
    if (s != VIRA_GPU_OK) throw Error(s, vira_gpu_last_error());
}

inline std::string version() { return vira_gpu_version(); }

// One GPU, one stream; owns meshes, BVHs, lights and materials (the parts of
// vira::Scene the renderer reads, scene.hpp:212-302).
class Context {
public:
    explicit Context(int device = 0, void* stream = nullptr) {
        vira_gpu_opts o{device, stream};
        check(vira_gpu_create(&h_, &o));
    }
    ~Context() {
        if (h_) (void)vira_gpu_destroy(h_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    Context(Context&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    Context& operator=(Context&& o) noexcept {
        if (this != &o) {
            if (h_) (void)vira_gpu_destroy(h_);
            h_ = std::exchange(o.h_, nullptr);
        }
        return *this;
    }
    vira_gpu_ctx* get() const { return h_; }

    void synthFbm(int w, int h, uint32_t seed, double amplitude, int octaves, double base_cells, float* heights_dev) {
        check(vira_gpu_synth_fbm(h_, w, h, seed, amplitude, octaves, base_cells, heights_dev));
    }
    // DEM::makeMesh (dem.ipp:138-211)
    int32_t demMesh(const float* heights_dev, int w, int h, const vira_gpu_projection& p,
                    const float* grey_albedo_dev = nullptr) {
        int32_t id = -1;
        check(vira_gpu_dem_mesh(h_, heights_dev, w, h, &p, grey_albedo_dev, &id));
        return id;
    }
    // Mesh(vertexBuffer, indexBuffer) (mesh.ipp:25-83)
    int32_t meshUpload(const std::vector<float>& positions, const std::vector<uint32_t>& indices,
                       const float* albedo = nullptr, const float* uv = nullptr, const float* normals = nullptr) {
        int32_t id = -1;
        check(vira_gpu_mesh_upload(h_, positions.data(), positions.size() / 3, indices.data(), indices.size() / 3,
                                   albedo, uv, normals, &id));
        return id;
    }
    void releaseMesh(int32_t mesh) { check(vira_gpu_mesh_release(h_, mesh)); }
    // refit = true keeps the BLAS's tree and recomputes its boxes instead of rebuilding it
    void demUpdate(int32_t mesh, const float* heights_dev, bool refit = false) {
        check(refit ? vira_gpu_dem_update_refit(h_, mesh, heights_dev) : vira_gpu_dem_update(h_, mesh, heights_dev));
    }
    // Mesh::buildBVH (mesh.ipp:396-418)
    vira_gpu_bvh_stats buildBLAS(int32_t mesh, vira_gpu_bvh_opts o = {4, 16, 4}) {
        check(vira_gpu_blas_build(h_, mesh, &o));
        vira_gpu_bvh_stats s{};
        check(vira_gpu_blas_stats(h_, mesh, &s));
        return s;
    }
    // New: the built BLAS refitted to the mesh's current vertices (tree kept, boxes recomputed)
    vira_gpu_bvh_stats refitBLAS(int32_t mesh) {
        check(vira_gpu_blas_refit(h_, mesh));
        vira_gpu_bvh_stats s{};
        check(vira_gpu_blas_stats(h_, mesh, &s));
        return s;
    }
    // Scene::buildTLAS (scene.ipp:712-812)
    void buildTLAS(const std::vector<vira_gpu_instance>& instances) {
        check(vira_gpu_tlas_build(h_, instances.data(), instances.size()));
    }
    void setLights(const std::vector<vira_gpu_light>& lights) {
        check(vira_gpu_lights_set(h_, lights.data(), (int)lights.size()));
    }
    void setMaterial(int32_t mesh, const vira_gpu_material& m) { check(vira_gpu_material_set(h_, mesh, &m)); }
    // Scene::setAmbient / setBackgroundEmission (scene.hpp:90-113): 8 ambient values (or
    // nullptr) and a W*H*8 device background image (or nullptr)
    void setEnvironment(const float* ambient, const float* background_dev = nullptr, int bg_width = 0,
                        int bg_height = 0) {
        check(vira_gpu_environment_set(h_, ambient, background_dev, bg_width, bg_height));
    }
    void synchronize() { check(vira_gpu_synchronize(h_)); }

    // --- camera sensor tail (camera.ipp:1175-1340) on W*H*8 device images
    void receivedPower(const vira_gpu_camera& cam, const vira_gpu_sensor& sensor, const float* radiance_dev,
                       float* power_dev, float* solid_angle_dev = nullptr) {
        check(vira_gpu_received_power(h_, &cam, &sensor, radiance_dev, power_dev, solid_angle_dev));
    }
    // PointSpreadFunction::makeKernel (psf.ipp): size*size*8 floats
    void psfKernel(const vira_gpu_psf& psf, int size, float* kernel_dev) {
        check(vira_gpu_psf_kernel(h_, &psf, size, kernel_dev));
    }
    // Image::convolve (image.ipp:993-1065)
    void convolve(const float* image_dev, int w, int h, const float* kernel_dev, int ksize, float* out_dev) {
        check(vira_gpu_convolve(h_, image_dev, w, h, kernel_dev, ksize, out_dev));
    }
    // Camera::simulateSensor: photons (optional) and the normalised image
    void simulateSensor(const vira_gpu_sensor_sim& cfg, const float* power_dev, int w, int h, float* photons_dev,
                        float* image_dev) {
        check(vira_gpu_simulate_sensor(h_, &cfg, power_dev, w, h, photons_dev, image_dev));
    }
    // denoiseSpectralRadianceEATWT (cpu_denoise.ipp:10-187), in place
    void denoise(const vira_gpu_denoise_opts& o, int w, int h, float* direct_dev, float* indirect_dev,
                 const float* albedo_dev, const float* depth_dev, const float* normal_dev, float* total_dev) {
        check(vira_gpu_denoise(h_, &o, w, h, direct_dev, indirect_dev, albedo_dev, depth_dev, normal_dev, total_dev));
    }

private:
    vira_gpu_ctx* h_ = nullptr;
};

// CPUPathTracerOptions (cpu_path_tracer.hpp:30-43) and the GPU additions.
struct PathTracerOptions {
    int samples = 1;
    int bounces = 0;
    bool russianRoulette = false;
    int rrStart = 1;
    uint32_t seed = 0;
    int integrator = 0;  // 0 auto, 1 per-pixel, 2 wavefront
    bool denoise = false;  // CPUPathTracerOptions::denoise (cpu_path_tracer.hpp:35)
    bool showBackground = false;          // show_background
    bool adaptiveSampling = false;        // adaptive_sampling
    int samplesPerBatch = 30;             // samples_per_batch
    float samplingTolerance = 0.05f;      // sampling_tolerance
    int samplesToDetectMiss = 30;         // samples_to_detect_miss
    bool countVisits = false;             // BVH visit counters into lastStats (per-pixel integrator)
};

// EATWTOptions (cpu_denoise.hpp:9-18).
inline vira_gpu_denoise_opts defaultDenoiserOptions() { return vira_gpu_denoise_opts{5, 5, 0.2f, 0.5f, 0.3f, 1e-6f}; }

// GPUPathTracer: CPUPathTracer's render(camera, scene) with the scene held by
// the context; renderPasses are caller-owned device buffers (W*H*8 floats per
// spectral pass, as RenderPasses::updateImages fills them, passes.ipp:81-127).
class PathTracer {
public:
    explicit PathTracer(Context& ctx) : ctx_(&ctx) {}
    PathTracerOptions options;
    vira_gpu_denoise_opts denoiserOptions = defaultDenoiserOptions();
    vira_gpu_outputs renderPasses{};
    vira_gpu_render_stats lastStats{};

    // With options.denoise the direct and indirect passes are denoised after the
    // render and the total pass becomes their sum (cpu_path_tracer.ipp:124-134).
    void render(const vira_gpu_camera& camera) {
        vira_gpu_render_opts o{};
        o.samples = options.samples;
        o.bounces = options.bounces;
        o.russian_roulette = options.russianRoulette ? 1 : 0;
        o.rr_start = options.rrStart;
        o.seed = options.seed;
        o.integrator = options.integrator;
        o.show_background = options.showBackground ? 1 : 0;
        o.adaptive_sampling = options.adaptiveSampling ? 1 : 0;
        o.samples_per_batch = options.samplesPerBatch;
        o.sampling_tolerance = options.samplingTolerance;
        o.samples_to_detect_miss = options.samplesToDetectMiss;
        o.count_visits = options.countVisits ? 1 : 0;
        check(vira_gpu_render(ctx_->get(), &camera, &o, &renderPasses, &lastStats));
        if (options.denoise)
            check(vira_gpu_denoise(ctx_->get(), &denoiserOptions, camera.width, camera.height,
                                   renderPasses.direct_dev, renderPasses.indirect_dev, renderPasses.albedo_dev,
                                   renderPasses.depth_dev, renderPasses.normal_dev, renderPasses.radiance_dev));
    }

private:
    Context* ctx_;
};

// CPUUnresolvedRenderer (cpu_unresolved_renderer.ipp:25-200): point sources
// splatted into a received-power image; `psf` carries the kernels (device
// pointers the caller owns), their edge maxima, the minimum power and K.
class UnresolvedRenderer {
public:
    explicit UnresolvedRenderer(Context& ctx) : ctx_(&ctx) {}
    vira_gpu_star_psf psf{};
    void render(const vira_gpu_camera& camera, const vira_gpu_sensor& sensor, const float* directions_dev,
                const float* irradiance_dev, size_t n, float* power_dev) {
        check(vira_gpu_render_stars(ctx_->get(), &camera, &sensor, &psf, directions_dev, irradiance_dev, n,
                                    power_dev));
    }

private:
    Context* ctx_;
};

// StarQuipu::read (star_quipu.ipp:20-29): a .qsc star catalogue into host records.
inline std::vector<vira_gpu_star> readStarQuipu(const std::string& path) {
    size_t n = 0;
    check(vira_gpu_qsc_read(path.c_str(), nullptr, 0, &n));
    std::vector<vira_gpu_star> stars(n);
    if (n) check(vira_gpu_qsc_read(path.c_str(), stars.data(), n, &n));
    return stars;
}

// StarCatalogue::makeStarLight (star_catalogue.ipp:55-68): n device records to
// unit directions (3 floats each) and band irradiance (8 floats each) at et.
inline void makeStarLight(Context& ctx, const vira_gpu_star* stars_dev, size_t n, double et, float* directions_dev,
                          float* irradiance_dev) {
    check(vira_gpu_star_light(ctx.get(), stars_dev, n, et, directions_dev, irradiance_dev));
}

// DEMQuipu (dem_quipu.hpp:34-81): a .qld DEM pyramid.  readBuffers builds the
// chosen level's mesh on the GPU and keeps it while the choice does not
// change (dem_quipu.ipp:131-163); the mesh belongs to the context.
class QuipuDEM {
public:
    explicit QuipuDEM(const std::string& path) { check(vira_gpu_qld_open(path.c_str(), &h_, &info_)); }
    ~QuipuDEM() {
        if (h_) (void)vira_gpu_qld_close(h_);
    }
    QuipuDEM(const QuipuDEM&) = delete;
    QuipuDEM& operator=(const QuipuDEM&) = delete;
    const vira_gpu_qld_info& info() const { return info_; }
    int levels() const { return info_.levels; }
    double gsd(int level) const {
        double g = 0;
        check(vira_gpu_qld_level_gsd(h_, level, &g));
        return g;
    }
    vira_gpu_qld_read_opts options = defaultOptions();
    static vira_gpu_qld_read_opts defaultOptions() {
        vira_gpu_qld_read_opts o{};
        o.read_albedos = 1;
        o.default_albedo = 0.06f;
        for (int k = 0; k < VIRA_GPU_NUM_BINS; ++k) o.albedo_profile[k] = 1.0f;
        return o;
    }
    // required_gsd: the coarsest level whose GSD is <= it (infinity: the coarsest)
    int32_t readBuffers(Context& ctx, double required_gsd) {
        int32_t level = 0;
        check(vira_gpu_qld_select(h_, required_gsd, &level));
        if (level == level_ && mesh_ >= 0) return mesh_;
        int32_t mesh = -1;
        check(vira_gpu_qld_read(ctx.get(), h_, level, &options, &mesh, nullptr, nullptr, &last_));
        if (mesh_ >= 0) ctx.releaseMesh(mesh_);
        mesh_ = mesh;
        level_ = level;
        return mesh_;
    }
    const vira_gpu_qld_level& lastLevel() const { return last_; }

private:
    vira_gpu_qld* h_ = nullptr;
    vira_gpu_qld_info info_{};
    vira_gpu_qld_level last_{};
    int32_t level_ = -1, mesh_ = -1;
};

// Batched requests in one persistent launch (vira_gpu_batch_submit).
inline vira_gpu_batch_stats renderBatch(Context& ctx, const std::vector<vira_gpu_request>& requests,
                                        const vira_gpu_batch_config* cfg = nullptr) {
    vira_gpu_batch_stats s{};
    check(vira_gpu_batch_submit(ctx.get(), requests.data(), requests.size(), cfg, &s));
    return s;
}

// Differentiable direct illumination: forward, then the adjoint.
inline vira_gpu_di_stats renderDirect(Context& ctx, const vira_gpu_camera& cam, const vira_gpu_di_opts& o,
                                      float* radiance_dev) {
    vira_gpu_di_stats s{};
    check(vira_gpu_di_render(ctx.get(), &cam, &o, radiance_dev, &s));
    return s;
}
inline vira_gpu_di_stats adjoint(Context& ctx, const vira_gpu_camera& cam, const vira_gpu_di_opts& o,
                                 const float* dL_dimage_dev, float* dL_dheights_dev, float* dL_dtexture_dev) {
    vira_gpu_di_stats s{};
    check(vira_gpu_adjoint(ctx.get(), &cam, &o, dL_dimage_dev, dL_dheights_dev, dL_dtexture_dev, &s));
    return s;
}

}  // namespace vira_gpu

#endif  // VIRA_GPU_HPP

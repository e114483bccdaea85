// render_dem.cpp -- the C++ host path end to end: a DEM rendered through the
// vira_gpu C++ wrapper (include/vira_gpu.hpp), as a vira-side GPUPathTracer
// would drive it.  Prints one JSON object: the scene it built (so another
// client can rebuild it) and checksums of the passes.
//
//   render_dem [dem_side] [image_side] [spp] [bounces]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vira_gpu.hpp"

namespace {
template <class T>
T* dev_alloc(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    return static_cast<T*>(p);
}
}  // namespace

int main(int argc, char** argv) {
    const int dem = argc > 1 ? std::atoi(argv[1]) : 65;
    const int img = argc > 2 ? std::atoi(argv[2]) : 64;
    const int spp = argc > 3 ? std::atoi(argv[3]) : 2;
    const int bounces = argc > 4 ? std::atoi(argv[4]) : 1;
    try {
        vira_gpu::Context ctx(0);
        float* heights = dev_alloc<float>((size_t)dem * dem);
        ctx.synthFbm(dem, dem, 0u, 500.0, 6, 8.0, heights);
        vira_gpu_projection proj{};
        proj.kind = 0;
        proj.spacing = 30.0;
        const int32_t mesh = ctx.demMesh(heights, dem, dem, proj);
        const vira_gpu_bvh_stats bs = ctx.buildBLAS(mesh);
        vira_gpu_instance inst{};
        inst.mesh_id = mesh;
        for (int k = 0; k < 16; ++k) inst.transform[k] = (k % 5 == 0) ? 1.0 : 0.0;
        ctx.buildTLAS({inst});
        // a distant point sun at 30 deg elevation giving 1 W/m^2 per band
        vira_gpu_light sun{};
        sun.kind = 0;
        const double d = 1.495978707e11, el = 30.0 * M_PI / 180.0;
        sun.position[0] = d * std::cos(el);
        sun.position[2] = d * std::sin(el);
        for (int k = 0; k < VIRA_GPU_NUM_BINS; ++k) sun.value[k] = (float)(d * d);
        ctx.setLights({sun});
        vira_gpu_material mat{};
        for (int k = 0; k < VIRA_GPU_NUM_BINS; ++k) mat.albedo[k] = 0.15f;
        ctx.setMaterial(mesh, mat);
        // pinhole, 40 deg field of view, nadir from 1.2x the DEM half-width / tan(20 deg)
        vira_gpu_camera cam{};
        cam.width = cam.height = img;
        const double f = 0.5 * img / std::tan(20.0 * M_PI / 180.0);
        const float kinv[6] = {(float)(1.0 / f), 0.f, 0.f, (float)(1.0 / f), (float)(-0.5 * img / f),
                               (float)(-0.5 * img / f)};
        std::memcpy(cam.kinv, kinv, sizeof(kinv));
        const float rot[9] = {1.f, 0.f, 0.f, 0.f, -1.f, 0.f, 0.f, 0.f, -1.f};
        std::memcpy(cam.rot, rot, sizeof(rot));
        cam.position[2] = 1.2 * 0.5 * (dem - 1) * 30.0 / std::tan(20.0 * M_PI / 180.0);
        cam.z_dir = 1.f;

        vira_gpu::PathTracer pt(ctx);
        pt.options.samples = spp;
        pt.options.bounces = bounces;
        pt.options.russianRoulette = true;
        const size_t npix = (size_t)img * img;
        pt.renderPasses.radiance_dev = dev_alloc<float>(npix * VIRA_GPU_NUM_BINS);
        pt.renderPasses.tri_id_dev = dev_alloc<uint32_t>(npix);
        pt.render(cam);
        std::vector<uint32_t> rad(npix * VIRA_GPU_NUM_BINS), tri(npix);
        if (hipMemcpy(rad.data(), pt.renderPasses.radiance_dev, rad.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(tri.data(), pt.renderPasses.tri_id_dev, tri.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        uint64_t hits = 0, tri_sum = 0, rad_hash = 1469598103934665603ull;
        for (uint32_t t : tri)
            if (t != 0xffffffffu) {
                hits++;
                tri_sum += t;
            }
        for (uint32_t b : rad) rad_hash = (rad_hash ^ b) * 1099511628211ull;  // FNV-1a over the float bits
        int bad_status = 0;
        try {
            ctx.buildBLAS(12345);
        } catch (const vira_gpu::Error& e) {
            bad_status = (int)e.status;
        }
        std::printf("{\"dem\": %d, \"image\": %d, \"spp\": %d, \"bounces\": %d, \"kinv\": [", dem, img, spp, bounces);
        for (int k = 0; k < 6; ++k) std::printf("%s%.9g", k ? ", " : "", cam.kinv[k]);
        std::printf("], \"position_z\": %.17g, \"bvh_nodes\": %llu, \"primary\": %llu, \"shadow\": %llu, "
                    "\"hits\": %llu, \"tri_sum\": %llu, \"radiance_fnv\": \"%016llx\", \"bad_mesh_status\": %d}\n",
                    cam.position[2], (unsigned long long)bs.num_nodes, (unsigned long long)pt.lastStats.primary_rays,
                    (unsigned long long)pt.lastStats.shadow_rays, (unsigned long long)hits,
                    (unsigned long long)tri_sum, (unsigned long long)rad_hash, bad_status);
        (void)hipFree(heights);
        (void)hipFree(pt.renderPasses.radiance_dev);
        (void)hipFree(pt.renderPasses.tri_id_dev);
    } catch (const vira_gpu::Error& e) {
        std::printf("{\"error\": \"%s\", \"status\": %d}\n", e.what(), (int)e.status);
        return 3;
    } catch (const std::exception& e) {
        std::printf("{\"error\": \"%s\"}\n", e.what());
        return 4;
    }
    return 0;
}

// quipu_scene.cpp -- the C++ host path for a terrain read from a quipu DEM
// pyramid: DEMQuipu::readBuffers picks the level for a GSD and the mesh is
// built on the GPU, then BLAS + TLAS, a render with options.denoise (the
// EATWT step of CPUPathTracer::render), received power, a few unresolved
// sources without PSF (CPUUnresolvedRenderer) and the photosite image
// (Camera::simulateSensor), all through include/vira_gpu.hpp.  Prints one JSON
// object with the level read and FNV-1a hashes of the resulting images.
//
//   quipu_scene <file.qld> [image_side] [required_gsd]
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
    if (hipMemset(p, 0, n * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMemset failed");
    return static_cast<T*>(p);
}
template <class T>
std::vector<T> to_host(const T* d, size_t n) {
    std::vector<T> h(n);
    if (hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("hipMemcpy failed");
    return h;
}
template <class T>
unsigned long long fnv(const std::vector<T>& v) {
    unsigned long long h = 1469598103934665603ull;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(v.data());
    for (size_t i = 0; i < v.size() * sizeof(T) / 4; ++i) h = (h ^ w[i]) * 1099511628211ull;
    return h;
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::printf("{\"error\": \"usage: quipu_scene <file.qld> [image_side] [required_gsd]\"}\n");
        return 2;
    }
    const int img = argc > 2 ? std::atoi(argv[2]) : 96;
    const double gsd = argc > 3 ? std::atof(argv[3]) : INFINITY;
    try {
        vira_gpu::Context ctx(0);
        vira_gpu::QuipuDEM dem(argv[1]);
        const int32_t mesh = dem.readBuffers(ctx, gsd);
        const vira_gpu_qld_level lv = dem.lastLevel();
        ctx.buildBLAS(mesh);
        vira_gpu_instance inst{};
        inst.mesh_id = mesh;
        for (int k = 0; k < 16; ++k) inst.transform[k] = (k % 5 == 0) ? 1.0 : 0.0;
        ctx.buildTLAS({inst});
        vira_gpu_light sun{};
        sun.kind = 0;
        const double d = 1.495978707e11, el = 35.0 * M_PI / 180.0;
        sun.position[0] = d * std::cos(el);
        sun.position[2] = d * std::sin(el);
        for (int k = 0; k < VIRA_GPU_NUM_BINS; ++k) sun.value[k] = (float)(d * d);
        ctx.setLights({sun});
        vira_gpu_material mat{};
        for (int k = 0; k < VIRA_GPU_NUM_BINS; ++k) mat.albedo[k] = 1.0f;  // the file's albedo is per vertex
        ctx.setMaterial(mesh, mat);
        // nadir pinhole, 40 degree field, high enough to see the level's whole extent
        vira_gpu_camera cam{};
        cam.width = cam.height = img;
        const double f = 0.5 * img / std::tan(20.0 * M_PI / 180.0);
        const float kinv[6] = {(float)(1.0 / f), 0.f, 0.f, (float)(1.0 / f), (float)(-0.5 * img / f),
                               (float)(-0.5 * img / f)};
        std::memcpy(cam.kinv, kinv, sizeof(kinv));
        const float rot[9] = {1.f, 0.f, 0.f, 0.f, -1.f, 0.f, 0.f, 0.f, -1.f};
        std::memcpy(cam.rot, rot, sizeof(rot));
        const double half = 0.5 * (lv.width - 1) * lv.gsd;
        cam.position[2] = 1.1 * half / std::tan(20.0 * M_PI / 180.0);
        cam.z_dir = 1.f;

        vira_gpu::PathTracer pt(ctx);
        pt.options.samples = 4;
        pt.options.bounces = 1;
        pt.options.denoise = true;
        const size_t npix = (size_t)img * img, nspec = npix * VIRA_GPU_NUM_BINS;
        pt.renderPasses.radiance_dev = dev_alloc<float>(nspec);
        pt.renderPasses.direct_dev = dev_alloc<float>(nspec);
        pt.renderPasses.indirect_dev = dev_alloc<float>(nspec);
        pt.renderPasses.albedo_dev = dev_alloc<float>(nspec);
        pt.renderPasses.normal_dev = dev_alloc<float>(npix * 3);
        pt.renderPasses.depth_dev = dev_alloc<float>(npix);
        pt.render(cam);

        // received power, three sources without PSF, then the photosite image
        vira_gpu_sensor sensor{};
        for (int k = 0; k < VIRA_GPU_NUM_BINS; ++k) sensor.optical_efficiency[k] = 1.0f;
        sensor.aperture_area = 2.5e-4f;
        float* power = dev_alloc<float>(nspec);
        ctx.receivedPower(cam, sensor, pt.renderPasses.radiance_dev, power);
        const float dirs_h[9] = {0.01f, 0.02f, -1.f, -0.05f, 0.03f, -1.f, 0.3f, 0.2f, 1.f};  // the last behind
        std::vector<float> irr_h(3 * VIRA_GPU_NUM_BINS, 1e-3f);
        float* dirs = dev_alloc<float>(9);
        float* irr = dev_alloc<float>(irr_h.size());
        if (hipMemcpy(dirs, dirs_h, sizeof(dirs_h), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(irr, irr_h.data(), irr_h.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        vira_gpu::UnresolvedRenderer stars(ctx);
        stars.psf.count = 0;  // a camera without PSF: each source's power lands on one pixel
        const float K[6] = {(float)f, 0.f, 0.f, (float)f, (float)(0.5 * img), (float)(0.5 * img)};
        std::memcpy(stars.psf.intrinsic, K, sizeof(K));
        stars.psf.cull_behind = 1;
        stars.render(cam, sensor, dirs, irr, 3, power);
        vira_gpu_sensor_sim sim{};
        sim.exposure_time = 0.01f;
        for (int k = 0; k < VIRA_GPU_NUM_BINS; ++k) {
            const double lam = (400.0 + 75.0 * (k + 0.5)) * 1e-9;
            sim.photon_energy[k] = (float)(6.62607015e-34 * 299792458.0 / lam);
            sim.quantum_efficiency[k] = 0.5f;
        }
        sim.gain = 0.5f;
        sim.max_adu = 4095.f;
        float* image = dev_alloc<float>(npix);
        ctx.simulateSensor(sim, power, img, img, nullptr, image);

        const auto rad = to_host(pt.renderPasses.radiance_dev, nspec);
        const auto pw = to_host(power, nspec);
        const auto im = to_host(image, npix);
        double isum = 0;
        for (float v : im) isum += v;
        std::printf("{\"level\": %d, \"gsd\": %.17g, \"width\": %d, \"height\": %d, \"height_bits\": %d, "
                    "\"compressed\": %d, \"levels\": %d, \"position_z\": %.17g, \"kinv\": [",
                    lv.level, lv.gsd, lv.width, lv.height, lv.height_bits, lv.compressed, dem.levels(),
                    cam.position[2]);
        for (int k = 0; k < 6; ++k) std::printf("%s%.9g", k ? ", " : "", cam.kinv[k]);
        std::printf("], \"f\": %.17g, \"primary\": %llu, \"radiance_fnv\": \"%016llx\", \"power_fnv\": \"%016llx\", "
                    "\"image_fnv\": \"%016llx\", \"image_sum\": %.9g}\n",
                    f, (unsigned long long)pt.lastStats.primary_rays, fnv(rad), fnv(pw), fnv(im), isum);
        for (void* p : {(void*)pt.renderPasses.radiance_dev, (void*)pt.renderPasses.direct_dev,
                        (void*)pt.renderPasses.indirect_dev, (void*)pt.renderPasses.albedo_dev,
                        (void*)pt.renderPasses.normal_dev, (void*)pt.renderPasses.depth_dev, (void*)power,
                        (void*)dirs, (void*)irr, (void*)image})
            (void)hipFree(p);
    } catch (const vira_gpu::Error& e) {
        std::printf("{\"error\": \"%s\", \"status\": %d}\n", e.what(), (int)e.status);
        return 3;
    } catch (const std::exception& e) {
        std::printf("{\"error\": \"%s\"}\n", e.what());
        return 4;
    }
    return 0;
}

// scene_params.hpp -- host-side conversion of the C-ABI scene state into the
// kernels' parameter blocks (shared by the renderers).
#pragma once

#include "context.hpp"
#include "shading.hpp"

namespace vg {

inline CamF cam_params(const vira_gpu_camera& cam) {
    CamF cf;
    cf.W = cam.width;
    cf.H = cam.height;
    for (int k = 0; k < 6; ++k) cf.kinv[k] = cam.kinv[k];
    for (int k = 0; k < 9; ++k) cf.rot[k] = cam.rot[k];
    cf.ox = (float)cam.position[0];
    cf.oy = (float)cam.position[1];
    cf.oz = (float)cam.position[2];
    cf.zdir = cam.z_dir;
    return cf;
}

inline MatF mat_params(const MeshG& m) {
    MatF mf;
    for (int k = 0; k < NB; ++k) mf.albedo[k] = m.mat.albedo[k];
    mf.tex = m.mat.texture_dev;
    mf.tw = m.mat.tex_width;
    mf.th = m.mat.tex_height;
    mf.smooth = m.mat.smooth_shading;
    return mf;
}

inline LightF light_params(const vira_gpu_light& L) {
    LightF f;
    f.kind = L.kind;
    f.px = (float)L.position[0];
    f.py = (float)L.position[1];
    f.pz = (float)L.position[2];
    f.radius = L.radius;
    for (int k = 0; k < NB; ++k) f.value[k] = L.value[k];
    return f;
}

// Copies the context's lights (rebased FP64 -> FP32) into buf; returns the kernel view.
inline Lights upload_lights(Ctx& c, DBuf& buf) {
    const int nl = (int)c.lights.size();
    LightF hl[VIRA_GPU_MAX_LIGHTS];
    for (int l = 0; l < nl; ++l) hl[l] = light_params(c.lights[l]);
    buf.alloc(sizeof(hl));
    VG_HIP(hipMemcpyAsync(buf.p, hl, sizeof(LightF) * std::max(1, nl), hipMemcpyHostToDevice, c.stream));
    return Lights{nl, buf.as<LightF>()};
}

// Image of the 8x8-tile pixel order used by the per-pixel kernels.
struct TileGrid {
    uint32_t tiles_x, tiles, nq;
    explicit TileGrid(const vira_gpu_camera& cam) {
        tiles_x = (uint32_t)((cam.width + 7) / 8);
        tiles = tiles_x * (uint32_t)((cam.height + 7) / 8);
        nq = tiles * 64u;
    }
};

}  // namespace vg

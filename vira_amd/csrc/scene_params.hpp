// scene_params.hpp -- host-side conversion of the C-ABI scene state into the
// kernels' parameter blocks (shared by the renderers).
#pragma once

#include "context.hpp"
#include "fused.hpp"
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
    mf.fnrm = m.fnrm.as<float4>();
    for (int k = 0; k < NB; ++k) mf.albedo[k] = m.mat.albedo[k];
    mf.tex = m.mat.texture_dev;
    mf.tw = m.mat.tex_width;
    mf.th = m.mat.tex_height;
    mf.smooth = m.mat.smooth_shading;
    mf.tch = m.mat.tex_channels == NB ? NB : 1;
    mf.nmap = m.mat.normal_map_dev;
    mf.nw = m.mat.nmap_width;
    mf.nh = m.mat.nmap_height;
    return mf;
}

inline OutPtrs out_params(const vira_gpu_outputs& o) {
    return OutPtrs{o.radiance_dev, o.direct_dev,    o.indirect_dev,      o.albedo_dev,  o.normal_dev,
                   o.depth_dev,    o.alpha_dev,     o.tri_id_dev,        o.instance_id_dev,
                   o.normal_camera_dev, o.mesh_id_dev, o.material_id_dev};
}

// The context's light environment for one render (show_background from its options).
inline EnvF env_params(const Ctx& c, const vira_gpu_render_opts& o) {
    EnvF e{};
    for (int k = 0; k < NB; ++k) e.ambient[k] = c.env_ambient[k];
    e.has_ambient = c.env_has_ambient ? 1 : 0;
    e.show_bg = o.show_background ? 1 : 0;
    e.bg = c.env_bg;
    e.bw = c.env_bw;
    e.bh = c.env_bh;
    return e;
}

// CPUPathTracerOptions' adaptive sampling; zero parameters stand for the
// reference's defaults (cpu_path_tracer.hpp:39-42).
inline Adaptive adaptive_params(const vira_gpu_render_opts& o) {
    Adaptive a;
    a.on = o.adaptive_sampling ? 1 : 0;
    a.batch = o.samples_per_batch > 0 ? o.samples_per_batch : 30;
    a.miss = o.samples_to_detect_miss > 0 ? o.samples_to_detect_miss : 30;
    a.tol = o.sampling_tolerance > 0.0f ? o.sampling_tolerance : 0.05f;
    return a;
}

// View normal matrix (the camera's, or the inverse of its rotation: the
// transpose) and the scene mesh's ids for the passes.
inline PassInfo pass_params(const vira_gpu_camera& cam, int32_t mesh_id, const MeshG& m) {
    PassInfo p;
    bool given = false;
    for (int k = 0; k < 9; ++k) given = given || cam.view_normal[k] != 0.0f;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) p.vn[r * 3 + k] = given ? cam.view_normal[r * 3 + k] : cam.rot[k * 3 + r];
    p.mesh_id = (uint32_t)mesh_id;
    p.material_id = m.mat.material_id;
    return p;
}

// The per-pixel renderer's variant: several lights or ambient light keep the
// vertex in LDS while the shadow rays are traced one after the other.
inline bool fused_multi(const Ctx& c) { return c.lights.size() > 1 || c.env_has_ambient; }

// A material using what the plain per-pixel variant leaves out (normal map,
// spectral albedo map).
inline bool material_feat(const vira_gpu_material& m) { return m.normal_map_dev || m.tex_channels == NB; }

// The per-pixel renderer's feature variant: background shown, adaptive
// sampling, or a material of the scene with a normal or spectral albedo map.
inline bool fused_feat(Ctx& c, const vira_gpu_render_opts& o) {
    if ((o.show_background && c.env_bg) || o.adaptive_sampling) return true;
    if (c.instanced) {
        for (const vira_gpu_instance& I : c.tlas)
            if (material_feat(c.mesh(I.mesh_id).mat)) return true;
        return false;
    }
    return material_feat(c.mesh(c.scene_mesh).mat);
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

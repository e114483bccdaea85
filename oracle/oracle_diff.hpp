// oracle_diff.hpp -- FP64 reference of the fixed-visibility direct-illumination
// render and its gradient with respect to DEM heights and albedo texels.
//
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).  The reference has no
// differentiable renderer; this is the checker of the GPU adjoint.  The
// forward model is the bounce-0 direct-light term of processIntersection
// (cpu_path_tracer.ipp:284-349,399-428) for point lights, evaluated on hit
// records (triangle id, barycentrics, per-light visibility) that are held
// fixed while the geometry moves:
//   p_v(h)  = DEM node position (dem.ipp:165-211) at height h_v, FP64, rebased;
//   fn      = normalize(cross(p1 - p0, p2 - p0))            (triangle.ipp:36-44);
//   n_v     = normalize(sum of unit face normals around v)   (mesh.ipp:199-233);
//   N       = smooth ? normalize(w0 n1 + w1 n2 + w2 n0) : fn;
//   P       = computeShadingPoint of the barycentric point (cpu_path_tracer.ipp:481-511),
//             without the ULP offset of the shadow-ray origin (ray.ipp:107-139);
//   c       = N . normalize(L - P),  E_k = I_k / |L - P|^2  (point_light.ipp:18-38);
//   L_k     = albedo_k * E_k * c^2 / (pi (1 + (c / pi)^2))  for c > 0 and a visible light
//             (Lambertian f = rho/pi max(c, 0), times cos, power heuristic with
//             light pdf 1 and BSDF pdf c/pi, cpu_path_tracer.ipp:412-428,472-478);
//   albedo  = vertex albedo (barycentric) x material albedo x bilinear texel
//             (image.ipp:589-670) at the interpolated uv;
//   image   = sum over samples / spp.
// The gradient is written by hand from the chain rule on these expressions,
// evaluated pixel by pixel with the normals rebuilt from the heights on the fly
// (a different organisation from the GPU's per-vertex passes).  The central
// finite difference of the L2 loss over exactly the pixels a vertex touches
// checks it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "oracle_core.hpp"

namespace vo {

struct DemDiff {
    int W = 0, H = 0;
    Projection proj{};
    std::vector<uint32_t> idx;       // triangles over grid node ids (dem_indices order)
    std::vector<float> uv;           // 2 per node
    std::vector<float> alb;          // NB per node (vertex albedo)
    std::vector<uint32_t> adj_off;   // node -> incident triangles (CSR, triangle order)
    std::vector<uint32_t> adj;
};

// Hit record of one (pixel, sample): id = 0xffffffff for a miss; vis bit l set
// when light l is sampled and unoccluded.
struct DiRecord {
    uint32_t id;
    float b1, b2;
    uint32_t vis;
};

struct DiScene {
    const DemDiff* dd;
    const DiRecord* rec;   // rec[s * npix + pixel]
    size_t npix;
    int spp;
    const Light* lights;
    int nl;
    const Material* mat;
};

DemDiff dem_diff_make(int W, int H, const Projection& p, const float* heights, const float* grey_albedo);

// FP64 image (NB per pixel) at heights h (one per grid node) and texels tex.
void di_forward64(const DiScene& s, const double* h, const double* tex, double* img, int nthreads);

// Gradient of sum(g * image) for the image gradient g (NB per pixel): dh per
// grid node, dtex per texel (either may be null).
void di_adjoint64(const DiScene& s, const double* h, const double* tex, const double* g, double* dh,
                  double* dtex);

// Central difference (step hstep) of the L2 loss sum((image - target)^2) with
// respect to the heights of the listed nodes, over the pixels each node touches.
void di_fd_heights64(const DiScene& s, const double* h, const double* tex, const double* target,
                     const int64_t* nodes, size_t n, double hstep, double* out, int nthreads);

}  // namespace vo

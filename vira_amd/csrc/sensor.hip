// sensor.hip -- received power per pixel (row a9, the last step of the render):
// Camera::calculateReceivedPower (camera.ipp:1175-1178)
//   P_k(i, j) = ((eta_k * A) * Omega_ij) * L_k(i, j)
// with Omega_ij the pixel's solid angle (Camera::initializePixelSolidAngle,
// camera.ipp:1594-1666): the four pixel-corner directions K^-1 (x, y, 1) in
// FP64 (the FP32 K^-1 widened, as camera.ipp:1566 does), normalised, and the
// two spherical triangles' areas by Girard's theorem (camera.ipp:1578-1592),
// rounded to FP32.
#include "context.hpp"
#include "shading.hpp"

namespace vg {
namespace {

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 dcross(D3 a, D3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ double ddot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ D3 dnormalize(D3 a) {
    const double s = __ddiv_rn(1.0, __dsqrt_rn(ddot(a, a)));
    return {a.x * s, a.y * s, a.z * s};
}
// unit tangent at p0 of the great circle towards p1
__device__ __forceinline__ D3 tangent(D3 p0, D3 p1) { return dnormalize(dcross(dcross(p0, dsub(p1, p0)), p0)); }

__device__ __forceinline__ double triangle_solid_angle(D3 c0, D3 c1, D3 c2) {
    const double a0 = acos(ddot(tangent(c0, c1), tangent(c0, c2)));
    const double a1 = acos(ddot(tangent(c1, c0), tangent(c1, c2)));
    const double a2 = acos(ddot(tangent(c2, c0), tangent(c2, c1)));
    return ((a0 + a1) + a2) - 3.14159265358979323846;
}

__device__ __forceinline__ D3 corner(const double k[6], double zdir, double x, double y) {
    const double hx = (k[0] * x + k[2] * y) + k[4];
    const double hy = (k[1] * x + k[3] * y) + k[5];
    return dnormalize(D3{hx, zdir * hy, zdir});
}

__global__ void k_received_power(int W, int H, CamF cam, vira_gpu_sensor sensor, const float* __restrict__ radiance,
                                 float* __restrict__ power, float* __restrict__ omega_out) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (size_t)W * H) return;
    const int i = (int)(p % (size_t)W), j = (int)(p / (size_t)W);
    double k[6];
    for (int m = 0; m < 6; ++m) k[m] = (double)cam.kinv[m];
    const double zd = (double)cam.zdir;
    const D3 c0 = corner(k, zd, i, j), c1 = corner(k, zd, i + 1, j);
    const D3 c2 = corner(k, zd, i + 1, j + 1), c3 = corner(k, zd, i, j + 1);
    const float omega = (float)(triangle_solid_angle(c0, c1, c2) + triangle_solid_angle(c0, c2, c3));
    if (omega_out) omega_out[p] = omega;
    if (power) {
#pragma unroll
        for (int b = 0; b < NB; ++b)
            power[NB * p + b] = ((sensor.optical_efficiency[b] * sensor.aperture_area) * omega) * radiance[NB * p + b];
    }
}

}  // namespace

void received_power(Ctx& c, const vira_gpu_camera& cam, const vira_gpu_sensor& s, const float* radiance,
                    float* power, float* omega) {
    VG_CHECK(cam.width > 0 && cam.height > 0, VIRA_GPU_ERR_INVALID, "received_power: empty image");
    VG_CHECK(!power || radiance, VIRA_GPU_ERR_INVALID, "received_power: power needs the radiance image");
    const size_t n = (size_t)cam.width * cam.height;
    CamF cf;
    cf.W = cam.width;
    cf.H = cam.height;
    for (int k = 0; k < 6; ++k) cf.kinv[k] = cam.kinv[k];
    cf.zdir = cam.z_dir;
    k_received_power<<<blocks_for(n, 256), 256, 0, c.stream>>>(cam.width, cam.height, cf, s, radiance, power, omega);
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace vg

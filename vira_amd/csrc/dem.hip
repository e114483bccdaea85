// dem.hip -- synthetic DEM inputs and mesh construction on the GPU.
//
// DEM::makeVertexBuffer (dem.ipp:165-211), imageToIndexBuffer
// (image_utils.ipp:364-412) and Mesh::calculateNormals (mesh.ipp:199-233)
// restated as gather kernels: every output element is produced by one thread
// from its inputs in the reference's summation order, so the results do not
// depend on scheduling.
#include <functional>

#include "context.hpp"

namespace vg {
namespace {

// ------------------------------------------------------------ fBm synthesis
__device__ __forceinline__ double fbm_lattice(long long ix, long long iy, uint32_t octave, uint32_t seed) {
    uint32_t h = lowbias32(seed);
    h = lowbias32(octave ^ h);
    h = lowbias32((uint32_t)iy ^ h);
    h = lowbias32((uint32_t)ix ^ h);
    return (double)(h >> 8) * (1.0 / 8388608.0) - 1.0;
}
__device__ __forceinline__ double fbm_fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
__device__ __forceinline__ double fbm_lerp(double a, double b, double t) { return a + t * (b - a); }

__global__ void k_fbm(int W, int H, uint32_t seed, double amplitude, int octaves, double base_cells,
                      float* out) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (size_t)W * H) return;
    int i = (int)(v % W), j = (int)(v / W);
    double sum = 0.0, amp = 1.0, norm = 0.0;
    double scale = 1.0 / base_cells;
    for (int o = 0; o < octaves; ++o) {
        double x = (double)i * scale, y = (double)j * scale;
        double fx0 = floor(x), fy0 = floor(y);
        long long ix = (long long)fx0, iy = (long long)fy0;
        double fx = x - fx0, fy = y - fy0;
        double v00 = fbm_lattice(ix, iy, (uint32_t)o, seed);
        double v10 = fbm_lattice(ix + 1, iy, (uint32_t)o, seed);
        double v01 = fbm_lattice(ix, iy + 1, (uint32_t)o, seed);
        double v11 = fbm_lattice(ix + 1, iy + 1, (uint32_t)o, seed);
        double sx = fbm_fade(fx), sy = fbm_fade(fy);
        double n = fbm_lerp(fbm_lerp(v00, v10, sx), fbm_lerp(v01, v11, sx), sy);
        sum = sum + amp * n;
        norm = norm + amp;
        amp = amp * 0.5;
        scale = scale * 2.0;
    }
    out[v] = (float)(amplitude * __ddiv_rn(sum, norm));
}

__global__ void k_vertex_albedo(size_t nv, uint32_t seed, float* out) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    uint32_t h = lowbias32((uint32_t)v ^ lowbias32(seed ^ 0xa5a5a5a5U));
    double u = (double)(h >> 8) * (1.0 / 16777216.0);
    out[v] = (float)(0.12 + 0.02 * (2.0 * u - 1.0));
}

// ------------------------------------------------------------ tessellation
// Candidate triangle c = 2*(i*(H-1) + j) + {0: A, 1: B} in imageToIndexBuffer order.
__device__ __forceinline__ void cand_vertices(uint32_t c, int W, int H, uint32_t v[3]) {
    uint32_t q = c >> 1;
    uint32_t i = q / (uint32_t)(H - 1), j = q % (uint32_t)(H - 1);
    if ((c & 1) == 0) {
        v[0] = i + (j + 1) * (uint32_t)W; v[1] = i + 1 + j * (uint32_t)W; v[2] = i + j * (uint32_t)W;
    } else {
        v[0] = i + 1 + j * (uint32_t)W; v[1] = i + (j + 1) * (uint32_t)W; v[2] = i + 1 + (j + 1) * (uint32_t)W;
    }
}

__global__ void k_cand_valid(const float* heights, int W, int H, size_t ncand, uint32_t* flag) {
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncand) return;
    uint32_t v[3];
    cand_vertices((uint32_t)c, W, H, v);
    flag[c] = (isfinite(heights[v[0]]) && isfinite(heights[v[1]]) && isfinite(heights[v[2]])) ? 1u : 0u;
}

__global__ void k_cand_write(const float* heights, int W, int H, size_t ncand, const uint32_t* flag,
                             const uint32_t* ids, uint32_t* idx, uint32_t* cand_to_tri) {
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncand) return;
    if (!flag[c]) { cand_to_tri[c] = 0xffffffffu; return; }
    uint32_t v[3];
    cand_vertices((uint32_t)c, W, H, v);
    uint32_t t = ids[c];
    idx[3 * (size_t)t] = v[0];
    idx[3 * (size_t)t + 1] = v[1];
    idx[3 * (size_t)t + 2] = v[2];
    cand_to_tri[c] = t;
}

// FP64 world position (dem_world of the oracle) minus the rebasing offset.
__device__ __forceinline__ void dem_world(const vira_gpu_projection& p, int W, int H, int i, int j, double h,
                                          double out[3]) {
    double cx = (double)(W - 1) * 0.5, cy = (double)(H - 1) * 0.5;
    if (p.kind == 0) {
        out[0] = ((double)i - cx) * p.spacing;
        out[1] = (cy - (double)j) * p.spacing;
        out[2] = h;
    } else {
        double a = __ddiv_rn(p.spacing, p.radius);
        double dx = 1.0, dy = ((double)i - cx) * a, dz = (cy - (double)j) * a;
        double inv = __ddiv_rn(1.0, __dsqrt_rn((dx * dx + dy * dy) + dz * dz));
        double r = p.radius + h;
        double s = r * inv;
        out[0] = p.center[0] + dx * s;
        out[1] = p.center[1] + dy * s;
        out[2] = p.center[2] + dz * s;
    }
    out[0] = out[0] - p.offset[0];
    out[1] = out[1] - p.offset[1];
    out[2] = out[2] - p.offset[2];
}

__global__ void k_dem_vertices(const float* heights, int W, int H, vira_gpu_projection proj, const float* grey,
                               GVertex* verts, double* pos64) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (size_t)W * H) return;
    int i = (int)(v % W), j = (int)(v / W);
    float h = heights[v];
    GVertex out;
    double p[3] = {__longlong_as_double(0x7ff0000000000000ll), 0.0, 0.0};
    if (!isfinite(h)) {
        out.px = out.py = out.pz = __int_as_float(0x7f800000);
    } else {
        dem_world(proj, W, H, i, j, (double)h, p);
        out.px = (float)p[0];
        out.py = (float)p[1];
        out.pz = (float)p[2];
    }
    pos64[3 * v] = p[0];
    pos64[3 * v + 1] = p[1];
    pos64[3 * v + 2] = p[2];
    float a = grey ? grey[v] : 1.0f;
    for (int k = 0; k < NB; ++k) out.alb[k] = a;
    out.nx = out.ny = out.nz = 0.0f;
    out.u = (float)__ddiv_rn((double)i, (double)(W - 1));
    out.v = (float)__ddiv_rn((double)j, (double)(H - 1));
    verts[v] = out;
}

__device__ __forceinline__ void finish_normal(GVertex& out, f3 n) {
    if (length(n) != 0.0f) n = normalize(n);
    out.nx = n.x;
    out.ny = n.y;
    out.nz = n.z;
}

// Grid normals: the (up to) six triangles around vertex (x, y) in increasing
// triangle-id order: B(x-1,y-1), A(x-1,y), B(x-1,y), A(x,y-1), B(x,y-1), A(x,y).
// Threads walk the vertices column by column (y fastest), the order of the
// quad table and of the triangle ids, so neighbouring threads read
// neighbouring entries of both.
__global__ void k_dem_normals(int W, int H, const uint32_t* cand_to_tri, const float4* fnrm, GVertex* verts) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)W * H) return;
    const int y = (int)(t % H), x = (int)(t / H);
    const size_t v = (size_t)y * W + x;
    const int dqx[6] = {-1, -1, -1, 0, 0, 0};
    const int dqy[6] = {-1, 0, 0, -1, -1, 0};
    const int ab[6] = {1, 0, 1, 0, 1, 0};
    f3 n = mk3(0.0f, 0.0f, 0.0f);
    for (int k = 0; k < 6; ++k) {
        int qi = x + dqx[k], qj = y + dqy[k];
        if (qi < 0 || qj < 0 || qi >= W - 1 || qj >= H - 1) continue;
        size_t c = 2 * ((size_t)qi * (size_t)(H - 1) + (size_t)qj) + (size_t)ab[k];
        uint32_t t = cand_to_tri[c];
        if (t == 0xffffffffu) continue;
        const float4 fn = fnrm[t];
        if (fn.w == 0.0f) continue;
        n = add(n, mk3(fn.x, fn.y, fn.z));
    }
    finish_normal(verts[v], n);
}

// ------------------------------------------------------------ generic meshes
__global__ void k_pack_vertices(size_t nv, const float* pos, const float* alb, const float* uv, const float* nrm,
                                GVertex* verts) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    GVertex o;
    o.px = pos[3 * v]; o.py = pos[3 * v + 1]; o.pz = pos[3 * v + 2];
    for (int k = 0; k < NB; ++k) o.alb[k] = alb ? alb[NB * v + k] : 1.0f;
    o.nx = nrm ? nrm[3 * v] : 0.0f;
    o.ny = nrm ? nrm[3 * v + 1] : 0.0f;
    o.nz = nrm ? nrm[3 * v + 2] : 0.0f;
    o.u = uv ? uv[2 * v] : 0.0f;
    o.v = uv ? uv[2 * v + 1] : 0.0f;
    verts[v] = o;
}

__global__ void k_adj_count(size_t nt, const uint32_t* idx, uint32_t* count) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    for (int k = 0; k < 3; ++k) atomicAdd(&count[idx[3 * t + k]], 1u);
}

__global__ void k_adj_fill(size_t nt, const uint32_t* idx, const uint32_t* offs, uint32_t* cursor, uint32_t* adj) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    for (int k = 0; k < 3; ++k) {
        uint32_t v = idx[3 * t + k];
        uint32_t slot = atomicAdd(&cursor[v], 1u);
        adj[offs[v] + slot] = (uint32_t)t;
    }
}

// Sorts each vertex's incident triangles by id (insertion sort: lists are short).
__global__ void k_adj_sort(size_t nv, const uint32_t* offs, const uint32_t* count, uint32_t* adj) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    uint32_t b = offs[v], n = count[v];
    for (uint32_t a = 1; a < n; ++a) {
        uint32_t key = adj[b + a];
        uint32_t k = a;
        while (k > 0 && adj[b + k - 1] > key) {
            adj[b + k] = adj[b + k - 1];
            --k;
        }
        adj[b + k] = key;
    }
}

// Sums each vertex's face normals in triangle order (sorted adjacency).
__global__ void k_adj_normals(size_t nv, const uint32_t* offs, const uint32_t* count, const uint32_t* adj,
                              const float4* fnrm, GVertex* verts) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    uint32_t b = offs[v], n = count[v];
    f3 acc = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t a = 0; a < n; ++a) {
        const float4 fn = fnrm[adj[b + a]];
        if (fn.w == 0.0f) continue;
        acc = add(acc, mk3(fn.x, fn.y, fn.z));
    }
    finish_normal(verts[v], acc);
}

// Face normals from FP64 positions (pos64, 3 per vertex).
__global__ void k_face_normals64(size_t nt, const uint32_t* __restrict__ idx, const double* __restrict__ pos64,
                                 float4* __restrict__ fnrm) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint32_t i0 = idx[3 * t], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
    fnrm[t] = face_normal64(pos64 + 3 * (size_t)i0, pos64 + 3 * (size_t)i1, pos64 + 3 * (size_t)i2);
}

// Face normals of a mesh given in FP32 (its positions widened to FP64).
__global__ void k_face_normals_f32(size_t nt, const uint32_t* __restrict__ idx, const GVertex* __restrict__ verts,
                                   float4* __restrict__ fnrm) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double q[3][3];
    for (int k = 0; k < 3; ++k) {
        const GVertex& V = verts[idx[3 * t + k]];
        q[k][0] = V.px;
        q[k][1] = V.py;
        q[k][2] = V.pz;
    }
    fnrm[t] = face_normal64(q[0], q[1], q[2]);
}

// DEM node positions for new heights (uv, albedo and normals untouched), and
// their FP64 values for the face normals.
__global__ void k_dem_positions(const float* heights, int W, int H, vira_gpu_projection proj, GVertex* verts,
                                double* pos64) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (size_t)W * H) return;
    float h = heights[v];
    double p[3] = {__longlong_as_double(0x7ff0000000000000ll), 0.0, 0.0};
    if (isfinite(h)) {
        dem_world(proj, W, H, (int)(v % W), (int)(v / W), (double)h, p);
        verts[v].px = (float)p[0];
        verts[v].py = (float)p[1];
        verts[v].pz = (float)p[2];
    }
    pos64[3 * v] = p[0];
    pos64[3 * v + 1] = p[1];
    pos64[3 * v + 2] = p[2];
}

// Counts nodes whose height changed between finite and non-finite (the
// triangle set of a DEM mesh is fixed once built).
__global__ void k_dem_mask_diff(const float* heights, size_t nv, const GVertex* verts, uint32_t* count) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    if ((bool)isfinite(heights[v]) == (bool)isinf(verts[v].px)) atomicAdd(count, 1u);
}

}  // namespace

void launch_synth_fbm(hipStream_t s, int W, int H, uint32_t seed, double amp, int oct, double base, float* out) {
    size_t n = (size_t)W * H;
    k_fbm<<<blocks_for(n, 256), 256, 0, s>>>(W, H, seed, amp, oct, base, out);
    VG_HIP(hipGetLastError());
}

void launch_synth_albedo(hipStream_t s, size_t nv, uint32_t seed, float* out) {
    k_vertex_albedo<<<blocks_for(nv, 256), 256, 0, s>>>(nv, seed, out);
    VG_HIP(hipGetLastError());
}

void dem_grid_mesh(Ctx& c, MeshG& m, const float* heights, int W, int H,
                   const std::function<void(GVertex*, double*)>& write_vertices) {
    VG_CHECK(W >= 2 && H >= 2, VIRA_GPU_ERR_INVALID, "DEM needs at least 2x2 samples");
    VG_CHECK((size_t)W * H < (1ull << 31), VIRA_GPU_ERR_INVALID, "DEM too large for 32-bit vertex ids");
    size_t nv = (size_t)W * H;
    size_t ncand = 2 * (size_t)(W - 1) * (size_t)(H - 1);
    VG_CHECK(ncand < (1ull << 27), VIRA_GPU_ERR_INVALID, "DEM exceeds 2^27 triangles");
    DBuf flag, ids, c2t;
    flag.alloc(ncand * 4);
    ids.alloc(ncand * 4);
    c2t.alloc(ncand * 4);
    k_cand_valid<<<blocks_for(ncand, 256), 256, 0, c.stream>>>(heights, W, H, ncand, flag.as<uint32_t>());
    VG_HIP(hipGetLastError());
    uint32_t nt = scan_exclusive_u32(c, flag.as<uint32_t>(), ids.as<uint32_t>(), ncand);
    m.nv = nv;
    m.nt = nt;
    m.grid_w = W;
    m.grid_h = H;
    m.verts.alloc(nv * sizeof(GVertex));
    m.idx.alloc(std::max<size_t>(1, (size_t)nt * 3 * 4));
    k_cand_write<<<blocks_for(ncand, 256), 256, 0, c.stream>>>(heights, W, H, ncand, flag.as<uint32_t>(),
                                                                ids.as<uint32_t>(), m.idx.as<uint32_t>(),
                                                                c2t.as<uint32_t>());
    VG_HIP(hipGetLastError());
    DBuf pos64;
    pos64.alloc(nv * 3 * sizeof(double));
    write_vertices(m.verts.as<GVertex>(), pos64.as<double>());
    m.fnrm.alloc(std::max<size_t>(1, (size_t)nt) * sizeof(float4));
    if (nt) {
        k_face_normals64<<<blocks_for(nt, 256), 256, 0, c.stream>>>(nt, m.idx.as<uint32_t>(), pos64.as<double>(),
                                                                    m.fnrm.as<float4>());
        VG_HIP(hipGetLastError());
    }
    k_dem_normals<<<blocks_for(nv, 256), 256, 0, c.stream>>>(W, H, c2t.as<uint32_t>(), m.fnrm.as<float4>(),
                                                              m.verts.as<GVertex>());
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(c.stream));
}

void dem_build_mesh(Ctx& c, MeshG& m, const float* heights, int W, int H, const vira_gpu_projection& p,
                    const float* grey) {
    dem_grid_mesh(c, m, heights, W, H, [&](GVertex* verts, double* pos64) {
        k_dem_vertices<<<blocks_for((size_t)W * H, 256), 256, 0, c.stream>>>(heights, W, H, p, grey, verts, pos64);
        VG_HIP(hipGetLastError());
    });
    m.is_dem = true;
    m.dem_w = W;
    m.dem_h = H;
    m.dem_proj = p;
}

void dem_update(Ctx& c, MeshG& m, const float* heights) {
    VG_CHECK(m.is_dem, VIRA_GPU_ERR_INVALID, "dem_update: mesh was not made by dem_mesh");
    const size_t nv = m.nv;
    DBuf cnt;
    cnt.alloc(4);
    VG_HIP(hipMemsetAsync(cnt.p, 0, 4, c.stream));
    k_dem_mask_diff<<<blocks_for(nv, 256), 256, 0, c.stream>>>(heights, nv, m.verts.as<GVertex>(),
                                                                cnt.as<uint32_t>());
    VG_HIP(hipGetLastError());
    uint32_t bad = 0;
    VG_HIP(hipMemcpyAsync(&bad, cnt.p, 4, hipMemcpyDeviceToHost, c.stream));
    VG_HIP(hipStreamSynchronize(c.stream));
    VG_CHECK(bad == 0, VIRA_GPU_ERR_INVALID,
             "dem_update: " + std::to_string(bad) + " heights changed between finite and non-finite");
    DBuf pos64;
    pos64.alloc(nv * 3 * sizeof(double));
    k_dem_positions<<<blocks_for(nv, 256), 256, 0, c.stream>>>(heights, m.dem_w, m.dem_h, m.dem_proj,
                                                                m.verts.as<GVertex>(), pos64.as<double>());
    VG_HIP(hipGetLastError());
    if (m.nt) {
        k_face_normals64<<<blocks_for(m.nt, 256), 256, 0, c.stream>>>(m.nt, m.idx.as<uint32_t>(), pos64.as<double>(),
                                                                      m.fnrm.as<float4>());
        VG_HIP(hipGetLastError());
    }
    mesh_adjacency(c, m);
    k_adj_normals<<<blocks_for(nv, 256), 256, 0, c.stream>>>(nv, m.adj_offs.as<uint32_t>(), m.adj_count.as<uint32_t>(),
                                                              m.adj.as<uint32_t>(), m.fnrm.as<float4>(),
                                                              m.verts.as<GVertex>());
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(c.stream));
}

void mesh_upload_pack(Ctx& c, MeshG& m, const float* pos, const float* alb, const float* uv, const float* nrm) {
    k_pack_vertices<<<blocks_for(m.nv, 256), 256, 0, c.stream>>>(m.nv, pos, alb, uv, nrm, m.verts.as<GVertex>());
    VG_HIP(hipGetLastError());
    m.fnrm.alloc(std::max<size_t>(1, m.nt) * sizeof(float4));
    if (m.nt) {
        k_face_normals_f32<<<blocks_for(m.nt, 256), 256, 0, c.stream>>>(m.nt, m.idx.as<uint32_t>(),
                                                                        m.verts.as<GVertex>(), m.fnrm.as<float4>());
        VG_HIP(hipGetLastError());
    }
}

void mesh_adjacency(Ctx& c, MeshG& m) {
    if (m.adj_built || m.nv == 0) return;
    m.adj_count.alloc(m.nv * 4);
    m.adj_offs.alloc(m.nv * 4);
    m.adj.alloc(std::max<size_t>(1, m.nt * 3 * 4));
    DBuf cursor;
    cursor.alloc(m.nv * 4);
    VG_HIP(hipMemsetAsync(m.adj_count.p, 0, m.nv * 4, c.stream));
    VG_HIP(hipMemsetAsync(cursor.p, 0, m.nv * 4, c.stream));
    if (m.nt) {
        k_adj_count<<<blocks_for(m.nt, 256), 256, 0, c.stream>>>(m.nt, m.idx.as<uint32_t>(),
                                                                  m.adj_count.as<uint32_t>());
        VG_HIP(hipGetLastError());
    }
    scan_exclusive_u32(c, m.adj_count.as<uint32_t>(), m.adj_offs.as<uint32_t>(), m.nv);
    if (m.nt) {
        k_adj_fill<<<blocks_for(m.nt, 256), 256, 0, c.stream>>>(m.nt, m.idx.as<uint32_t>(), m.adj_offs.as<uint32_t>(),
                                                                 cursor.as<uint32_t>(), m.adj.as<uint32_t>());
        VG_HIP(hipGetLastError());
    }
    k_adj_sort<<<blocks_for(m.nv, 256), 256, 0, c.stream>>>(m.nv, m.adj_offs.as<uint32_t>(),
                                                             m.adj_count.as<uint32_t>(), m.adj.as<uint32_t>());
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(c.stream));
    m.adj_built = true;
}

void mesh_compute_normals(Ctx& c, MeshG& m) {
    if (m.nv == 0) return;
    mesh_adjacency(c, m);
    k_adj_normals<<<blocks_for(m.nv, 256), 256, 0, c.stream>>>(m.nv, m.adj_offs.as<uint32_t>(),
                                                                m.adj_count.as<uint32_t>(), m.adj.as<uint32_t>(),
                                                                m.fnrm.as<float4>(), m.verts.as<GVertex>());
    VG_HIP(hipGetLastError());
    VG_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace vg

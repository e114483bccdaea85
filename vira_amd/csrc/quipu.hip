// quipu.hip -- the quipu DEM pyramid reader (.qld) feeding the GPU mesh and
// BVH builders.
//
// DEMQuipu (dem_quipu.hpp:34-81, dem_quipu.ipp:29-236) with the field codecs
// of quipu_io.ipp: the header (identifier, class id, transformation, level
// count, albedo flag, normal cone), the (GSD, offset) table, and per level the
// DEM projection, the quantised heights and the albedo.  DEMQuipu::readBuffers
// picks the coarsest level whose GSD is within the request
// (dem_quipu.ipp:137-151).
//
// Work split: the file is parsed and its LZ4 blocks decoded on the host (one
// block per image: LZ4's back-references make a single block a serial chain,
// and the format has no chunking to decode in parallel), the quantised
// samples go to the device as they are stored (1 or 2 bytes each), and the
// GPU dequantises them (fixedToFloat, image_utils.ipp:508-527), places every
// vertex with the level's projection (DEMProjection::transformToWorld,
// dem_projection.ipp:102-144), attaches the albedo, and builds the triangles
// and normals with the DEM tessellation of dem.hip.
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "context.hpp"
#include "lz4_block.hpp"

namespace vg {
namespace {

enum : uint16_t {
    kIdFloat = 1, kIdDouble = 2, kIdU8 = 3, kIdU16 = 4, kIdU32 = 5, kIdU64 = 6, kIdDemPyramid = 2002,
};

// Bounds-checked little-endian reader over the file, mapped read-only: image
// blocks are decoded straight from the mapping, with no copy of the file.
struct Rd {
    std::string path;
    int fd = -1;
    const uint8_t* base = nullptr;
    uint64_t size = 0, at = 0;
    explicit Rd(const std::string& p) : path(p) {
        fd = ::open(p.c_str(), O_RDONLY);
        VG_CHECK(fd >= 0, VIRA_GPU_ERR_INVALID, "quipu: cannot open " + p);
        struct stat st;
        if (::fstat(fd, &st) != 0) {
            ::close(fd);
            throw Error(VIRA_GPU_ERR_INVALID, "quipu: cannot stat " + p);
        }
        size = (uint64_t)st.st_size;
        if (size > 0) {
            void* m = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
            if (m == MAP_FAILED) {
                ::close(fd);
                throw Error(VIRA_GPU_ERR_INVALID, "quipu: cannot map " + p);
            }
            base = static_cast<const uint8_t*>(m);
        }
    }
    Rd(const Rd&) = delete;
    Rd& operator=(const Rd&) = delete;
    ~Rd() {
        if (base) ::munmap(const_cast<uint8_t*>(base), size);
        if (fd >= 0) ::close(fd);
    }
    uint64_t pos() const { return at; }
    void seek(uint64_t to) {
        VG_CHECK(to <= size, VIRA_GPU_ERR_INVALID, "quipu: offset beyond the end of " + path);
        at = to;
    }
    // the next n bytes of the file (checked), and the position moved past them
    const uint8_t* take(uint64_t n) {
        VG_CHECK(n <= size - at, VIRA_GPU_ERR_INVALID, "quipu: truncated file " + path);
        const uint8_t* p = base + at;
        at += n;
        return p;
    }
    void bytes(void* dst, uint64_t n) {
        const uint8_t* p = take(n);
        if (n) std::memcpy(dst, p, n);
    }
    template <class T>
    T get() {
        T v;
        bytes(&v, sizeof(T));
        return v;
    }
};

// readVec / readMat (quipu_io.ipp:343-480): class id, sizes, values as float or
// double (integer element types are read as the reference allows for vectors).
double read_elem(Rd& r, uint16_t id) {
    switch (id) {
        case kIdFloat: return (double)r.get<float>();
        case kIdDouble: return r.get<double>();
        case kIdU8: return (double)r.get<uint8_t>();
        case kIdU16: return (double)r.get<uint16_t>();
        case kIdU32: return (double)r.get<uint64_t>();  // the reference reads VIRA_UINT32 vectors as size_t
        case kIdU64: return (double)r.get<uint64_t>();
        default: throw Error(VIRA_GPU_ERR_INVALID, "quipu: vector with class id " + std::to_string(id));
    }
}
void read_vec(Rd& r, int n, double* out) {
    const uint16_t id = r.get<uint16_t>();
    const int nl = r.get<uint8_t>();
    VG_CHECK(nl == n, VIRA_GPU_ERR_INVALID,
             "quipu: vector of length " + std::to_string(nl) + " where " + std::to_string(n) + " is expected");
    for (int k = 0; k < n; ++k) out[k] = read_elem(r, id);
}
void read_mat3(Rd& r, double m[3][3]) {  // m[column][row], GLM order
    const uint16_t id = r.get<uint16_t>();
    const int n = r.get<uint8_t>(), mm = r.get<uint8_t>();
    VG_CHECK(n == 3 && mm == 3, VIRA_GPU_ERR_INVALID, "quipu: rotation is not 3x3");
    VG_CHECK(id == kIdFloat || id == kIdDouble, VIRA_GPU_ERR_INVALID, "quipu: matrix with class id " + std::to_string(id));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m[i][j] = read_elem(r, id);
}

// readTransformation + Rotation(mat3) + makeTransformationMatrix
// (quipu_io.ipp:503-519, rotation.ipp:15-24, reference_frame.ipp:488-500).
void read_transformation(Rd& r, double M[16]) {
    double p[3], s[3], R[3][3];
    read_vec(r, 3, p);
    read_mat3(r, R);
    read_vec(r, 3, s);
    const double det = R[0][0] * (R[1][1] * R[2][2] - R[2][1] * R[1][2]) -
                       R[1][0] * (R[0][1] * R[2][2] - R[2][1] * R[0][2]) +
                       R[2][0] * (R[0][1] * R[1][2] - R[1][1] * R[0][2]);
    VG_CHECK(std::fabs(det - 1.0) <= 1e-3, VIRA_GPU_ERR_INVALID,
             "quipu: the stored rotation is not a rotation (determinant " + std::to_string(det) + ")");
    for (int c = 0; c < 3; ++c) {
        for (int k = 0; k < 3; ++k) M[4 * c + k] = R[c][k] * s[c];
        M[4 * c + 3] = 0.0;
    }
    M[12] = p[0];
    M[13] = p[1];
    M[14] = p[2];
    M[15] = 1.0;
}

// DEMProjection fields (quipu_io.ipp:606-628).
struct Proj {
    uint32_t res[2];
    double pixel_scale[2];
    uint32_t tiles[2], tile_row, tile_col;
    double xoff, yoff, dx, dy;
    bool ogr;
    std::string wkt;
    double rot[2], tie[2];
};
Proj read_projection(Rd& r) {
    Proj p{};
    p.res[0] = r.get<uint32_t>();
    p.res[1] = r.get<uint32_t>();
    p.pixel_scale[0] = r.get<double>();
    p.pixel_scale[1] = r.get<double>();
    p.tiles[0] = r.get<uint32_t>();
    p.tiles[1] = r.get<uint32_t>();
    p.tile_row = r.get<uint32_t>();
    p.tile_col = r.get<uint32_t>();
    p.xoff = r.get<double>();
    p.yoff = r.get<double>();
    p.dx = r.get<double>();
    p.dy = r.get<double>();
    p.ogr = r.get<uint8_t>() != 0;
    if (p.ogr) {
        const uint32_t n = r.get<uint32_t>();
        VG_CHECK(n < (1u << 20), VIRA_GPU_ERR_INVALID, "quipu: projection string too long");
        p.wkt.resize(n);
        if (n) r.bytes(&p.wkt[0], n);
        p.rot[0] = r.get<double>();
        p.rot[1] = r.get<double>();
        p.tie[0] = r.get<double>();
        p.tie[1] = r.get<double>();
    }
    return p;
}

// One stored image (writeImage/readImage, quipu_io.ipp:744-811): class id,
// resolution, compressed flag, then either size_t bytes + uint64 block size +
// one LZ4 block of the pixels in memory order (i + j*W), or the raw pixels
// written with i outer and j inner.  The pixels are decoded in memory order
// into host memory that `dest(bytes)` returns (page-locked staging, so their
// copy to the device is one DMA).
struct Img {
    uint32_t w = 0, h = 0;
    int elem = 0;
    bool compressed = false;
    uint8_t* raw = nullptr;  // w * h * elem bytes, owned by the caller of read_image
    size_t bytes() const { return (size_t)w * h * (size_t)elem; }
};
template <class Dest>
Img read_image(Rd& r, uint16_t expect_id, int elem, double& decode_ms, Dest&& dest) {
    Img im;
    const uint16_t id = r.get<uint16_t>();
    VG_CHECK(id == expect_id, VIRA_GPU_ERR_INVALID,
             "quipu: image with class id " + std::to_string(id) + " where " + std::to_string(expect_id) + " is expected");
    im.w = r.get<uint32_t>();
    im.h = r.get<uint32_t>();
    im.elem = elem;
    im.compressed = r.get<uint8_t>() != 0;
    VG_CHECK((uint64_t)im.w * im.h < (1ull << 31), VIRA_GPU_ERR_INVALID, "quipu: image of 2^31 pixels or more");
    const uint64_t n = (uint64_t)im.w * im.h * (uint64_t)elem;
    const uint64_t left = r.size - std::min(r.size, r.pos());
    if (im.compressed) {
        const uint64_t stated = r.get<uint64_t>();
        VG_CHECK(stated == n, VIRA_GPU_ERR_INVALID, "quipu: compressed image size disagrees with its resolution");
        const uint64_t cn = r.get<uint64_t>();
        VG_CHECK(cn <= left, VIRA_GPU_ERR_INVALID, "quipu: truncated LZ4 block");
        // an LZ4 block expands at most ~255-fold (a match length byte of 255 per input byte)
        VG_CHECK(n <= 255 * cn + 16, VIRA_GPU_ERR_INVALID, "quipu: image larger than its LZ4 block can hold");
        const uint8_t* comp = r.take(cn);
        im.raw = dest(n);
        const auto t0 = std::chrono::steady_clock::now();
        const long long got = lz4_decode_block(comp, cn, im.raw, n);
        decode_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        VG_CHECK(got == (long long)n, VIRA_GPU_ERR_INVALID, "quipu: LZ4 block failed to decompress");
    } else {
        VG_CHECK(n <= left, VIRA_GPU_ERR_INVALID, "quipu: image larger than the rest of its file");
        // stored with i outer, j inner: transposed into memory order
        const uint8_t* src = r.take(n);
        im.raw = dest(n);
        for (uint64_t i = 0; i < im.w; ++i)
            for (uint64_t j = 0; j < im.h; ++j)
                std::memcpy(&im.raw[(i + j * im.w) * elem], &src[(i * im.h + j) * elem], elem);
    }
    return im;
}

// ------------------------------------------------------------- device side
struct QProj {
    int kind;  // 0 planar (non-OGR), 1 equirectangular (OGR), 2 polar stereographic (OGR, sphere)
    double pixel_scale[2], xoff, yoff, dx, dy;
    double start_x, start_y, scale_x, scale_y;  // DEMProjection::initToWorld
    double rot[2], tie[2];
    double a, b;                                // ellipsoid semi-axes
    double lon0, lat_ts, lat0, fe, fn;          // equirectangular parameters (degrees, metres)
    double pole, akm1;                          // polar stereographic: +1 north / -1 south, 2 k0
    double offset[3];
};

// fixedToFloat (image_utils.ipp:508-527): 0 is "no data" (+inf), else
// ((max - min) * (v / (m - 1))) + min in FP32.
template <class T>
__global__ void k_fixed_to_float(const T* __restrict__ q, size_t n, float lo, float hi, float* __restrict__ out) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const T v = q[k];
    if (v == 0) {
        out[k] = __int_as_float(0x7f800000);
    } else {
        const float m = (float)(T)~(T)0;
        float f = __fdiv_rn((float)v, m - 1.0f);
        out[k] = ((hi - lo) * f) + lo;
    }
}

// DEMProjection::transformToWorld (dem_projection.ipp:102-144) in FP64; the
// equirectangular inverse is the spherical one of PROJ's eqc
// (x = a (lon - lon0) cos(lat_ts), y = a (lat - lat0)), then
// utils::ellipsoidToCartesian (utils.ipp:291-313) with c = a.
__device__ void q_world(const QProj& p, int i, int j, double z, double out[3]) {
    if (p.kind == 0) {
        const double Ix = (double)i + __ddiv_rn(p.xoff, p.dx);
        const double Iy = (double)j + __ddiv_rn(p.yoff, p.dy);
        const double xs = Ix * p.scale_x, ys = -Iy * p.scale_y;
        out[0] = p.dx * p.pixel_scale[0] * (p.start_x + (2.0 * xs));
        out[1] = p.dy * p.pixel_scale[1] * (-p.start_y + (2.0 * ys));
        out[2] = z;
        return;
    }
    const double Ix = (p.dx * (double)i) + p.xoff + 0.5;
    const double Iy = (p.dy * (double)j) + p.yoff + 0.5;
    const double X = p.tie[0] + (Ix * p.pixel_scale[0]) + (Iy * p.rot[0]);
    const double Y = p.tie[1] + (Ix * p.rot[1]) + (Iy * p.pixel_scale[1]);
    const double d2r = 3.14159265358979323846 / 180.0;
    double lon, lat;
    if (p.kind == 2) {
        // PROJ's spherical polar stere inverse: (x, y) / a, rh = hypot,
        // c = 2 atan(rh / 2k0), phi = +-asin(cos c), lam = atan2(x, -+y)
        const double x = __ddiv_rn(X - p.fe, p.a), y = __ddiv_rn(Y - p.fn, p.a);
        const double rh = hypot(x, y);
        const double cosc = cos(2.0 * atan(__ddiv_rn(rh, p.akm1)));
        if (rh <= 1e-10) {
            lat = 90.0 * p.pole;
            lon = p.lon0;
        } else {
            lat = asin(p.pole * cosc) / d2r;
            lon = p.lon0 + atan2(x, -p.pole * y) / d2r;
        }
    } else {
        lon = p.lon0 + __ddiv_rn(X - p.fe, p.a * cos(p.lat_ts * d2r)) / d2r;
        lat = p.lat0 + __ddiv_rn(Y - p.fn, p.a) / d2r;
    }
    const double ee2 = __ddiv_rn(p.a * p.a - p.b * p.b, p.a * p.a);
    const double clat = cos(lat * d2r), slat = sin(lat * d2r), clon = cos(lon * d2r), slon = sin(lon * d2r);
    const double V = __ddiv_rn(p.a, __dsqrt_rn(1.0 - ee2 * (clat * clat) * (slon * slon)));
    out[0] = ((V + z) * clon * clat) - p.offset[0];
    out[1] = ((V * (1.0 - ee2) + z) * slon * clat) - p.offset[1];
    out[2] = ((V + z) * slat) - p.offset[2];
}

// albedo: 0 none (default value), 1 constant spectrum, 2 grey image x
// profile, 3 eight grey channels
__global__ void k_qld_vertices(const float* __restrict__ heights, int W, int H, QProj p, int albedo_kind,
                               const float* __restrict__ albedo, vira_gpu_qld_read_opts o, GVertex* __restrict__ verts,
                               double* __restrict__ pos64) {
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (size_t)W * H) return;
    const int i = (int)(v % (size_t)W), j = (int)(v / (size_t)W);
    const float h = heights[v];
    GVertex out;
    for (int k = 0; k < NB; ++k) out.alb[k] = 0.0f;
    double w[3] = {__longlong_as_double(0x7ff0000000000000ll), 0.0, 0.0};
    if (!isfinite(h)) {
        out.px = out.py = out.pz = __int_as_float(0x7f800000);
    } else {
        q_world(p, i, j, (double)h, w);
        out.px = (float)w[0];
        out.py = (float)w[1];
        out.pz = (float)w[2];
        for (int k = 0; k < NB; ++k) {
            float a = o.default_albedo;
            if (albedo_kind == 1) a = albedo[k];
            else if (albedo_kind == 2) a = o.albedo_profile[k] * albedo[v];
            else if (albedo_kind == 3) a = albedo[(size_t)k * W * H + v];
            out.alb[k] = a;
        }
    }
    out.nx = out.ny = out.nz = 0.0f;
    out.u = (float)__ddiv_rn((double)i, (double)(W - 1));
    out.v = (float)__ddiv_rn((double)j, (double)(H - 1));
    verts[v] = out;
    pos64[3 * v] = w[0];
    pos64[3 * v + 1] = w[1];
    pos64[3 * v + 2] = w[2];
}

// A tolerant scan of a WKT1 projection string: PROJECTION["name"] and
// PARAMETER["name", value] (case-insensitive names) and the SPHEROID axes.
std::string lower(std::string s) {
    for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}
bool wkt_param(const std::string& w, const char* name, double& v) {
    const std::string lw = lower(w), key = std::string("parameter[\"") + name + "\"";
    const size_t at = lw.find(key);
    if (at == std::string::npos) return false;
    const size_t comma = w.find(',', at + key.size());
    if (comma == std::string::npos) return false;
    v = std::strtod(w.c_str() + comma + 1, nullptr);
    return true;
}
QProj make_qproj(const Proj& P, const double offset[3]) {
    QProj q{};
    q.pixel_scale[0] = P.pixel_scale[0];
    q.pixel_scale[1] = P.pixel_scale[1];
    q.xoff = P.xoff;
    q.yoff = P.yoff;
    q.dx = P.dx;
    q.dy = P.dy;
    for (int k = 0; k < 3; ++k) q.offset[k] = offset[k];
    if (!P.ogr) {
        q.kind = 0;
        const double tx = ((double)P.res[0] - 1.) / 2., ty = ((double)P.res[1] - 1.) / 2.;
        q.start_x = -tx;
        q.start_y = -ty;
        q.scale_x = (tx + 0.5) / (double)P.res[0];
        q.scale_y = (ty + 0.5) / (double)P.res[1];
        return q;
    }
    q.kind = 1;
    q.rot[0] = P.rot[0];
    q.rot[1] = P.rot[1];
    q.tie[0] = P.tie[0];
    q.tie[1] = P.tie[1];
    const std::string lw = lower(P.wkt);
    const size_t pj = lw.find("projection[\"");
    VG_CHECK(pj != std::string::npos, VIRA_GPU_ERR_UNSUPPORTED, "quipu: projection string has no PROJECTION");
    const size_t e = lw.find('"', pj + 12);
    const std::string name = lw.substr(pj + 12, e - (pj + 12));
    const bool eqc = name == "equirectangular" || name == "equidistant_cylindrical" || name == "plate_carree";
    const bool ps_n = name == "stereographic_north_pole", ps_s = name == "stereographic_south_pole";
    const bool ps = name == "polar_stereographic" || ps_n || ps_s;
    VG_CHECK(eqc || ps, VIRA_GPU_ERR_UNSUPPORTED,
             "quipu: OGR projection '" + name + "' is not supported (equirectangular and polar stereographic only)");
    const size_t sp = lw.find("spheroid[\"");
    VG_CHECK(sp != std::string::npos, VIRA_GPU_ERR_INVALID, "quipu: projection string has no SPHEROID");
    const size_t c1 = P.wkt.find(',', sp);
    VG_CHECK(c1 != std::string::npos, VIRA_GPU_ERR_INVALID, "quipu: malformed SPHEROID");
    char* endp = nullptr;
    q.a = std::strtod(P.wkt.c_str() + c1 + 1, &endp);
    const size_t c2 = P.wkt.find(',', (size_t)(endp - P.wkt.c_str()));
    const double inv_f = c2 == std::string::npos ? 0.0 : std::strtod(P.wkt.c_str() + c2 + 1, nullptr);
    VG_CHECK(q.a > 0, VIRA_GPU_ERR_INVALID, "quipu: SPHEROID semi-major axis must be positive");
    q.b = inv_f == 0.0 ? q.a : q.a * (1.0 - 1.0 / inv_f);
    q.lon0 = q.lat_ts = q.lat0 = q.fe = q.fn = 0.0;
    wkt_param(P.wkt, "central_meridian", q.lon0) || wkt_param(P.wkt, "longitude_of_center", q.lon0);
    wkt_param(P.wkt, "standard_parallel_1", q.lat_ts) || wkt_param(P.wkt, "latitude_of_true_scale", q.lat_ts);
    wkt_param(P.wkt, "latitude_of_origin", q.lat0) || wkt_param(P.wkt, "latitude_of_center", q.lat0);
    wkt_param(P.wkt, "false_easting", q.fe);
    wkt_param(P.wkt, "false_northing", q.fn);
    q.pole = 1.0;
    q.akm1 = 2.0;
    if (ps) {
        // GDAL's reading of WKT1 polar stereographic: a latitude of origin of
        // +-90 is the pole with scale_factor k0; any other value is the
        // latitude of true scale (the ESRI names give it as
        // standard_parallel_1), the pole on its side.  PROJ's stere then uses
        // akm1 = 2 k0, or cos(phits) / tan(pi/4 - phits/2) for |phits| < 90.
        VG_CHECK(q.a == q.b, VIRA_GPU_ERR_UNSUPPORTED,
                 "quipu: polar stereographic is supported on a sphere only (inverse flattening 0)");
        double lat_o = ps_s ? -90.0 : 90.0, k0 = 1.0;
        wkt_param(P.wkt, "latitude_of_origin", lat_o) || wkt_param(P.wkt, "standard_parallel_1", lat_o);
        wkt_param(P.wkt, "scale_factor", k0);
        wkt_param(P.wkt, "central_meridian", q.lon0) ||
            wkt_param(P.wkt, "straight_vertical_longitude_from_pole", q.lon0) ||
            wkt_param(P.wkt, "longitude_of_origin", q.lon0);
        q.pole = ps_n ? 1.0 : ps_s ? -1.0 : (lat_o < 0.0 ? -1.0 : 1.0);
        const double phits = std::fabs(lat_o) * (3.14159265358979323846 / 180.0);
        VG_CHECK(std::fabs(lat_o) <= 90.0 && k0 > 0.0, VIRA_GPU_ERR_INVALID,
                 "quipu: polar stereographic latitude or scale factor out of range");
        q.akm1 = std::fabs(phits - 3.14159265358979323846 / 2) >= 1e-10
                     ? std::cos(phits) / std::tan(3.14159265358979323846 / 4 - 0.5 * phits)
                     : 2.0 * k0;
        q.kind = 2;
    }
    return q;
}

// One level read from the file: heights (quantised or FP32) and albedo.
struct Level {
    Proj proj;
    float hmm[2];
    int height_bits;
    Img heights;
    int albedo_kind = 0;  // as k_qld_vertices
    float constant[NB];
    std::vector<Img> grey;       // 1 (grey image) or 8 (channels)
    std::vector<float> grey_mm;  // min/max per grey image
};

// The level's images go to the context's page-locked staging buffers (grown
// as needed and kept): heights in one, albedo images one after another in the
// other.  A read synchronises before it returns, so the next one may reuse them.
Level read_level(Ctx& c, const vira_gpu_qld& q, uint64_t offset, bool read_albedos, double& decode_ms) {
    Rd r(q.path);
    r.seek(offset);
    Level L;
    L.proj = read_projection(r);
    // readDEMHeights (quipu_io.ipp:662-687)
    L.hmm[0] = r.get<float>();
    L.hmm[1] = r.get<float>();
    const uint16_t prec = r.get<uint16_t>();
    VG_CHECK(prec == kIdU8 || prec == kIdU16 || prec == kIdFloat, VIRA_GPU_ERR_INVALID,
             "quipu: height precision class id " + std::to_string(prec));
    L.height_bits = prec == kIdU8 ? 8 : (prec == kIdU16 ? 16 : 32);
    L.heights = read_image(r, prec, L.height_bits / 8, decode_ms, [&](size_t n) {
        c.qld_heights.alloc(n);
        return c.qld_heights.as<uint8_t>();
    });
    VG_CHECK(L.heights.w >= 2 && L.heights.h >= 2, VIRA_GPU_ERR_INVALID, "quipu: level smaller than 2x2");
    // albedo (dem_quipu.ipp:176-209): grey images of the heights' size, one byte per pixel
    const size_t npix = (size_t)L.heights.w * L.heights.h;
    auto grey_slot = [&](size_t k) {
        return [&c, npix, k](size_t n) {
            VG_CHECK(n == npix, VIRA_GPU_ERR_INVALID, "quipu: albedo resolution differs from the heights'");
            return c.qld_albedo.as<uint8_t>() + k * npix;
        };
    };
    if (q.has_albedos && read_albedos) {
        const uint8_t type = r.get<uint8_t>();
        if (type == 0) {
            L.albedo_kind = 1;
            r.bytes(L.constant, sizeof(L.constant));
        } else if (type == 1) {
            L.albedo_kind = 2;
            c.qld_albedo.alloc(npix);
            L.grey_mm.push_back(r.get<float>());
            L.grey_mm.push_back(r.get<float>());
            L.grey.push_back(read_image(r, kIdU8, 1, decode_ms, grey_slot(0)));
        } else if (type == 2) {
            L.albedo_kind = 3;
            const uint32_t nc = r.get<uint32_t>();
            VG_CHECK(nc == NB, VIRA_GPU_ERR_INVALID,
                     "quipu: spectral albedo with " + std::to_string(nc) + " channels (8 expected)");
            c.qld_albedo.alloc(NB * npix);
            (void)r.get<uint8_t>();  // the channels' compressed flag, repeated per image
            for (uint32_t k = 0; k < nc; ++k) {
                L.grey_mm.push_back(r.get<float>());
                L.grey_mm.push_back(r.get<float>());
                L.grey.push_back(read_image(r, kIdU8, 1, decode_ms, grey_slot(k)));
            }
        } else {
            throw Error(VIRA_GPU_ERR_INVALID, "quipu: albedo type " + std::to_string(type));
        }
        for (const Img& g : L.grey)
            VG_CHECK(g.w == L.heights.w && g.h == L.heights.h, VIRA_GPU_ERR_INVALID,
                     "quipu: albedo resolution differs from the heights'");
    }
    return L;
}

template <class T>
void launch_fixed(Ctx& c, const void* q, size_t n, float lo, float hi, float* out) {
    k_fixed_to_float<T><<<blocks_for(n, 256), 256, 0, c.stream>>>((const T*)q, n, lo, hi, out);
    VG_HIP(hipGetLastError());
}

}  // namespace

vira_gpu_qld* qld_open(const char* path, vira_gpu_qld_info* info) {
    VG_CHECK(path != nullptr, VIRA_GPU_ERR_INVALID, "quipu: null path");
    Rd r(path);
    char ident[5];
    r.bytes(ident, 5);
    VG_CHECK(std::memcmp(ident, "QUIPU", 5) == 0, VIRA_GPU_ERR_INVALID,
             std::string(path) + " exists, but is not a valid Quipu");
    const uint16_t cls = r.get<uint16_t>();
    VG_CHECK(cls == kIdDemPyramid, VIRA_GPU_ERR_INVALID,
             "quipu: class id " + std::to_string(cls) + " is not a DEM pyramid (2002)");
    auto q = std::make_unique<vira_gpu_qld>();
    q->path = path;
    q->file_size = r.size;
    read_transformation(r, q->M);
    q->levels = r.get<uint16_t>();
    q->has_albedos = r.get<uint8_t>() != 0;
    double n3[3];
    read_vec(r, 3, n3);
    for (int k = 0; k < 3; ++k) q->normal[k] = (float)n3[k];
    q->cone = r.get<float>();
    VG_CHECK(q->levels >= 1, VIRA_GPU_ERR_INVALID, "quipu: no levels");
    for (int i = 0; i < q->levels; ++i) {
        q->gsd.push_back(r.get<double>());
        q->offset.push_back(r.get<uint64_t>());
        VG_CHECK(q->offset.back() < r.size, VIRA_GPU_ERR_INVALID, "quipu: level offset beyond the end of the file");
    }
    if (info) {
        info->levels = q->levels;
        info->has_albedos = q->has_albedos ? 1 : 0;
        for (int k = 0; k < 16; ++k) info->transformation[k] = q->M[k];
        for (int k = 0; k < 3; ++k) info->normal[k] = q->normal[k];
        info->cone_angle = q->cone;
    }
    return q.release();
}

// DEMQuipu::readBuffers' choice (dem_quipu.ipp:137-151).
int qld_select(const vira_gpu_qld& q, double required_gsd) {
    int sel = 0;
    for (int i = 1; i < q.levels; ++i) {
        if (q.gsd[i] <= required_gsd) sel = i;
        else break;
    }
    return sel;
}

void qld_read(Ctx& c, const vira_gpu_qld& q, int level, const vira_gpu_qld_read_opts& o, float* heights_dev,
              float* albedo_dev, int32_t* mesh_id_out, vira_gpu_qld_level* info,
              const std::function<int32_t()>& new_mesh) {
    VG_CHECK(level >= 0 && level < q.levels, VIRA_GPU_ERR_INVALID, "quipu: level out of range");
    const auto t0 = std::chrono::steady_clock::now();
    double decode_ms = 0.0;
    Level L = read_level(c, q, q.offset[level], o.read_albedos != 0, decode_ms);
    const auto t1 = std::chrono::steady_clock::now();
    const int W = (int)L.heights.w, H = (int)L.heights.h;
    const size_t n = (size_t)W * H;
    VG_CHECK(n < (1ull << 31), VIRA_GPU_ERR_INVALID, "quipu: level too large for 32-bit vertex ids");
    // quantised samples to the device, dequantised there
    DBuf hq, hbuf, abuf, aq;
    hq.alloc(L.heights.bytes());
    VG_HIP(hipMemcpyAsync(hq.p, L.heights.raw, L.heights.bytes(), hipMemcpyHostToDevice, c.stream));
    float* heights = heights_dev;
    if (!heights) {
        hbuf.alloc(n * sizeof(float));
        heights = (float*)hbuf.p;
    }
    if (L.height_bits == 8) launch_fixed<uint8_t>(c, hq.p, n, L.hmm[0], L.hmm[1], heights);
    else if (L.height_bits == 16) launch_fixed<uint16_t>(c, hq.p, n, L.hmm[0], L.hmm[1], heights);
    else VG_HIP(hipMemcpyAsync(heights, hq.p, n * sizeof(float), hipMemcpyDeviceToDevice, c.stream));
    // albedo: the constant spectrum, or grey images dequantised on the device
    int kind = L.albedo_kind;
    const float* alb = nullptr;
    if (kind == 1) {
        abuf.alloc(NB * sizeof(float));
        VG_HIP(hipMemcpyAsync(abuf.p, L.constant, sizeof(L.constant), hipMemcpyHostToDevice, c.stream));
        alb = (const float*)abuf.p;
    } else if (kind >= 2) {
        const size_t ng = L.grey.size();
        abuf.alloc(ng * n * sizeof(float));
        aq.alloc(ng * n);
        for (size_t g = 0; g < ng; ++g) {
            VG_HIP(hipMemcpyAsync((uint8_t*)aq.p + g * n, L.grey[g].raw, n, hipMemcpyHostToDevice, c.stream));
            launch_fixed<uint8_t>(c, (uint8_t*)aq.p + g * n, n, L.grey_mm[2 * g], L.grey_mm[2 * g + 1],
                                  (float*)abuf.p + g * n);
        }
        alb = (const float*)abuf.p;
        // DEMQuipuReaderOptions::albedoProfile == 1 keeps the grey image as is (dem_quipu.ipp:190-198)
    }
    if (albedo_dev) {
        // W*H*8 per-vertex albedo as the mesh would carry it (k_qld_vertices' rule), for callers that
        // want the decoded data without a mesh
        DBuf tmpv, tmp64;
        tmpv.alloc(n * sizeof(GVertex));
        tmp64.alloc(n * 3 * sizeof(double));
        QProj qp = make_qproj(L.proj, o.offset);
        k_qld_vertices<<<blocks_for(n, 256), 256, 0, c.stream>>>(heights, W, H, qp, kind, alb, o,
                                                                 (GVertex*)tmpv.p, tmp64.as<double>());
        VG_HIP(hipGetLastError());
        VG_HIP(hipMemcpy2DAsync(albedo_dev, NB * sizeof(float), (const char*)tmpv.p + offsetof(GVertex, alb),
                                sizeof(GVertex), NB * sizeof(float), n, hipMemcpyDeviceToDevice, c.stream));
        VG_HIP(hipStreamSynchronize(c.stream));
    }
    int32_t id = -1;
    if (mesh_id_out) {
        id = new_mesh();
        MeshG& m = c.mesh(id);
        const QProj qp = make_qproj(L.proj, o.offset);
        dem_grid_mesh(c, m, heights, W, H, [&](GVertex* verts, double* pos64) {
            k_qld_vertices<<<blocks_for(n, 256), 256, 0, c.stream>>>(heights, W, H, qp, kind, alb, o, verts, pos64);
            VG_HIP(hipGetLastError());
        });
        *mesh_id_out = id;
    }
    VG_HIP(hipStreamSynchronize(c.stream));
    const auto t2 = std::chrono::steady_clock::now();
    if (info) {
        info->level = level;
        info->gsd = q.gsd[level];
        info->width = W;
        info->height = H;
        info->height_bits = L.height_bits;
        info->compressed = L.heights.compressed ? 1 : 0;
        info->albedo_type = L.albedo_kind - 1;
        info->is_ogr = L.proj.ogr ? 1 : 0;
        info->read_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        info->decode_ms = decode_ms;
        info->build_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
}

}  // namespace vg

// oracle_diff.cpp -- FP64 fixed-visibility direct illumination and its adjoint.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp and oracle_diff.hpp).
#include "oracle_diff.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <unordered_set>

namespace vo {
namespace {

using D3 = V3<double>;
constexpr double kPiD = 3.14159265358979323846;

inline D3 node_pos(const DemDiff& d, const double* h, uint32_t v) {
    int i = (int)(v % (uint32_t)d.W), j = (int)(v / (uint32_t)d.W);
    D3 p = dem_world(d.proj, d.W, d.H, i, j, h[v]);
    return {p.x - d.proj.offset[0], p.y - d.proj.offset[1], p.z - d.proj.offset[2]};
}

// d position / d height: +z on the flat grid, the radial direction on the sphere.
inline D3 node_up(const DemDiff& d, uint32_t v) {
    if (d.proj.kind == 0) return {0.0, 0.0, 1.0};
    int i = (int)(v % (uint32_t)d.W), j = (int)(v / (uint32_t)d.W);
    double cx = (double)(d.W - 1) * 0.5, cy = (double)(d.H - 1) * 0.5, a = d.proj.spacing / d.proj.radius;
    D3 q{1.0, ((double)i - cx) * a, (cy - (double)j) * a};
    return q * (1.0 / length(q));
}

struct Face {
    D3 p[3];
    D3 e, fn;
    double le;
    bool ok;
};

inline Face face_at(const DemDiff& d, const double* h, uint32_t f) {
    Face F;
    for (int a = 0; a < 3; ++a) F.p[a] = node_pos(d, h, d.idx[3 * f + a]);
    F.e = cross(F.p[1] - F.p[0], F.p[2] - F.p[0]);
    F.le = length(F.e);
    F.ok = F.le > 0.0 && std::isfinite(F.le);
    F.fn = F.ok ? F.e * (1.0 / F.le) : D3{0, 0, 0};
    return F;
}

// Sum of unit face normals around node v (mesh.ipp:199-233, degenerate faces skipped).
inline D3 normal_sum(const DemDiff& d, const double* h, uint32_t v) {
    D3 s{0, 0, 0};
    for (uint32_t k = d.adj_off[v]; k < d.adj_off[v + 1]; ++k) {
        Face F = face_at(d, h, d.adj[k]);
        if (F.ok) s = s + F.fn;
    }
    return s;
}

// Bilinear texel taps (image.ipp:589-670): indices and weights.
inline void tex_taps(const Material& m, double u, double v, int64_t idx[4], double w[4]) {
    double rx = (double)m.tex_w, ry = (double)m.tex_h;
    double px = std::fmod(u * rx, rx), py = std::fmod(v * ry, ry);
    if (px < 0) px += rx;
    if (py < 0) py += ry;
    if (!(px >= 0)) px = 0;
    if (!(py >= 0)) py = 0;
    int x0 = std::min((int)std::floor(px), m.tex_w - 1), y0 = std::min((int)std::floor(py), m.tex_h - 1);
    int x1 = std::min(x0 + 1, m.tex_w - 1), y1 = std::min(y0 + 1, m.tex_h - 1);
    double dx = px - x0, dy = py - y0;
    idx[0] = x0 + (int64_t)y0 * m.tex_w;
    idx[1] = x1 + (int64_t)y0 * m.tex_w;
    idx[2] = x0 + (int64_t)y1 * m.tex_w;
    idx[3] = x1 + (int64_t)y1 * m.tex_w;
    w[0] = (1 - dx) * (1 - dy);
    w[1] = dx * (1 - dy);
    w[2] = (1 - dx) * dy;
    w[3] = dx * dy;
}

// Gradient sinks of one record: node gradients of positions, texel gradients.
struct Sink {
    virtual void pos(uint32_t v, const D3& g) = 0;
    virtual void texel(int64_t t, double g) = 0;
    virtual ~Sink() = default;
};

// Adds the gradient of a unit face normal fn = e/|e| (e = cross(p1-p0, p2-p0))
// with respect to the face's corner positions.
inline void face_normal_backward(const DemDiff& d, const Face& F, uint32_t f, const D3& gfn, Sink& sink) {
    if (!F.ok) return;
    D3 ge = (gfn - dot(gfn, F.fn) * F.fn) * (1.0 / F.le);
    D3 a = F.p[1] - F.p[0], b = F.p[2] - F.p[0];
    D3 ga = cross(b, ge), gb = cross(ge, a);
    sink.pos(d.idx[3 * f + 1], ga);
    sink.pos(d.idx[3 * f + 2], gb);
    sink.pos(d.idx[3 * f + 0], D3{0, 0, 0} - (ga + gb));
}

// Contribution of one record (NB bins, before the 1/spp average); with G the
// upstream gradient of those bins, also pushes its gradient into the sink.
void record_eval(const DiScene& s, const double* h, const double* tex, const DiRecord& r, double out[NB],
                 const double* G, Sink* sink) {
    for (int k = 0; k < NB; ++k) out[k] = 0;
    if (r.id == 0xffffffffu) return;
    const DemDiff& d = *s.dd;
    const Material& mat = *s.mat;
    const uint32_t v[3] = {d.idx[3 * r.id], d.idx[3 * r.id + 1], d.idx[3 * r.id + 2]};
    const double w0 = r.b1, w1 = r.b2, w2 = 1.0 - (double)r.b1 - (double)r.b2;
    const double om[3] = {w2, w0, w1};  // weights of corners 0, 1, 2 (embree_tlas.ipp:160-166)
    Face F = face_at(d, h, r.id);
    D3 ns[3], n[3];
    double ls[3];
    for (int a = 0; a < 3; ++a) {
        ns[a] = normal_sum(d, h, v[a]);
        ls[a] = length(ns[a]);
        n[a] = ls[a] > 0 ? ns[a] * (1.0 / ls[a]) : ns[a];
    }
    D3 m = (om[1] * n[1] + om[2] * n[2]) + om[0] * n[0];
    double lm = length(m);
    D3 N = mat.smooth ? m * (1.0 / lm) : F.fn;
    D3 P0 = (om[1] * F.p[1] + om[2] * F.p[2]) + om[0] * F.p[0];
    double del[3];
    D3 sp{0, 0, 0};
    for (int a = 0; a < 3; ++a) {
        del[a] = dot(P0 - F.p[a], n[a]);
        sp = sp + om[a] * (P0 - del[a] * n[a]);
    }
    const bool use = dot(sp - P0, N) > 0.0;
    D3 P = use ? sp : P0;
    double texv = 1.0;
    int64_t ti[4];
    double tw[4];
    if (mat.tex) {
        double u = (om[1] * d.uv[2 * v[1]] + om[2] * d.uv[2 * v[2]]) + om[0] * d.uv[2 * v[0]];
        double vv = (om[1] * d.uv[2 * v[1] + 1] + om[2] * d.uv[2 * v[2] + 1]) + om[0] * d.uv[2 * v[0] + 1];
        tex_taps(mat, u, vv, ti, tw);
        texv = 0;
        for (int q = 0; q < 4; ++q) texv += tw[q] * tex[ti[q]];
    }
    double av[NB], mv[NB], alb[NB];
    for (int k = 0; k < NB; ++k) {
        av[k] = (om[1] * d.alb[NB * v[1] + k] + om[2] * d.alb[NB * v[2] + k]) + om[0] * d.alb[NB * v[0] + k];
        mv[k] = mat.tex ? texv * mat.albedo[k] : mat.albedo[k];
        alb[k] = av[k] * mv[k];
    }
    D3 gN{0, 0, 0}, gP{0, 0, 0};
    double gtex = 0;
    for (int l = 0; l < s.nl; ++l) {
        if (!((r.vis >> l) & 1u)) continue;
        const Light& L = s.lights[l];
        if (L.kind != 0) continue;
        D3 toL = D3{L.pos[0], L.pos[1], L.pos[2]} - P;
        double dist = length(toL);
        D3 dir = toL * (1.0 / dist);
        double c = dot(N, dir);
        if (!(c > 0.0)) continue;
        double Fc = (c * c) / (kPiD * (1.0 + (c / kPiD) * (c / kPiD)));
        double dFc = 2.0 * kPiD * kPiD * kPiD * c / ((kPiD * kPiD + c * c) * (kPiD * kPiD + c * c));
        double gc = 0, gdist = 0;
        for (int k = 0; k < NB; ++k) {
            double rad = (double)L.value[k] / (dist * dist);
            out[k] += alb[k] * rad * Fc;
            if (G) {
                gc += G[k] * alb[k] * rad * dFc;
                gdist += G[k] * alb[k] * rad * Fc * (-2.0 / dist);
                if (mat.tex) gtex += G[k] * av[k] * mat.albedo[k] * rad * Fc;
            }
        }
        if (!G) continue;
        gN = gN + gc * dir;
        D3 gdir = gc * N;
        gP = gP - (gdir - dot(gdir, dir) * dir) * (1.0 / dist) - gdist * dir;
    }
    if (!G || !sink) return;
    if (mat.tex)
        for (int q = 0; q < 4; ++q) sink->texel(ti[q], gtex * tw[q]);
    D3 gp[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gn[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    D3 gP0 = gP;
    if (use) {
        double sw = om[0] + om[1] + om[2];
        gP0 = sw * gP;
        for (int a = 0; a < 3; ++a) {
            double ng = dot(n[a], gP);
            gP0 = gP0 - (om[a] * ng) * n[a];
            gp[a] = gp[a] + (om[a] * ng) * n[a];
            gn[a] = gn[a] - om[a] * (ng * (P0 - F.p[a]) + del[a] * gP);
        }
    }
    for (int a = 0; a < 3; ++a) gp[a] = gp[a] + om[a] * gP0;
    D3 gfn{0, 0, 0};
    if (mat.smooth) {
        D3 gm = (gN - dot(gN, N) * N) * (1.0 / lm);
        for (int a = 0; a < 3; ++a) gn[a] = gn[a] + om[a] * gm;
    } else {
        gfn = gN;
    }
    for (int a = 0; a < 3; ++a) sink->pos(v[a], gp[a]);
    face_normal_backward(d, F, r.id, gfn, *sink);
    for (int a = 0; a < 3; ++a) {
        if (!(ls[a] > 0)) continue;
        D3 gs = (gn[a] - dot(gn[a], n[a]) * n[a]) * (1.0 / ls[a]);
        for (uint32_t k = d.adj_off[v[a]]; k < d.adj_off[v[a] + 1]; ++k) {
            uint32_t f = d.adj[k];
            Face Ff = face_at(d, h, f);
            face_normal_backward(d, Ff, f, gs, *sink);
        }
    }
}

void pixel_eval(const DiScene& s, const double* h, const double* tex, size_t p, double img[NB]) {
    for (int k = 0; k < NB; ++k) img[k] = 0;
    double c[NB];
    for (int smp = 0; smp < s.spp; ++smp) {
        record_eval(s, h, tex, s.rec[(size_t)smp * s.npix + p], c, nullptr, nullptr);
        for (int k = 0; k < NB; ++k) img[k] += c[k];
    }
    for (int k = 0; k < NB; ++k) img[k] /= (double)s.spp;
}

struct DenseSink : Sink {
    const DemDiff& d;
    double* dh;
    double* dtex;
    DenseSink(const DemDiff& dd, double* a, double* b) : d(dd), dh(a), dtex(b) {}
    void pos(uint32_t v, const D3& g) override {
        if (dh) dh[v] += dot(g, node_up(d, v));
    }
    void texel(int64_t t, double g) override {
        if (dtex) dtex[t] += g;
    }
};

// fn(i, worker) over i in [0, n) on nthreads workers (worker 0 is the caller).
template <class F>
void parallel_for(size_t n, int nthreads, size_t chunk, F fn) {
    std::atomic<size_t> next{0};
    auto work = [&](int tid) {
        while (true) {
            size_t i0 = next.fetch_add(chunk);
            if (i0 >= n) break;
            for (size_t i = i0; i < std::min(n, i0 + chunk); ++i) fn(i, tid);
        }
    };
    int nt = std::max(1, nthreads);
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& t : th) t.join();
}

}  // namespace

DemDiff dem_diff_make(int W, int H, const Projection& p, const float* heights, const float* grey_albedo) {
    DemDiff d;
    d.W = W;
    d.H = H;
    d.proj = p;
    dem_indices(heights, W, H, d.idx);
    size_t nv = (size_t)W * H, nt = d.idx.size() / 3;
    d.uv.resize(2 * nv);
    for (int j = 0; j < H; ++j)
        for (int i = 0; i < W; ++i) {
            size_t v = (size_t)j * W + i;
            d.uv[2 * v] = (float)((double)i / (double)(W - 1));
            d.uv[2 * v + 1] = (float)((double)j / (double)(H - 1));
        }
    d.alb.assign((size_t)NB * nv, 1.0f);
    if (grey_albedo)
        for (size_t v = 0; v < nv; ++v)
            for (int k = 0; k < NB; ++k) d.alb[NB * v + k] = grey_albedo[v];
    d.adj_off.assign(nv + 1, 0);
    for (size_t t = 0; t < nt; ++t)
        for (int a = 0; a < 3; ++a) d.adj_off[d.idx[3 * t + a] + 1]++;
    for (size_t v = 0; v < nv; ++v) d.adj_off[v + 1] += d.adj_off[v];
    d.adj.resize(d.adj_off[nv]);
    std::vector<uint32_t> cur(d.adj_off.begin(), d.adj_off.end() - 1);
    for (size_t t = 0; t < nt; ++t)
        for (int a = 0; a < 3; ++a) d.adj[cur[d.idx[3 * t + a]]++] = (uint32_t)t;
    return d;
}

void di_forward64(const DiScene& s, const double* h, const double* tex, double* img, int nthreads) {
    parallel_for(s.npix, nthreads, 256, [&](size_t p, int) { pixel_eval(s, h, tex, p, img + NB * p); });
}

void di_adjoint64(const DiScene& s, const double* h, const double* tex, const double* g, double* dh,
                  double* dtex) {
    DenseSink sink(*s.dd, dh, dtex);
    double c[NB], G[NB];
    for (size_t p = 0; p < s.npix; ++p) {
        for (int k = 0; k < NB; ++k) G[k] = g[NB * p + k] / (double)s.spp;
        for (int smp = 0; smp < s.spp; ++smp)
            record_eval(s, h, tex, s.rec[(size_t)smp * s.npix + p], c, G, &sink);
    }
}

void di_fd_heights64(const DiScene& s, const double* h, const double* tex, const double* target,
                     const int64_t* nodes, size_t n, double hstep, double* out, int nthreads) {
    const DemDiff& d = *s.dd;
    const size_t nt = d.idx.size() / 3;
    // triangle -> pixels whose records hit it
    std::vector<uint32_t> off(nt + 1, 0), pix;
    for (int smp = 0; smp < s.spp; ++smp)
        for (size_t p = 0; p < s.npix; ++p) {
            uint32_t id = s.rec[(size_t)smp * s.npix + p].id;
            if (id != 0xffffffffu) off[id + 1]++;
        }
    for (size_t t = 0; t < nt; ++t) off[t + 1] += off[t];
    pix.resize(off[nt]);
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (int smp = 0; smp < s.spp; ++smp)
        for (size_t p = 0; p < s.npix; ++p) {
            uint32_t id = s.rec[(size_t)smp * s.npix + p].id;
            if (id != 0xffffffffu) pix[cur[id]++] = (uint32_t)p;
        }
    const size_t nv = (size_t)d.W * d.H;
    const int nthr = std::max(1, std::min(nthreads, (int)n));
    std::vector<std::vector<double>> hcopy(nthr);
    parallel_for(n, nthr, 1, [&](size_t q, int tid) {
        const uint32_t v = (uint32_t)nodes[q];
        // faces whose normals or corners move with h_v: faces around v and around its neighbours
        std::unordered_set<uint32_t> ring{v}, faces;
        for (uint32_t k = d.adj_off[v]; k < d.adj_off[v + 1]; ++k)
            for (int a = 0; a < 3; ++a) ring.insert(d.idx[3 * d.adj[k] + a]);
        for (uint32_t u : ring)
            for (uint32_t k = d.adj_off[u]; k < d.adj_off[u + 1]; ++k) faces.insert(d.adj[k]);
        std::unordered_set<uint32_t> pset;
        for (uint32_t f : faces)
            for (uint32_t k = off[f]; k < off[f + 1]; ++k) pset.insert(pix[k]);
        std::vector<double>& hh = hcopy[tid];
        if (hh.empty()) hh.assign(h, h + nv);
        double loss[2];
        for (int sgn = 0; sgn < 2; ++sgn) {
            hh[v] = h[v] + (sgn ? -hstep : hstep);
            double L = 0, img[NB];
            for (uint32_t p : pset) {
                pixel_eval(s, hh.data(), tex, p, img);
                for (int k = 0; k < NB; ++k) {
                    double r = img[k] - target[NB * p + k];
                    L += r * r;
                }
            }
            loss[sgn] = L;
        }
        hh[v] = h[v];
        out[q] = (loss[0] - loss[1]) / (2.0 * hstep);
    });
}

}  // namespace vo

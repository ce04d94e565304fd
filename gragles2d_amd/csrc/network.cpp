// network.cpp -- seeded microstructures and the mode-0 network file.
//
// Reader: grainhdl::readMicrostructure (src/grainhdl.cpp:605-661): per grain a header line
// "ID nVertices q0 q1 q2 q3" followed by nVertices "x y" lines (unit square), grains are
// numbered by their position in the file.  Writer: LSbox::plot_full_grain
// (src/box.cpp:1552-1596) as called by save_Full_Microstructure_for_Restart.
//
// Generator: stands in for grainhdl::VOROMicrostructure (src/grainhdl.cpp:513-603), whose
// seed points come from an unseeded rand() and whose cells come from a prebuilt voro++.  Here
// the cells are computed by half-plane clipping of the unit square (Voronoi) or with power
// bisectors (lognormal weights), from a SplitMix64 stream, so every run is reproducible.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ggx.h"
#include "ggx_internal.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
    }
};

struct P2 { double x, y; };

// Sutherland-Hodgman clip of a convex polygon by a*x + b*y <= c.
void clip(std::vector<P2>& poly, double a, double b, double c, std::vector<P2>& tmp) {
    tmp.clear();
    size_t n = poly.size();
    for (size_t i = 0; i < n; ++i) {
        const P2& P = poly[i];
        const P2& Q = poly[(i + 1) % n];
        double fp = a * P.x + b * P.y - c, fq = a * Q.x + b * Q.y - c;
        bool pin = fp <= 0, qin = fq <= 0;
        if (pin) tmp.push_back(P);
        if (pin != qin) {
            double t = fp / (fp - fq);
            tmp.push_back(P2{P.x + t * (Q.x - P.x), P.y + t * (Q.y - P.y)});
        }
    }
    poly.swap(tmp);
}

}  // namespace

extern "C" int ggx_network_new(ggx_network** out) {
    if (!out) return GGX_E_ARG;
    *out = new ggx_network();
    return GGX_OK;
}

extern "C" void ggx_network_free(ggx_network* net) { delete net; }

extern "C" int ggx_network_count(const ggx_network* net) { return net ? (int)net->ids.size() : GGX_E_ARG; }

extern "C" int ggx_network_grain(const ggx_network* net, int32_t idx, int32_t* id, int32_t* nverts,
                                 const double** xy, const double** quat) {
    if (!net || idx < 0 || idx >= (int32_t)net->ids.size()) return GGX_E_ARG;
    if (id) *id = net->ids[idx];
    if (nverts) *nverts = (int32_t)(net->xy[idx].size() / 2);
    if (xy) *xy = net->xy[idx].data();
    if (quat) *quat = net->q[idx].data();
    return GGX_OK;
}

extern "C" int ggx_network_add(ggx_network* net, int32_t id, int32_t nverts, const double* xy,
                               const double* quat) {
    if (!net || nverts < 2 || !xy || id < 1) return GGX_E_ARG;
    net->ids.push_back(id);
    net->xy.emplace_back(xy, xy + 2 * (size_t)nverts);
    std::array<double, 4> q{1, 0, 0, 0};
    if (quat) for (int k = 0; k < 4; ++k) q[k] = quat[k];
    net->q.push_back(q);
    return GGX_OK;
}

namespace {

// One tessellation from `n` seeds; empty power cells are dropped.
ggx_network* tessellate(int32_t n, uint64_t seed, double lognormal_sigma) {
    SplitMix64 rng(seed);
    std::vector<P2> pts((size_t)n);
    for (int i = 0; i < n; ++i) { pts[i].x = rng.uniform(); pts[i].y = rng.uniform(); }
    // orientations: LSbox default draw, Quaternion::randomOriShoemakeQuat (Quaternion.cpp:291-307)
    std::vector<std::array<double, 4>> q((size_t)n);
    for (int i = 0; i < n; ++i) {
        double s = rng.uniform();
        double s1 = std::sqrt(1 - s), s2 = std::sqrt(s);
        double t1 = 2 * 3.14159265358979323846 * rng.uniform();
        double t2 = 2 * 3.14159265358979323846 * rng.uniform();
        double a = std::fabs(s1 * std::sin(t1)), b = std::fabs(s1 * std::cos(t1));
        double c = std::fabs(s2 * std::sin(t2)), d = std::fabs(s2 * std::cos(t2));
        double nrm = std::sqrt(a * a + b * b + c * c + d * d);
        q[i] = {a / nrm, b / nrm, c / nrm, d / nrm};
    }
    // power weights: w_i = (0.2 * spacing * s_i)^2 with s_i lognormal(0, sigma); the factor 0.2
    // makes the standard deviation of log(equivalent radius) equal sigma at sigma = 0.35
    double spacing = 1.0 / std::sqrt((double)n);
    std::vector<double> w((size_t)n, 0.0);
    double wmax = 0.0;
    if (lognormal_sigma > 0) {
        for (int i = 0; i < n; ++i) {
            double si = std::exp(lognormal_sigma * rng.normal());
            w[i] = (0.2 * spacing * si) * (0.2 * spacing * si);
            wmax = std::max(wmax, w[i]);
        }
    }
    // uniform bins of the seeds
    int G = std::max(1, (int)std::ceil(std::sqrt((double)n)));
    double bs = 1.0 / G;
    std::vector<int> bin_start((size_t)G * G + 1, 0), bin_items((size_t)n);
    auto bin_of = [&](const P2& p) {
        int bx = std::min(G - 1, std::max(0, (int)(p.x / bs)));
        int by = std::min(G - 1, std::max(0, (int)(p.y / bs)));
        return by * G + bx;
    };
    for (int i = 0; i < n; ++i) bin_start[bin_of(pts[i]) + 1]++;
    for (int b = 0; b < G * G; ++b) bin_start[b + 1] += bin_start[b];
    {
        std::vector<int> fill(bin_start.begin(), bin_start.end() - 1);
        for (int i = 0; i < n; ++i) bin_items[fill[bin_of(pts[i])]++] = i;
    }
    std::vector<std::vector<P2>> cells((size_t)n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        std::vector<P2> poly{{0, 0}, {1, 0}, {1, 1}, {0, 1}}, tmp;
        const P2 p = pts[i];
        int b = bin_of(p), bx = b % G, by = b / G;
        for (int ring = 0; ring <= G; ++ring) {
            for (int yy = by - ring; yy <= by + ring; ++yy) {
                if (yy < 0 || yy >= G) continue;
                for (int xx = bx - ring; xx <= bx + ring; ++xx) {
                    if (xx < 0 || xx >= G) continue;
                    if (std::max(std::abs(xx - bx), std::abs(yy - by)) != ring) continue;
                    int bb = yy * G + xx;
                    for (int k = bin_start[bb]; k < bin_start[bb + 1]; ++k) {
                        int j = bin_items[k];
                        if (j == i) continue;
                        // |x-p|^2 - w_i <= |x-q|^2 - w_j  <=>  2 x.(q-p) <= |q|^2-|p|^2 - w_j + w_i
                        const P2 qq = pts[j];
                        double a = 2 * (qq.x - p.x), bcoef = 2 * (qq.y - p.y);
                        double c = (qq.x * qq.x + qq.y * qq.y) - (p.x * p.x + p.y * p.y) - w[j] + w[i];
                        clip(poly, a, bcoef, c, tmp);
                        if (poly.empty()) break;
                    }
                    if (poly.empty()) break;
                }
                if (poly.empty()) break;
            }
            if (poly.empty()) break;
            double rad = 0;
            for (const P2& v : poly) rad = std::max(rad, std::hypot(v.x - p.x, v.y - p.y));
            double reach = rad + std::sqrt(std::max(0.0, rad * rad - w[i] + wmax));
            if (ring * bs > reach) break;
        }
        // drop near-duplicate consecutive vertices
        std::vector<P2> clean;
        for (size_t k = 0; k < poly.size(); ++k) {
            const P2& v = poly[k];
            if (!clean.empty() && std::fabs(v.x - clean.back().x) < 1e-12 && std::fabs(v.y - clean.back().y) < 1e-12) continue;
            clean.push_back(v);
        }
        while (clean.size() > 1 && std::fabs(clean[0].x - clean.back().x) < 1e-12 && std::fabs(clean[0].y - clean.back().y) < 1e-12)
            clean.pop_back();
        if (clean.size() >= 3) cells[i] = clean;
    }
    ggx_network* net = new ggx_network();
    int id = 1;
    for (int i = 0; i < n; ++i) {
        if (cells[i].empty()) continue;
        std::vector<double> xy;
        for (const P2& v : cells[i]) { xy.push_back(v.x); xy.push_back(v.y); }
        xy.push_back(cells[i][0].x);
        xy.push_back(cells[i][0].y);
        net->ids.push_back(id++);
        net->xy.push_back(std::move(xy));
        net->q.push_back(q[i]);
    }
    return net;
}

}  // namespace

extern "C" int ggx_network_generate(int32_t n, uint64_t seed, double lognormal_sigma, ggx_network** out) {
    if (n < 1 || !out) return GGX_E_ARG;
    // a power diagram can leave cells empty: add seeds until `n` cells are non-empty
    int32_t seeds = n;
    ggx_network* net = tessellate(seeds, seed, lognormal_sigma);
    for (int it = 0; it < 16 && (int32_t)net->ids.size() < n; ++it) {
        seeds += n - (int32_t)net->ids.size();
        delete net;
        net = tessellate(seeds, seed, lognormal_sigma);
    }
    *out = net;
    return GGX_OK;
}

extern "C" int ggx_network_read(const char* path, int32_t ngrains, ggx_network** out) {
    if (!path || !out || ngrains < 1) return GGX_E_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return GGX_E_IO;  // reference: "Could not read from specified file !" + exit(2)
    ggx_network* net = new ggx_network();
    for (int nn = 1; nn <= ngrains; ++nn) {
        int id = 0, nv = 0;
        double q[4];
        if (std::fscanf(f, "%d\t", &id) != 1) break;
        if (std::fscanf(f, "%d\t %lf\t %lf\t%lf\t%lf\n", &nv, &q[0], &q[1], &q[2], &q[3]) != 5) break;
        std::vector<double> xy;
        for (int j = 0; j < nv; ++j) {
            double x, y;
            if (std::fscanf(f, "%lf\t %lf\n", &x, &y) != 2) { nv = j; break; }
            xy.push_back(x);
            xy.push_back(y);
        }
        if (std::fscanf(f, "\n") < 0) {}
        // grains are numbered by file position (grainhdl.cpp:627-657)
        net->ids.push_back(nn);
        net->xy.push_back(std::move(xy));
        net->q.push_back({q[0], q[1], q[2], q[3]});
    }
    std::fclose(f);
    *out = net;
    return GGX_OK;
}

extern "C" int ggx_network_write(const char* path, const ggx_network* net) {
    if (!path || !net) return GGX_E_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return GGX_E_IO;
    for (size_t g = 0; g < net->ids.size(); ++g) {
        size_t nv = net->xy[g].size() / 2;
        std::fprintf(f, "%d\t%zu\t%.17g\t%.17g\t%.17g\t%.17g\n", net->ids[g], nv, net->q[g][0], net->q[g][1],
                     net->q[g][2], net->q[g][3]);
        for (size_t k = 0; k < nv; ++k) std::fprintf(f, "%.17g\t%.17g\n", net->xy[g][2 * k], net->xy[g][2 * k + 1]);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    return GGX_OK;
}

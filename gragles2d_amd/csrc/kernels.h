// kernels.h -- CDNA4 (gfx950) kernels of the per-grain level-set timestep.
//
// Every kernel is batched: a launch covers all live grains of a step (one workgroup per
// grain unless stated otherwise); the per-grain launch mode of the baseline issues the same
// kernels with a one-grain work list.  Boxes live in device arenas addressed by per-grain
// offsets; T is the field type (double for the FP64 path, float for the FP32 path).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ggx_internal.h"

namespace ggx {

struct GridConst {
    int m;        // ngridpoints
    int gb;       // grid_blowup
    int R;        // realDomainSize
    int nslots;   // max grain id + 1 (grains.size())
    double h, delta, dt;
};

template <class T>
__device__ __forceinline__ int sgn_of(T v) {  // utilities.h:27-30
    return (T(0) < v) - (v < T(0));
}

// Flat index of DimensionalBuffer::getValueAt (dimensionalBuffer.h:68-73): 32-bit unsigned
// arithmetic, checked against the element count only.  Returns -1 when the reference would
// throw std::out_of_range.
__device__ __forceinline__ long long flat_index(int row, int col, int x0, int y0, int w, long long size) {
    uint32_t k = (uint32_t)row - (uint32_t)y0;
    k *= (uint32_t)w;
    k += (uint32_t)col - (uint32_t)x0;
    return ((long long)k < size) ? (long long)k : -1;
}

// ------------------------------------------------------------------------------------------
// Deletion of grains that vanished in the previous step (grainhdl::level_set,
// grainhdl.cpp:840-843) and reset of the per-step bump allocators.
// ------------------------------------------------------------------------------------------
__global__ void k_prep(int nslots, int* __restrict__ flags, unsigned long long* __restrict__ tops) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) { tops[0] = 0ull; tops[1] = 0ull; }
    if (g > 0 && g < nslots) {
        int f = flags[g];
        if ((f & kPresent) && !(f & kExists)) flags[g] = f & ~kPresent;
    }
}

// ------------------------------------------------------------------------------------------
// Comparison lists, LSbox::computeSecondOrderNeighbours (box.cpp:1027-1068): one thread per
// live grain.  cmp = existing entries of the previous list, then every existing
// neighbour-of-neighbour whose input box intersects (closed intervals, checkIntersection
// box.cpp:881-888), deduplicated in first-seen order, excluding self and the boundary.
// ------------------------------------------------------------------------------------------
__global__ void k_lists(int nlive, const int* __restrict__ live, const Box* __restrict__ box,
                        const int* __restrict__ flags, const int* __restrict__ sec_cnt,
                        const int* __restrict__ sec, int* __restrict__ cmp_cnt, int* __restrict__ cmp,
                        int* __restrict__ err) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlive) return;
    int g = live[t];
    Box bg = box[g];
    int* out = cmp + (size_t)g * kCmpCap;
    const int* sg = sec + (size_t)g * kCmpCap;
    int ns = sec_cnt[g];
    int cnt = 0;
    bool over = false;
    for (int a = 0; a < ns; ++a) {
        int nb = sg[a];
        if (flags[nb] & kExists) {
            if (cnt < kCmpCap) out[cnt++] = nb; else over = true;
        }
    }
    for (int a = 0; a < ns; ++a) {
        int nb = sg[a];
        int nn = sec_cnt[nb];
        const int* snb = sec + (size_t)nb * kCmpCap;
        for (int b = 0; b < nn; ++b) {
            int c = snb[b];
            if (!(flags[c] & kExists)) continue;
            Box bc = box[c];
            if (bg.x0 > bc.x0 + bc.n || bg.x0 + bg.n < bc.x0 || bg.y0 > bc.y0 + bc.n || bg.y0 + bg.n < bc.y0) continue;
            if (c == g || c == 0) continue;
            bool in = false;
            for (int k = 0; k < cnt; ++k)
                if (out[k] == c) { in = true; break; }
            if (!in) {
                if (cnt < kCmpCap) out[cnt++] = c; else over = true;
            }
        }
    }
    if (over) err[g] = kErrCmpCap;
    cmp_cnt[g] = cnt;
}

// ------------------------------------------------------------------------------------------
// Convolution as an exact dense circulant (LSbox::convolutionGeneratorFFTW, box.cpp:622-695,
// Gaussian mode).  The Gaussian symbol G(i,j) = g(i) g(j) is separable and even, so
// IDFT(G . DFT(phi)) = K_n phi K_n with the symmetric circulant
// K_n[m] = sum_k g(k) cos(2 pi k m / n), computed on the host per box size with all n taps
// (the symbol is cut at Nyquist, so K_n has a slowly decaying tail and cannot be banded).
// Rows first, then columns.  Each lane produces a 4-wide run of outputs so that one LDS read
// of the input feeds four FMAs.  kk holds K_n twice (length 2n) so no modulo is needed.
// ------------------------------------------------------------------------------------------
template <class T, int S>
__device__ __forceinline__ void circ_run(const T* __restrict__ line, int stride, int n, int x0,
                                         const T* __restrict__ kk, T* acc) {
    // acc[s] = sum_j line[j] * K[(x0 + s - j) mod n],  kk[t] = K[t mod n] for t in [0, 2n)
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = T(0);
    for (int j = 0; j < n; ++j) {
        T v = line[j * stride];
        const T* kr = kk + (x0 - j + n);
#pragma unroll
        for (int s = 0; s < S; ++s) acc[s] += kr[s] * v;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_conv_circulant(const int* __restrict__ live, const Box* __restrict__ box,
                                                        const long long* __restrict__ off, const T* __restrict__ src,
                                                        T* __restrict__ dst, const long long* __restrict__ kern_off,
                                                        const T* __restrict__ kern, int lds_bytes) {
    constexpr int S = 4;
    HIP_DYNAMIC_SHARED(unsigned char, smem)
    int g = live[blockIdx.x];
    int n = box[g].n;
    long long o = off[g];
    const T* s = src + o;
    T* d = dst + o;
    T* kk = (T*)smem;  // 2n + S entries
    const T* kg = kern + kern_off[n];
    for (int t = threadIdx.x; t < 2 * n + S; t += blockDim.x) kk[t] = kg[t % n];
    T* A = kk + ((2 * n + S + 1) & ~1);
    long long avail = (long long)lds_bytes / (long long)sizeof(T) - ((2 * n + S + 1) & ~1);
    long long nn = (long long)n * n;
    int runs = (n + S - 1) / S;  // output runs per line
    __syncthreads();
    if (2 * nn <= avail) {
        T* B = A + nn;
        for (long long p = threadIdx.x; p < nn; p += blockDim.x) A[p] = s[p];
        __syncthreads();
        for (int it = threadIdx.x; it < n * runs; it += blockDim.x) {
            int y = it / runs, x0 = (it - y * runs) * S;
            T acc[S];
            circ_run<T, S>(A + (long long)y * n, 1, n, x0, kk, acc);
#pragma unroll
            for (int q = 0; q < S; ++q)
                if (x0 + q < n) B[(long long)y * n + x0 + q] = acc[q];
        }
        __syncthreads();
        for (int it = threadIdx.x; it < n * runs; it += blockDim.x) {
            int x = it / runs, y0 = (it - x * runs) * S;
            T acc[S];
            circ_run<T, S>(B + x, n, n, y0, kk, acc);
#pragma unroll
            for (int q = 0; q < S; ++q)
                if (y0 + q < n) d[(long long)(y0 + q) * n + x] = acc[q];
        }
        return;
    }
    // strip mode: rows in strips of H rows, then columns in strips of H columns
    int H = (int)(avail / (2LL * n));
    if (H < 1) H = 1;  // the host guarantees room for two lines
    for (int y0 = 0; y0 < n; y0 += H) {
        int hh = min(H, n - y0);
        T* B = A + (long long)hh * n;
        for (long long p = threadIdx.x; p < (long long)hh * n; p += blockDim.x) A[p] = s[(long long)y0 * n + p];
        __syncthreads();
        for (int it = threadIdx.x; it < hh * runs; it += blockDim.x) {
            int y = it / runs, x0 = (it - y * runs) * S;
            T acc[S];
            circ_run<T, S>(A + (long long)y * n, 1, n, x0, kk, acc);
#pragma unroll
            for (int q = 0; q < S; ++q)
                if (x0 + q < n) B[(long long)y * n + x0 + q] = acc[q];
        }
        __syncthreads();
        for (long long p = threadIdx.x; p < (long long)hh * n; p += blockDim.x) d[(long long)y0 * n + p] = B[p];
        __syncthreads();
    }
    __threadfence();
    __syncthreads();
    for (int xs = 0; xs < n; xs += H) {
        int ww = min(H, n - xs);
        T* B = A + (long long)ww * n;
        // A[x][y] (column-major strip) = d[y][xs + x]
        for (long long p = threadIdx.x; p < (long long)ww * n; p += blockDim.x) {
            int y = (int)(p / ww), x = (int)(p - (long long)y * ww);
            A[(long long)x * n + y] = d[(long long)y * n + xs + x];
        }
        __syncthreads();
        for (int it = threadIdx.x; it < ww * runs; it += blockDim.x) {
            int x = it / runs, y0 = (it - x * runs) * S;
            T acc[S];
            circ_run<T, S>(A + (long long)x * n, 1, n, y0, kk, acc);
#pragma unroll
            for (int q = 0; q < S; ++q)
                if (y0 + q < n) B[(long long)x * n + y0 + q] = acc[q];
        }
        __syncthreads();
        for (long long p = threadIdx.x; p < (long long)ww * n; p += blockDim.x) {
            int y = (int)(p / ww), x = (int)(p - (long long)y * ww);
            d[(long long)y * n + xs + x] = B[(long long)x * n + y];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// Comparison + set-comparison, LSbox::executeComparison (box.cpp:890-943),
// boundaryCondition (968-1025), executeSetComparison (838-879).  One workgroup per grain,
// pixel-parallel.  For each pixel: max over the comparison list in list order with strict >
// (first wins ties), then the domain wall as grain 0; IDLocal keeps its stale flat value
// where nobody wins (IDChunkMinimal::clear() is a no-op, dimensionalBufferIDLocal.h:113-117;
// resize keeps the vector prefix, dimensionalBuffer.h:108-115).  The compared field
// c = (|t| < 0.7 delta) ? (t - max)/2 : t overwrites the grain's slot of the phi arena.
// ------------------------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_compare(const int* __restrict__ live, const Box* __restrict__ box,
                                                 const long long* __restrict__ off, const int* __restrict__ cmp_cnt,
                                                 const int* __restrict__ cmp, const T* __restrict__ tilde,
                                                 T* __restrict__ cfield, const uint32_t* __restrict__ idl_old,
                                                 uint32_t* __restrict__ idl_new, const long long* __restrict__ off_prev,
                                                 const int* __restrict__ n_prev, int4* __restrict__ newbox,
                                                 int* __restrict__ flags, GridConst G) {
    int g = live[blockIdx.x];
    Box b = box[g];
    long long o = off[g];
    int n = b.n;
    __shared__ int s_id[kCmpCap];
    __shared__ int s_x0[kCmpCap], s_y0[kCmpCap], s_n[kCmpCap];
    __shared__ long long s_off[kCmpCap];
    __shared__ int s_bb[4];
    int cnt = cmp_cnt[g];
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        int nb = cmp[(size_t)g * kCmpCap + k];
        Box bn = box[nb];
        s_id[k] = nb;
        s_x0[k] = bn.x0; s_y0[k] = bn.y0; s_n[k] = bn.n;
        s_off[k] = off[nb];
    }
    if (threadIdx.x == 0) {
        s_bb[0] = b.y0 + n;  // newYMin = out.maxY
        s_bb[1] = b.y0;      // newYMax = out.minY
        s_bb[2] = b.x0 + n;  // newXMin
        s_bb[3] = b.x0;      // newXMax
    }
    __syncthreads();
    // BoundaryIntersection (box.cpp:945-962) on the output buffer bounds (== box b)
    int lo = 2 * G.gb, hi = G.m - 2 * G.gb;
    bool bnd = !(b.x0 > lo && b.x0 + n < hi && b.y0 > lo && b.y0 + n < hi);
    long long nn = (long long)n * n;
    long long np = (long long)n_prev[g] * n_prev[g];
    long long op = off_prev[g];
    double band = 0.7 * G.delta;
    int ymin_l = 0x7fffffff, ymax_l = -0x7fffffff, xmin_l = 0x7fffffff, xmax_l = -0x7fffffff;
    for (long long p = threadIdx.x; p < nn; p += blockDim.x) {
        int yy = (int)(p / n), xx = (int)(p - (long long)yy * n);
        int i = b.y0 + yy, j = b.x0 + xx;
        T mx = T(-1.0);
        int id = -1;
        for (int k = 0; k < cnt; ++k) {
            int ly = i - s_y0[k], lx = j - s_x0[k];
            int nk = s_n[k];
            if ((unsigned)ly < (unsigned)nk && (unsigned)lx < (unsigned)nk) {
                T v = tilde[s_off[k] + (long long)ly * nk + lx];
                if (v > mx) { mx = v; id = s_id[k]; }
            }
        }
        if (bnd) {
            int gb = G.gb, m = G.m;
            int dXMin = -(j - gb), dYMin = -(i - gb);
            int dXMax = (j - (m - gb)), dYMax = (i - (m - gb));
            int dX = (abs(dXMin) < abs(dXMax)) ? dXMin : dXMax;
            int dY = (abs(dYMin) < abs(dYMax)) ? dYMin : dYMax;
            double dist = 0.0;
            if (dX > 0 && dY > 0) dist = sqrt((double)dX * dX + dY * dY);
            else if (dX < 0 && dY > 0) dist = dY;
            else if (dX > 0 && dY < 0) dist = dX;
            else if (dX < 0 && dY < 0) dist = max(dX, dY);
            else if (dX == 0) dist = (dY > 0) ? (double)dY : 0.0;
            else dist = (dX > 0) ? (double)dX : 0.0;  // dY == 0
            if ((T)(dist * G.h) > mx) { mx = (T)(dist * G.h); id = 0; }
        }
        uint32_t idv;
        if (id >= 0) idv = (uint32_t)id;
        else idv = (p < np) ? idl_old[op + p] : 0u;
        idl_new[o + p] = idv;
        T t = tilde[o + p];
        T c;
        if (fabs((double)t) < band) c = (T)(0.5 * ((double)t - (double)mx));
        else c = t;
        cfield[o + p] = c;
        if (c >= T(0)) {
            ymin_l = min(ymin_l, i); ymax_l = max(ymax_l, i);
            xmin_l = min(xmin_l, j); xmax_l = max(xmax_l, j);
        }
    }
    if (ymin_l != 0x7fffffff) {
        atomicMin(&s_bb[0], ymin_l); atomicMax(&s_bb[1], ymax_l);
        atomicMin(&s_bb[2], xmin_l); atomicMax(&s_bb[3], xmax_l);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int nYMin = s_bb[0] - G.gb, nYMax = s_bb[1] + G.gb, nXMin = s_bb[2] - G.gb, nXMax = s_bb[3] + G.gb;
        if (nXMin >= nXMax || nYMin >= nYMax) {
            nXMin = b.x0; nXMax = b.x0 + n; nYMin = b.y0; nYMax = b.y0 + n;
        }
        newbox[g] = make_int4(nXMin, nXMax, nYMin, nYMax);
        int f = flags[g];
        flags[g] = bnd ? (f | kBoundary) : (f & ~kBoundary);
    }
}

// ------------------------------------------------------------------------------------------
// Contour, volume, motion check, new box and direct neighbours:
// LSbox::extractContour (box.cpp:1089-1153) with MarchingSquaresAlgorithm::generateContour
// (marchingSquares.cpp:28-188), isMotionRegular (box.cpp:1155-1190), clearContourGrainArea
// (1883-1904), then computeVolumeAndEnergy (1201-1228) -> buildDirectNeighbours
// (grainBoundary.cpp:85-199).  One workgroup per grain: the inside/outside test of the box is
// kept as a bitmask in LDS so lane 0's trace never waits on memory for its direction; the
// neighbour probes run on all lanes.
// ------------------------------------------------------------------------------------------
struct ContourOut {
    double2* pts;             // contour arena
    long long* coff;          // per grain: offset into pts
    int* ccap;                // per grain: capacity reserved this step
    int* clen;                // per grain: contour length (in: old, out: new)
    double* volume;
    double* perimeter;
    double* energy;
    double* meanda;
    int* nb_cnt;
    int* nb_id;
    double* nb_len;
    Box* box_next;
    long long* off_next;
    int* flags;
    int* err;
    unsigned long long* tops;  // [0] field arena top (elements), [1] contour arena top
    long long field_cap;
    long long contour_cap;
};

enum { MS_TL = 1, MS_TR = 2, MS_BL = 4, MS_BR = 8 };
enum { D_UP = 0, D_DOWN = 1, D_LEFT = 2, D_RIGHT = 3, D_NONE = 4 };

template <class T>
struct Tracer {
    const T* c;          // compared field of the box (global)
    const uint32_t* bits;  // inside bitmask (LDS)
    int x0, y0, n;
    __device__ bool inside(int row, int col) const {  // marchingSquares.cpp:189-196
        int r = row - y0, q = col - x0;
        if ((unsigned)r >= (unsigned)n || (unsigned)q >= (unsigned)n) return false;
        int k = r * n + q;
        return (bits[k >> 5] >> (k & 31)) & 1u;
    }
    __device__ double val(int row, int col) const { return (double)c[(row - y0) * n + (col - x0)]; }
};

// marchingSquares.cpp:204-265
template <class T>
__device__ double2 ms_point(const Tracer<T>& tr, int dir, int top, int bottom, int left, int right) {
    double2 r = make_double2(-1.0, -1.0);
    double delta, position;
    int miny = tr.y0, maxy = tr.y0 + tr.n, minx = tr.x0, maxx = tr.x0 + tr.n;
    switch (dir) {
        case D_LEFT:
            r.x = left;
            if (top < miny) r.y = bottom;
            else {
                delta = tr.val(bottom, left) - tr.val(top, left);
                position = bottom - tr.val(bottom, left) / delta;
                r.y = position;
            }
            break;
        case D_RIGHT:
            r.x = right;
            if (bottom >= maxy) r.y = top;
            else {
                delta = tr.val(top, right) - tr.val(bottom, right);
                position = top + tr.val(top, right) / delta;
                r.y = position;
            }
            break;
        case D_UP:
            r.y = top;
            if (right >= maxx) r.x = left;
            else {
                delta = tr.val(top, left) - tr.val(top, right);
                position = left + tr.val(top, left) / delta;
                r.x = position;
            }
            break;
        case D_DOWN:
            r.y = bottom;
            if (left < minx) r.x = right;
            else {
                delta = tr.val(bottom, right) - tr.val(bottom, left);
                position = right - tr.val(bottom, right) / delta;
                r.x = position;
            }
            break;
        default: break;
    }
    return r;
}

struct TraceResult {
    int len;          // number of points (closed)
    int status;       // 0 no inside point, 1 ok, -1 iteration cap, -2 capacity
    double vol_sum;   // shoelace sum (grainBoundary.cpp:205-209)
    double minx, maxx, miny, maxy;
};

// MarchingSquaresAlgorithm::generateContour, run by one lane.  Points go to `out` (cap).
template <class T>
__device__ TraceResult ms_trace(const Tracer<T>& tr, int sx, int sy, double2* out, int cap) {
    TraceResult res;
    res.len = 0; res.status = 1; res.vol_sum = 0.0;
    res.minx = res.maxx = res.miny = res.maxy = 0.0;
    int left = sx - 1, right = sx, top = sy - 1, bottom = sy;
    int next = D_NONE, prev = D_NONE;
    long long itcap = 8LL * (tr.n + 2) * (tr.n + 2) + 16;
    long long it = 0;
    double2 last = make_double2(0, 0);
    auto push = [&](double2 p) {  // insertPoint (marchingSquares.cpp:197-203)
        if (res.len == 0 || p.x != last.x || p.y != last.y) {
            if (res.len >= cap) { res.status = -2; return; }
            if (res.len > 0) res.vol_sum += (last.y + p.y) * (last.x - p.x);
            if (res.len == 0) { res.minx = res.maxx = p.x; res.miny = res.maxy = p.y; }
            else {
                res.minx = fmin(res.minx, p.x); res.maxx = fmax(res.maxx, p.x);
                res.miny = fmin(res.miny, p.y); res.maxy = fmax(res.maxy, p.y);
            }
            out[res.len++] = p;
            last = p;
        }
    };
    do {
        int mask = 0;
        if (tr.inside(top, left)) mask |= MS_TL;
        if (tr.inside(top, right)) mask |= MS_TR;
        if (tr.inside(bottom, left)) mask |= MS_BL;
        if (tr.inside(bottom, right)) mask |= MS_BR;
        bool gen = true;
        switch (mask) {
            case MS_TL: next = D_UP; break;
            case MS_TR: next = D_RIGHT; break;
            case MS_TL | MS_TR: next = D_RIGHT; break;
            case MS_BL: next = D_LEFT; break;
            case MS_BL | MS_TL: next = D_UP; break;
            case MS_BL | MS_TR: next = (prev == D_UP) ? D_RIGHT : D_LEFT; gen = false; break;
            case MS_BL | MS_TR | MS_TL: next = D_RIGHT; break;
            case MS_BR: next = D_DOWN; break;
            case MS_BR | MS_TL: next = (prev == D_RIGHT) ? D_DOWN : D_UP; gen = false; break;
            case MS_BR | MS_TR: next = D_DOWN; break;
            case MS_BR | MS_TR | MS_TL: next = D_DOWN; break;
            case MS_BR | MS_BL: next = D_LEFT; break;
            case MS_BR | MS_BL | MS_TL: next = D_UP; break;
            case MS_BR | MS_BL | MS_TR: next = D_LEFT; break;
            default: gen = false; break;
        }
        if (gen) {
            push(ms_point(tr, next, top, bottom, left, right));
            if (res.status < 0) return res;
        }
        switch (next) {
            case D_UP: top--; bottom--; break;
            case D_DOWN: top++; bottom++; break;
            case D_LEFT: left--; right--; break;
            case D_RIGHT: left++; right++; break;
            default: break;
        }
        prev = next;
        if (++it > itcap) { res.status = -1; return res; }
    } while (left != sx - 1 || bottom != sy);
    if (res.len > 0) {
        double2 f = out[0];
        if (f.x != last.x || f.y != last.y) {
            if (res.len >= cap) { res.status = -2; return res; }
            res.vol_sum += (last.y + f.y) * (last.x - f.x);
            out[res.len++] = f;
            last = f;
        }
    }
    return res;
}

template <class T>
__global__ void __launch_bounds__(256) k_contour(const int* __restrict__ live, const Box* __restrict__ box,
                                                 const long long* __restrict__ off, T* __restrict__ cfield,
                                                 const uint32_t* __restrict__ idl, const int4* __restrict__ newbox,
                                                 ContourOut O, GridConst G, int loop, int analysis_timestep) {
    HIP_DYNAMIC_SHARED(unsigned char, g_smem)
    int g = live[blockIdx.x];
    Box b = box[g];
    int n = b.n;
    long long o = off[g];
    T* c = cfield + o;
    long long nn = (long long)n * n;
    int nwords = (int)((nn + 31) >> 5);
    uint32_t* bits = (uint32_t*)g_smem;
    int ccap = 16 * n + 64;
    {
        long long c2 = 2LL * (n + 1) * (n + 1) + 2;
        if (c2 < ccap) ccap = (int)c2;
    }
    int* s_cand = (int*)(g_smem + (size_t)nwords * 4);
    double* s_len = (double*)(g_smem + (((size_t)nwords * 4 + (size_t)ccap * 4 + 15) & ~(size_t)15));
    __shared__ int s_start;
    __shared__ int s_state;       // 0 trace, 1 clear+retrace, 2 done
    __shared__ TraceResult s_tr;
    __shared__ long long s_coff;
    __shared__ int s_err;
    __shared__ int s_regular;
    __shared__ int s_exists;
    __shared__ int s_clear[4];    // rect for clearContourGrainArea

    // reserve contour space
    if (threadIdx.x == 0) {
        unsigned long long co = atomicAdd(&O.tops[1], (unsigned long long)ccap);
        s_coff = (long long)co;
        s_err = (co + (unsigned long long)ccap > (unsigned long long)O.contour_cap) ? kErrContourCap : 0;
        s_start = 0x7fffffff;
        s_regular = 1;
        s_exists = 1;
    }
    // inside bitmask, value >= 0 (marchingSquares.cpp:195)
    for (int w = threadIdx.x; w < nwords; w += blockDim.x) {
        uint32_t word = 0;
        for (int bit = 0; bit < 32; ++bit) {
            long long k = ((long long)w << 5) + bit;
            if (k < nn && c[k] >= T(0)) word |= (1u << bit);
        }
        bits[w] = word;
    }
    __syncthreads();
    if (s_err) {
        if (threadIdx.x == 0) { O.err[g] = s_err; }
        return;
    }
    double2* pts = O.pts + s_coff;
    Tracer<T> tr{c, bits, b.x0, b.y0, n};
    int old_len = O.clen[g];
    double old_vol = O.volume[g];
    int retries = 0;
    while (true) {
        // first inside pixel in row-major order (marchingSquares.cpp:32-42)
        if (threadIdx.x == 0) s_start = 0x7fffffff;
        __syncthreads();
        int loc = 0x7fffffff;
        for (int w = threadIdx.x; w < nwords; w += blockDim.x) {
            uint32_t word = bits[w];
            if (word) { loc = (w << 5) + __ffs(word) - 1; break; }
        }
        if (loc != 0x7fffffff) atomicMin(&s_start, loc);
        __syncthreads();
        if (threadIdx.x == 0) {
            if (s_start == 0x7fffffff) {
                s_exists = 0;
                s_state = 2;
            } else {
                int sy = b.y0 + s_start / n, sx = b.x0 + s_start % n;
                TraceResult r = ms_trace(tr, sx, sy, pts, ccap);
                s_tr = r;
                if (r.status == -1) { s_err = kErrTrace; s_exists = 0; s_state = 2; }
                else if (r.status == -2) { s_err = kErrContourCap; s_exists = 0; s_state = 2; }
                else {
                    double newVolume = fabs(r.vol_sum) * G.h * G.h * 0.5;
                    // isMotionRegular (box.cpp:1155-1190)
                    double mullins = (newVolume - old_vol) / G.dt * (3 / 3.14159265358979323846);
                    double contourRatio = ((double)r.len) / old_len;
                    double volumeRatio = newVolume / old_vol;
                    int regular;
                    if (mullins > -6.0) regular = 1;
                    else if (volumeRatio < 0.10) regular = 0;
                    else if (contourRatio < 0.35) regular = 0;
                    else if (contourRatio > 0.7) regular = 1;
                    else { s_regular = 0; regular = 1; }
                    if (regular) s_state = 2;
                    else {
                        s_state = 1;
                        if (r.len == 0) s_state = 1;
                        s_clear[0] = (int)(r.miny - 1);
                        s_clear[1] = (int)(r.minx - 1);
                    }
                }
            }
        }
        __syncthreads();
        if (s_state == 2) break;
        if (++retries > 256) {
            if (threadIdx.x == 0) { s_err = kErrTrace; s_exists = 0; }
            __syncthreads();
            break;
        }
        // clearContourGrainArea: i in [int(miny-1), maxy+1), j in [int(minx-1), maxx+1)
        {
            int i0 = s_clear[0], j0 = s_clear[1];
            double maxy = s_tr.maxy, maxx = s_tr.maxx;
            int ih = 0, jw = 0;
            while (i0 + ih < maxy + 1) ih++;
            while (j0 + jw < maxx + 1) jw++;
            int bad = 0;
            for (int p = threadIdx.x; p < ih * jw; p += blockDim.x) {
                int ii = i0 + p / jw, jj = j0 + p % jw;
                long long k = flat_index(ii, jj, b.x0, b.y0, n, nn);
                if (k < 0) { bad = 1; continue; }
                if (c[k] >= T(0)) {
                    c[k] = (T)(-G.h);
                    atomicAnd(&bits[k >> 5], ~(1u << (k & 31)));
                }
            }
            if (bad) atomicOr(&s_err, kErrClearBounds);
            __threadfence_block();
            __syncthreads();
            if (s_err) {
                if (threadIdx.x == 0) s_exists = 0;
                __syncthreads();
                break;
            }
        }
    }
    __syncthreads();
    int exists = s_exists && !s_err;
    if (threadIdx.x == 0) {
        int f = O.flags[g];
        f = s_regular ? (f | kRegular) : (f & ~kRegular);
        if (!exists) f &= ~kExists;
        O.flags[g] = f;
        if (s_err) O.err[g] = s_err;
        O.box_next[g] = b;
        O.coff[g] = s_coff;
        O.ccap[g] = ccap;
    }
    if (!exists) {
        if (threadIdx.x == 0) O.clen[g] = (s_start == 0x7fffffff) ? O.clen[g] : s_tr.len;
        return;
    }
    int len = s_tr.len;
    if (threadIdx.x == 0) {
        // bounds clamp (box.cpp:1118-1135) and out->resize + resizeToSquare (1145-1148)
        int4 nb = newbox[g];
        int nXMin = nb.x, nXMax = nb.y, nYMin = nb.z, nYMax = nb.w;
        int m = G.m;
        if (nXMin < 0) nXMin = 0;
        if (nYMin < 0) nYMin = 0;
        if (nXMax > m) nXMax = m;
        if (nYMax > m) nYMax = m;
        int xmin = nXMin, xmax = nXMax, ymin = nYMin, ymax = nYMax;
        int grid_size = m - 1;
        unsigned int height = (unsigned int)(ymax - ymin), width = (unsigned int)(xmax - xmin);
        if (width != height) {
            if (width > height) { int diff = (int)(width - height); ymin -= diff / 2; ymax += diff / 2 + diff % 2; }
            else { int diff = (int)(height - width); xmin -= diff / 2; xmax += diff / 2 + diff % 2; }
            if (xmin < 0) { int dl = -xmin; xmax += dl; xmin += dl; }
            else if (xmax > grid_size) { int dl = grid_size - xmax; xmax += dl; xmin += dl; }
            if (ymin < 0) { int dl = -ymin; ymax += dl; ymin += dl; }
            else if (ymax > grid_size) { int dl = grid_size - ymax; ymax += dl; ymin += dl; }
        }
        int n2 = xmax - xmin;
        Box nbx;
        nbx.x0 = xmin; nbx.y0 = ymin; nbx.n = n2; nbx.pad = ymax - ymin;
        long long need = ((long long)n2 * n2 + kAlign - 1) / kAlign * kAlign;
        unsigned long long fo = atomicAdd(&O.tops[0], (unsigned long long)need);
        if ((long long)(fo + need) > O.field_cap) O.err[g] = kErrArena;
        O.box_next[g] = nbx;
        O.off_next[g] = (long long)fo;
        O.clen[g] = len;
    }
    // perimeter (grainBoundary.cpp:224-231) and segment lengths, all lanes
    __threadfence_block();
    __syncthreads();
    for (int i = threadIdx.x; i < len - 1; i += blockDim.x) {
        double2 a = pts[i], bb = pts[i + 1];
        double dx = a.x - bb.x, dy = a.y - bb.y;
        s_len[i] = sqrt(dx * dx + dy * dy);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double per = 0;
        for (int i = 0; i < len - 1; ++i) per += s_len[i] * G.h;
        O.perimeter[g] = per;
    }
    if (!s_regular) return;  // computeVolumeAndEnergy returns early (box.cpp:1202-1203)
    // buildDirectNeighbours probes (grainBoundary.cpp:96-120), all lanes
    for (int i = threadIdx.x; i < len - 1; i += blockDim.x) {
        double2 p = pts[i];
        int pxGrid = (int)p.x, pyGrid = (int)p.y;
        int cand = -1;
        bool ok = true;
        long long k = flat_index(pyGrid, pxGrid, b.x0, b.y0, n, nn);
        if (k < 0) ok = false;
        else if (c[k] > T(0)) {
            pxGrid = (int)(p.x + 1);
            k = flat_index(pyGrid, pxGrid, b.x0, b.y0, n, nn);
            if (k < 0) ok = false;
            else if (c[k] > T(0)) {
                pyGrid = (int)(p.y + 1);
                k = flat_index(pyGrid, pxGrid, b.x0, b.y0, n, nn);
                if (k < 0) ok = false;
                else if (c[k] > T(0)) pxGrid = (int)p.x;
            }
        }
        if (ok) {  // LSbox::getNeighbourAt (box.cpp:2190-2195), getGrainByID (grainhdl.h:153-160)
            if (pyGrid < 0 || pyGrid >= b.y0 + n || pxGrid < 0 || pxGrid >= b.x0 + n) ok = false;
            else {
                k = flat_index(pyGrid, pxGrid, b.x0, b.y0, n, nn);
                if (k < 0) ok = false;
                else {
                    uint32_t nid = idl[o + k];
                    if (nid == 0u) cand = 0;
                    else if (nid < (uint32_t)G.nslots && (O.flags[nid] & kPresent)) cand = (int)nid;
                    else ok = false;
                }
            }
        }
        s_cand[i] = ok ? cand : -1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int cnt = 0;
        int* ids = O.nb_id + (size_t)g * kNbCap;
        double* lens = O.nb_len + (size_t)g * kNbCap;
        int stop = len - 1;
        int over = 0;
        for (int i = 0; i < len - 1; ++i) {
            int cd = s_cand[i];
            if (cd < 0) { stop = i; break; }
            double ll = s_len[i] * G.h;
            int e = 0;
            for (; e < cnt; ++e)
                if (ids[e] == cd) break;
            if (e == cnt) {
                if (cnt < kNbCap) { ids[cnt] = cd; lens[cnt] = 0.0; cnt++; }
                else { over = 1; e = -1; }
            }
            if (e >= 0) lens[e] += ll;
        }
        O.nb_cnt[g] = cnt;
        if (over) O.err[g] = kErrNbCap;
        if (stop < len - 1) {  // markAsInvalidMotion (box.cpp:1939-1945)
            O.flags[g] &= ~kRegular;
        }
        // computeEnergy (grainBoundary.cpp:215-222): energy 1 on processed vertices, -1 on the
        // rest (SPoint(-1,-1,-1,-1) from generatePoint)
        double e = 0;
        for (int i = 0; i < len - 1; ++i) e += s_len[i] * (i < stop ? 1.0 : -1.0) * G.h;
        O.energy[g] = e;
        double newVolume = fabs(fabs(s_tr.vol_sum) * G.h * G.h * 0.5);
        double dA = newVolume - old_vol;
        dA = dA / G.dt;
        dA = dA * (3 / 3.14159265358979323846);
        O.volume[g] = newVolume;
        unsigned int lp = (unsigned int)loop;
        unsigned long at = analysis_timestep > 0 ? (unsigned long)analysis_timestep : 1ul;
        double md = O.meanda[g];
        if ((unsigned long)(lp - 1u) % at == 0 || lp == 0) md = 0;
        O.meanda[g] = md + dA;
    }
}

// ------------------------------------------------------------------------------------------
// Redistancing, LSbox::executeRedistancing (box.cpp:1259-1399): out <- -1 on the new box,
// four sequential sweeps (y forward with zero-crossing interpolation, y backward, x forward
// with interpolation, x backward) driven by the compared field on the old box, clamp to
// [-delta, delta].  One workgroup per grain: lanes own columns (y sweeps) or rows (x sweeps),
// the output tile lives in LDS with an odd row pitch.
// ------------------------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_redist(const int* __restrict__ live, const Box* __restrict__ box,
                                                const long long* __restrict__ off, const T* __restrict__ cfield,
                                                const Box* __restrict__ box_next, const long long* __restrict__ off_next,
                                                const int* __restrict__ flags, T* __restrict__ out_field,
                                                GridConst G, int lds_bytes) {
    HIP_DYNAMIC_SHARED(unsigned char, g_smem)
    int g = live[blockIdx.x];
    if (!(flags[g] & kExists)) return;
    Box bi = box[g], bo = box_next[g];
    const T* in = cfield + off[g];
    T* gout = out_field + off_next[g];
    int ni = bi.n, no = bo.n;
    int pitch = no | 1;
    bool lds = (long long)pitch * no * (long long)sizeof(T) <= (long long)lds_bytes;
    T* o = lds ? (T*)g_smem : gout;
    if (!lds) pitch = no;
    double h = G.h;
    // output buffer window [oxmin, oxmax) x [oymin, oymax); input window likewise
    int oxmin = bo.x0, oxmax = bo.x0 + no, oymin = bo.y0, oymax = bo.y0 + no;
    int ixmn = bi.x0, ixmx = bi.x0 + ni, iymn = bi.y0, iymx = bi.y0 + ni;
    int ixmin = ixmn < oxmin ? oxmin : ixmn;
    int iymin = iymn < oymin ? oymin : iymn;
    int ixmax = ixmx < oxmax ? ixmx : oxmax;
    int iymax = iymx < oymax ? iymx : oymax;
#define OV(i, j) o[((i) - oymin) * pitch + ((j) - oxmin)]
#define IV(i, j) ((double)in[((i) - iymn) * ni + ((j) - ixmn)])
    for (int p = threadIdx.x; p < no * no; p += blockDim.x) {
        int yy = p / no, xx = p - yy * no;
        o[yy * pitch + xx] = T(-1.0);
    }
    __syncthreads();
    // y-direction forward (box.cpp:1291-1319)
    for (int j = ixmin + threadIdx.x; j < oxmax; j += blockDim.x) {
        for (int i = iymin; i < oymax - 1; i++) {
            double cand;
            if (j < ixmax && i < iymax - 1) {
                double a = IV(i, j), bnext = IV(i + 1, j);
                if (a * bnext <= 0.0) {
                    double slope = (bnext - a) / h;
                    double d0 = -a / slope;
                    if (fabs((double)OV(i, j)) > fabs(d0)) OV(i, j) = (T)(-d0 * sgn_of(slope));
                }
                cand = (double)OV(i, j) + (sgn_of(in[(i + 1 - iymn) * ni + (j - ixmn)]) * h);
            } else {
                cand = (double)OV(i, j) + (sgn_of(OV(i + 1, j)) * h);
            }
            if (fabs(cand) < fabs((double)OV(i + 1, j))) OV(i + 1, j) = (T)cand;
        }
    }
    __syncthreads();
    // y-direction backward (1321-1338)
    for (int j = ixmin + threadIdx.x; j < oxmax; j += blockDim.x) {
        for (int i = iymax - 1; i > oymin; i--) {
            double cand;
            if (j < ixmax && i > iymin) cand = (double)OV(i, j) + (sgn_of(in[(i - 1 - iymn) * ni + (j - ixmn)]) * h);
            else cand = (double)OV(i, j) + (sgn_of(OV(i - 1, j)) * h);
            if (fabs(cand) < fabs((double)OV(i - 1, j))) OV(i - 1, j) = (T)cand;
        }
    }
    __syncthreads();
    // x-direction forward (1339-1367)
    for (int i = iymin + threadIdx.x; i < oymax; i += blockDim.x) {
        for (int j = ixmin; j < oxmax - 1; j++) {
            double cand;
            if (j < ixmax - 1 && i < iymax) {
                double a = IV(i, j), bnext = IV(i, j + 1);
                if (a * bnext <= 0.0) {
                    double slope = (bnext - a) / h;
                    double d0 = -a / slope;
                    if (fabs((double)OV(i, j)) > fabs(d0)) OV(i, j) = (T)(-d0 * sgn_of(slope));
                }
                cand = (double)OV(i, j) + (sgn_of(in[(i - iymn) * ni + (j + 1 - ixmn)]) * h);
            } else {
                cand = (double)OV(i, j) + (sgn_of(OV(i, j + 1)) * h);
            }
            if (fabs(cand) < fabs((double)OV(i, j + 1))) OV(i, j + 1) = (T)cand;
        }
    }
    __syncthreads();
    // x-direction backward (1369-1388)
    for (int i = iymin + threadIdx.x; i < oymax; i += blockDim.x) {
        for (int j = ixmax - 1; j > oxmin; j--) {
            double cand;
            if (j > ixmin && i < iymax) cand = (double)OV(i, j) + (sgn_of(in[(i - iymn) * ni + (j - 1 - ixmn)]) * h);
            else cand = (double)OV(i, j) + sgn_of(OV(i, j - 1)) * h;
            if (fabs(cand) < fabs((double)OV(i, j - 1))) OV(i, j - 1) = (T)cand;
        }
    }
    __syncthreads();
#undef OV
#undef IV
    // clampValues (dimensionalBufferReal.h:43-56)
    T lo = (T)(-G.delta), hi = (T)G.delta;
    for (int p = threadIdx.x; p < no * no; p += blockDim.x) {
        int yy = p / no, xx = p - yy * no;
        T v = o[yy * pitch + xx];
        if (v < lo) v = lo;
        else if (v > hi) v = hi;
        gout[p] = v;
    }
}

// ------------------------------------------------------------------------------------------
// Initial signed distance, LSbox::calculateDistanceFunction (box.cpp:342-401): crossing-number
// point-in-polygon and the minimum point-segment distance, clamped to delta.  One workgroup
// per grain, one lane per pixel.  Contraction is off so the values match a scalar CPU build.
// ------------------------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_init_sdf(const int* __restrict__ live, const Box* __restrict__ box,
                                                  const long long* __restrict__ off, const int* __restrict__ voff,
                                                  const int* __restrict__ vcnt, const double2* __restrict__ verts,
                                                  T* __restrict__ field, GridConst G) {
#pragma clang fp contract(off)
    int g = live[blockIdx.x];
    Box b = box[g];
    int n = b.n;
    const double2* cg = verts + voff[g];
    int cs = vcnt[g];
    T* o = field + off[g];
    double h = G.h;
    int gb = G.gb;
    for (int p = threadIdx.x; p < n * n; p += blockDim.x) {
        int yy = p / n, xx = p - yy * n;
        int i = b.y0 + yy, j = b.x0 + xx;
        double tx = (j - gb) * h, ty = (i - gb) * h;
        bool inside = false;
        for (int k = 1, l = 0; k < cs; k++) {
            double2 L = cg[l], K = cg[k];
            if (((L.y > ty) != (K.y > ty)) && (tx < (K.x - L.x) * (ty - L.y) / (K.y - L.y) + L.x)) inside = !inside;
            l = k;
        }
        double minDist = 1000000.0;
        for (int k = 1, l = 0; k < cs; k++) {
            double2 L = cg[l], K = cg[k];
            double ux = K.x - L.x, uy = K.y - L.y;
            double lambda = (tx - L.x) * ux + (ty - L.y) * uy;
            lambda /= ux * ux + uy * uy;
            double dist;
            if (lambda <= 0) {
                double dx = tx - L.x, dy = ty - L.y;
                dist = sqrt(dx * dx + dy * dy);
            } else if (lambda >= 1) {
                double dx = K.x - tx, dy = K.y - ty;
                dist = sqrt(dx * dx + dy * dy);
            } else {
                double px = L.x + ux * lambda, py = L.y + uy * lambda;
                double dx = tx - px, dy = ty - py;
                dist = sqrt(dx * dx + dy * dy);
            }
            minDist = (dist < minDist) ? dist : minDist;
            l = k;
        }
        if (minDist > G.delta) minDist = G.delta;
        o[p] = (T)(inside ? minDist : -minDist);
    }
}

}  // namespace ggx

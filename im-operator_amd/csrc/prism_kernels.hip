// prism_kernels.hip -- PRISM edge-directed demosaic (src/algorithms/prism.cpp:
// 25-316) and RCD (src/algorithms/rcd.cpp:7-248) on gfx950 in two LDS-tiled
// passes.  RCD's stages are PRISM's S0-S6 without the Laplacian term of the
// green estimate (prism.cpp:132-136), and it has no colour-difference median
// (S7-S9): pass B sends S6's planes straight to the border fill.
//
//   k_prism_a  raw -> CFA (decode + black level, S0), VH map (S1), low-pass
//              (S2), PQ map (S4) and green at R/B (S3); writes four FP32
//              planes: cfa, vh, pq, c1.
//   k_prism_b  those planes -> R/B at B/R (S5), R and B at G (S6), the 3x3
//              colour-difference medians (S7-S9), the width-4 border fill,
//              rounding and the interleaved RGB write (S10, S11).
//
// Each workgroup owns a 32 x 16 tile of output pixels and computes every
// intermediate it needs over the tile plus an apron in LDS: S3 reads CFA up
// to 5 pixels away (VH at +-1 of a +-4 window), and S5 -> S6 -> median reach
// 7 pixels into the four planes.  The last tile of each row / column is
// shifted inward so that every tile lies inside the frame; the width-4 border
// fill then always finds its source pixel (at most 4 pixels inward) in the
// same tile, and overlapping tiles write identical values.
//
// Float reproducibility: every stage keeps the reference's operation order
// (the stage functions take accessors, so the same expressions serve LDS and
// global planes).  The reference library is compiled by GCC with -mavx2
// -mfma, which fuses the first product of each a*b + c*d sum into one
// multiply-add; those sites use fmaf() here and every other operation is a
// separately rounded IEEE op (contraction is off for this file; division is
// correctly rounded).  Regions where the reference leaves a plane at its
// initial value (VH = PQ = 0.5, low-pass 0, chroma natives / 0 outside
// [4, h-4) x [4, w-4)) keep those values.
#include "device_common.hpp"
#include "median9.hpp"
#include "plain.hpp"
#include "window_decode.hpp"

#pragma clang fp contract(off)

namespace imop_hip {

namespace {

constexpr float kEps = 1e-5f;
constexpr float kEpsSq = 1e-10f;

constexpr int TW = 32, TH = 16;  // output tile (measured best of 128x16, 64x16, 64x8, 32x32, 32x16)

struct PrismPlanes {
    float *cfa, *c1, *vh, *pq;
};

__host__ __device__ inline size_t plane_stride(int w, int h) {
    size_t n = (size_t)w * (size_t)h;
    return (n + 63u) & ~(size_t)63u;  // 256-byte aligned planes
}

__device__ __forceinline__ PrismPlanes planes_of(void* ws, int w, int h) {
    const size_t s = plane_stride(w, h);
    float* b = (float*)ws;
    return PrismPlanes{b, b + s, b + 2 * s, b + 3 * s};
}

// std::min / std::max argument order: first argument wins ties.
__device__ __forceinline__ float fmin_first(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax_first(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float clamp01(float v) { return fmax_first(0.0f, fmin_first(1.0f, v)); }

__device__ __forceinline__ bool inside(int x, int y, int w, int h, int b) {
    return y >= b && y < h - b && x >= b && x < w - b;
}

// Tile of this workgroup and its origin.  The 1D grid is remapped so that
// each of the 8 XCDs walks a contiguous raster run of tiles (workgroups are
// dealt round-robin to the XCDs; with nblk = 8q + r, XCD x runs q + (x < r)
// of them): neighbouring tiles, which re-read each other's aprons, then share
// one L2.  The last tile of a row / column is shifted inward.
__device__ __forceinline__ void tile_origin(int w, int h, int& x0, int& y0) {
    const int nblk = gridDim.x;
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int t = xcd * (nblk >> 3) + imin(xcd, nblk & 7) + slot;
    const int gx = (w + TW - 1) / TW;
    const int ty = t / gx, tx = t - ty * gx;
    x0 = imin(tx * TW, imax(w - TW, 0));
    y0 = imin(ty * TH, imax(h - TH, 0));
}

// k-th site of one class (rb: R/B sites, else G sites) in a region of RW
// (even) x RH pixels whose top-left pixel is (xr, yr): every row holds RW / 2
// of each class at alternating columns, so the threads of a wave all run the
// same stage instead of half of them idling at the other class's sites.
__device__ __forceinline__ void nth_site(int k, int RW, int xr, int yr, int pattern, bool rb, int& i, int& j) {
    const int half = RW >> 1;
    j = k / half;
    const int m = k - j * half;
    const int y = yr + j;
    const int prb = ((y & 1) == r_row_of(pattern)) ? r_col_of(pattern) : 1 - r_col_of(pattern);
    const int par = rb ? prb : 1 - prb;
    i = 2 * m + ((par - xr) & 1);
}

// ------------------------------------------------------------ stage bodies
// C(dx, dy): CFA at the pixel's offset; likewise the other accessors.

// S1 (prism.cpp:84-99), inside [4, h-4) x [4, w-4)
template <class Cf>
__device__ __forceinline__ float stage_vh(Cf C) {
    const float c = C(0, 0);
    float vs = 0.0f, hs = 0.0f;
    for (int k = -4; k <= 4; ++k) {
        vs += fabsf(c - C(0, k));
        hs += fabsf(c - C(k, 0));
    }
    vs = fmax_first(vs, kEpsSq);
    hs = fmax_first(hs, kEpsSq);
    return vs / (vs + hs);
}

// S2 (prism.cpp:101-112), R/B sites inside [2, h-2) x [2, w-2)
template <class Cf>
__device__ __forceinline__ float stage_lpf(Cf C) {
    float s4 = C(0, -1) + C(0, 1) + C(-1, 0) + C(1, 0);
    float d4 = C(-1, -1) + C(1, -1) + C(-1, 1) + C(1, 1);
    return fmaf(0.0625f, d4, fmaf(0.25f, C(0, 0), 0.125f * s4));
}

// S4 (prism.cpp:165-179), R/B sites inside [4, h-4) x [4, w-4)
template <class Cf>
__device__ __forceinline__ float stage_pq(Cf C) {
    const float c = C(0, 0);
    float ps = 0.0f, qs = 0.0f;
    for (int k = -4; k <= 4; ++k) {
        ps += fabsf(c - C(k, k));
        qs += fabsf(c - C(-k, k));
    }
    ps = fmax_first(ps, kEpsSq);
    qs = fmax_first(qs, kEpsSq);
    return ps / (ps + qs);
}

// S3: green at R/B sites (prism.cpp:114-163; rcd.cpp:88-127 without the
// Laplacian), inside [4, h-4) x [4, w-4)
template <bool RCD, class Cf, class Vf, class Lf>
__device__ __forceinline__ float stage_s3(Cf C, Vf V, Lf L) {
    float vc = V(0, 0);
    float vn = 0.25f * (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1));
    float disc = (fabsf(0.5f - vc) < fabsf(0.5f - vn)) ? vn : vc;
    float c0 = C(0, 0);
    float gN = kEps + fabsf(C(0, -1) - C(0, 1)) + fabsf(c0 - C(0, -2)) + fabsf(C(0, -1) - C(0, -3));
    float gS = kEps + fabsf(C(0, 1) - C(0, -1)) + fabsf(c0 - C(0, 2)) + fabsf(C(0, 1) - C(0, 3));
    float gW = kEps + fabsf(C(-1, 0) - C(1, 0)) + fabsf(c0 - C(-2, 0)) + fabsf(C(-1, 0) - C(-3, 0));
    float gE = kEps + fabsf(C(1, 0) - C(-1, 0)) + fabsf(c0 - C(2, 0)) + fabsf(C(1, 0) - C(3, 0));
    float l0 = L(0, 0), lN = L(0, -2), lS = L(0, 2), lW = L(-2, 0), lE = L(2, 0);
    float eN = C(0, -1) * (1.0f + (l0 - lN) / (kEps + l0 + lN));
    float eS = C(0, 1) * (1.0f + (l0 - lS) / (kEps + l0 + lS));
    float eW = C(-1, 0) * (1.0f + (l0 - lW) / (kEps + l0 + lW));
    float eE = C(1, 0) * (1.0f + (l0 - lE) / (kEps + l0 + lE));
    float vEst = fmaf(gS, eN, gN * eS) / (gN + gS);
    float hEst = fmaf(gW, eE, gE * eW) / (gE + gW);
    if (!RCD) {
        float lapV = (2.0f * c0 - C(0, -2) - C(0, 2)) * 0.25f;
        float lapH = (2.0f * c0 - C(-2, 0) - C(2, 0)) * 0.25f;
        vEst += lapV;
        hEst += lapH;
    }
    return clamp01(fmaf(disc, hEst, (1.0f - disc) * vEst));
}

// S5: the opposite chroma at an R/B site (prism.cpp:181-225), inside
// [4, h-4) x [4, w-4).  X = the opposite colour's plane, G = green.
template <class Xf, class Gf, class Qf>
__device__ __forceinline__ float stage_s5(Xf X, Gf G, Qf Q) {
    float pc = Q(0, 0);
    float pn = 0.25f * (Q(-1, -1) + Q(1, -1) + Q(-1, 1) + Q(1, 1));
    float disc = (fabsf(0.5f - pc) < fabsf(0.5f - pn)) ? pn : pc;
    float g0 = G(0, 0);
    float gNW = kEps + fabsf(X(-1, -1) - X(1, 1)) + fabsf(g0 - G(-2, -2)) + fabsf(X(-1, -1) - X(-3, -3));
    float gNE = kEps + fabsf(X(1, -1) - X(-1, 1)) + fabsf(g0 - G(2, -2)) + fabsf(X(1, -1) - X(3, -3));
    float gSW = kEps + fabsf(X(-1, 1) - X(1, -1)) + fabsf(g0 - G(-2, 2)) + fabsf(X(-1, 1) - X(-3, 3));
    float gSE = kEps + fabsf(X(1, 1) - X(-1, -1)) + fabsf(g0 - G(2, 2)) + fabsf(X(1, 1) - X(3, 3));
    float eNW = X(-1, -1) - G(-1, -1);
    float eNE = X(1, -1) - G(1, -1);
    float eSW = X(-1, 1) - G(-1, 1);
    float eSE = X(1, 1) - G(1, 1);
    float pEst = fmaf(gNW, eSE, gSE * eNW) / (gNW + gSE);
    float qEst = fmaf(gNE, eSW, gSW * eNE) / (gNE + gSW);
    return clamp01(fmaf(disc, qEst, fmaf(1.0f - disc, pEst, g0)));
}

// S6: one chroma (X) at a G site (prism.cpp:227-264), inside
// [4, h-4) x [4, w-4); disc from the VH map.
template <class Xf, class Gf>
__device__ __forceinline__ float stage_s6(Xf X, Gf G, float disc) {
    float g0 = G(0, 0);
    float gN = kEps + fabsf(g0 - G(0, -2)) + fabsf(X(0, -1) - X(0, 1)) + fabsf(X(0, -1) - X(0, -3));
    float gS = kEps + fabsf(g0 - G(0, 2)) + fabsf(X(0, 1) - X(0, -1)) + fabsf(X(0, 1) - X(0, 3));
    float gW = kEps + fabsf(g0 - G(-2, 0)) + fabsf(X(-1, 0) - X(1, 0)) + fabsf(X(-1, 0) - X(-3, 0));
    float gE = kEps + fabsf(g0 - G(2, 0)) + fabsf(X(1, 0) - X(-1, 0)) + fabsf(X(1, 0) - X(3, 0));
    float dN = X(0, -1) - G(0, -1);
    float dS = X(0, 1) - G(0, 1);
    float dW = X(-1, 0) - G(-1, 0);
    float dE = X(1, 0) - G(1, 0);
    float vEst = fmaf(gS, dN, gN * dS) / (gN + gS);
    float hEst = fmaf(gW, dE, gE * dW) / (gE + gW);
    return clamp01(fmaf(1.0f - disc, vEst, fmaf(disc, hEst, g0)));
}


// LDS plane of a tile with apron A: element (x, y) of the frame.
template <int A>
struct Tile {
    static constexpr int W = TW + 2 * A, H = TH + 2 * A;
};

// (float)v / m rounded to nearest for an integer sample v in [0, 65535] and
// m = 2^bd - 1, bd in 1..16, from y = 1 / m rounded to nearest: the product
// and one fused remainder correction.  Checked against the correctly rounded
// quotient for every such v and m on the CPU (tests/cfa_norm.c).
__device__ __forceinline__ float norm_sample(int v, float m, float y) {
    const float a = (float)v;
    const float q0 = a * y;
    return fmaf(fmaf(-m, q0, a), y, q0);
}

// S0 over the RW x RH window whose top-left frame position is (xs, ys), 0
// outside the frame: windows inside the frame and above its last row take
// the per-layout rows of window_inside; the others, and the GigE and generic
// bit layouts, sample_at per pixel.
template <int RW, int RH>
__device__ __forceinline__ void s0_window(const FrameParams& f, float* dst, int xs, int ys, float scale, float rscale) {
    if (window_inside<RW, RH>(f, xs, ys, [&](int j, int c, int v) { dst[j * RW + c] = norm_sample(v, scale, rscale); }))
        return;
    const int w = f.w, h = f.h;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane >= RW) return;
    const int x = xs + lane;
    const bool xok = (unsigned)x < (unsigned)w;
    for (int j = wv; j < RH; j += 4) {
        const int y = ys + j;
        dst[j * RW + lane] = xok && (unsigned)y < (unsigned)h ? norm_sample(sample_at(f, x, y), scale, rscale) : 0.0f;
    }
}

// LDS copy (row pitch RW) of the RW x RH window of a frame-sized plane whose
// top-left frame position is (xs, ys); 0 outside the frame.  Each wave takes
// two rows per step (lanes 0-31 the first, 32-63 the second), each lane two
// consecutive columns: no index division, and the bounds checks only in
// windows that cross the frame edge.  RW even, at most 64; 256 threads;
// frames of fewer than 2^31 pixels (32-bit offsets).
template <int RW, int RH>
__device__ __forceinline__ void fill_plane(float* dst, const float* __restrict__ src, int xs, int ys, int w,
                                           int h) {
    static_assert(RW % 2 == 0 && RW <= 64, "fill_plane: RW");
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = 2 * (lane & 31);
    if (c >= RW) return;
    const bool interior = xs >= 0 && ys >= 0 && xs + RW <= w && ys + RH <= h;  // workgroup-uniform
    for (int r = 2 * wv + (lane >> 5); r < RH; r += 8) {
        const int x = xs + c, y = ys + r;
        float2 v;
        if (interior) {
            const float* p = src + (uint32_t)(y * w + x);
            v = make_float2(p[0], p[1]);
        } else {
            const bool yok = (unsigned)y < (unsigned)h;
            v.x = yok && (unsigned)x < (unsigned)w ? src[(uint32_t)(y * w + x)] : 0.0f;
            v.y = yok && (unsigned)(x + 1) < (unsigned)w ? src[(uint32_t)(y * w + x + 1)] : 0.0f;
        }
        *(float2*)(dst + r * RW + c) = v;
    }
}

// ------------------------------------------------------------------ pass A

constexpr int AC = 5;  // CFA apron of pass A

template <bool RCD>
__global__ __launch_bounds__(256) void k_prism_a(FrameParams f, void* ws) {
    __shared__ __align__(16) float s_cfa[Tile<AC>::H][Tile<AC>::W];  // tile + 5
    __shared__ __align__(16) float s_vh[Tile<1>::H][Tile<1>::W];     // tile + 1
    __shared__ __align__(16) float s_lp[Tile<2>::H][Tile<2>::W];     // tile + 2
    const int w = f.w, h = f.h;
    int x0, y0;
    tile_origin(w, h, x0, y0);
    const PrismPlanes P = planes_of(ws, w, h);
    const float scale = (float)f.maxv, rscale = 1.0f / scale;

    // S0: normalised CFA (prism.cpp:58-78)
    s0_window<Tile<AC>::W, Tile<AC>::H>(f, &s_cfa[0][0], x0 - AC, y0 - AC, scale, rscale);
    __syncthreads();
    // S1 over tile + 1, S2 over tile + 2
    for (int k = threadIdx.x; k < Tile<1>::H * Tile<1>::W; k += blockDim.x) {
        const int j = k / Tile<1>::W, i = k - j * Tile<1>::W;
        const int x = x0 - 1 + i, y = y0 - 1 + j;
        const int cj = j + AC - 1, ci = i + AC - 1;
        s_vh[j][i] = inside(x, y, w, h, 4) ? stage_vh([&](int dx, int dy) { return s_cfa[cj + dy][ci + dx]; }) : 0.5f;
    }
    // (S3 reads the low-pass only at R/B sites)
    for (int k = threadIdx.x; k < Tile<2>::H * Tile<2>::W / 2; k += blockDim.x) {
        int i, j;
        nth_site(k, Tile<2>::W, x0 - 2, y0 - 2, f.pattern, true, i, j);
        const int x = x0 - 2 + i, y = y0 - 2 + j;
        const int cj = j + AC - 2, ci = i + AC - 2;
        s_lp[j][i] = inside(x, y, w, h, 2) ? stage_lpf([&](int dx, int dy) { return s_cfa[cj + dy][ci + dx]; })
                                           : 0.0f;
    }
    __syncthreads();
    // tile: CFA and VH everywhere; at G sites green = CFA and PQ = 0.5
    for (int k = threadIdx.x; k < TH * TW; k += blockDim.x) {
        const int j = k / TW, i = k - j * TW;
        const int x = x0 + i, y = y0 + j;
        if (x >= w || y >= h) continue;
        const size_t o = (size_t)y * w + x;
        const float c = s_cfa[j + AC][i + AC];
        P.cfa[o] = c;
        P.vh[o] = s_vh[j + 1][i + 1];
        if (site_of(f.pattern, y, x) == 1) {
            P.pq[o] = 0.5f;
            P.c1[o] = c;
        }
    }
    // tile R/B sites: S4 PQ and S3 green
    for (int k = threadIdx.x; k < TH * TW / 2; k += blockDim.x) {
        int i, j;
        nth_site(k, TW, x0, y0, f.pattern, true, i, j);
        const int x = x0 + i, y = y0 + j;
        if (x >= w || y >= h) continue;
        const size_t o = (size_t)y * w + x;
        const int cj = j + AC, ci = i + AC;
        auto C = [&](int dx, int dy) { return s_cfa[cj + dy][ci + dx]; };
        const bool in4 = inside(x, y, w, h, 4);
        P.pq[o] = in4 ? stage_pq(C) : 0.5f;
        P.c1[o] = in4 ? stage_s3<RCD>(C, [&](int dx, int dy) { return s_vh[j + 1 + dy][i + 1 + dx]; },
                                 [&](int dx, int dy) { return s_lp[j + 2 + dy][i + 2 + dx]; })
                      : 0.0f;
    }
}

// ------------------------------------------------------------------ pass B


// pass B reads CFA with a 7-pixel apron, green 6, PQ 5, VH 2; chroma planes
// cover the tile + 4
template <bool RCD>
__global__ __launch_bounds__(256) void k_prism_b(FrameParams f, void* ws) {
    __shared__ __align__(16) float s_cfa[Tile<7>::H][Tile<7>::W];
    __shared__ __align__(16) float s_g[Tile<6>::H][Tile<6>::W];
    __shared__ __align__(16) float s_pq[Tile<5>::H][Tile<5>::W];   // later: R after the median (tile)
    __shared__ __align__(16) float s_vh[Tile<2>::H][Tile<2>::W];   // later: B after the median (tile)
    __shared__ __align__(16) float s_r[Tile<4>::H][Tile<4>::W];    // chroma planes over tile + 4
    __shared__ __align__(16) float s_b[Tile<4>::H][Tile<4>::W];
    const int w = f.w, h = f.h;
    int x0, y0;
    tile_origin(w, h, x0, y0);
    const PrismPlanes P = planes_of(ws, w, h);

    // (frames on this path have fewer than 2^31 pixels: 32-bit offsets)
    fill_plane<Tile<7>::W, Tile<7>::H>(&s_cfa[0][0], P.cfa, x0 - 7, y0 - 7, w, h);
    fill_plane<Tile<6>::W, Tile<6>::H>(&s_g[0][0], P.c1, x0 - 6, y0 - 6, w, h);
    fill_plane<Tile<5>::W, Tile<5>::H>(&s_pq[0][0], P.pq, x0 - 5, y0 - 5, w, h);
    fill_plane<Tile<2>::W, Tile<2>::H>(&s_vh[0][0], P.vh, x0 - 2, y0 - 2, w, h);
    __syncthreads();

    // chroma planes over tile + 4: at R/B sites the native and S5's opposite
    // chroma, at G sites 0 until S6 (the two loops write disjoint entries)
    for (int k = threadIdx.x; k < Tile<4>::H * Tile<4>::W / 2; k += blockDim.x) {
        int i, j;
        nth_site(k, Tile<4>::W, x0 - 4, y0 - 4, f.pattern, true, i, j);
        const int x = x0 - 4 + i, y = y0 - 4 + j;
        const int cj = j + 3, ci = i + 3;  // s_cfa coordinates (apron 7)
        const int gj = j + 2, gi = i + 2;  // s_g (apron 6)
        const int qj = j + 1, qi = i + 1;  // s_pq (apron 5)
        const float nat = s_cfa[cj][ci];
        float opp = 0.0f;
        if (inside(x, y, w, h, 4))
            opp = stage_s5([&](int dx, int dy) { return s_cfa[cj + dy][ci + dx]; },
                           [&](int dx, int dy) { return s_g[gj + dy][gi + dx]; },
                           [&](int dx, int dy) { return s_pq[qj + dy][qi + dx]; });
        const bool red = site_of(f.pattern, y, x) == 0;
        s_r[j][i] = red ? nat : opp;
        s_b[j][i] = red ? opp : nat;
    }
    for (int k = threadIdx.x; k < Tile<4>::H * Tile<4>::W / 2; k += blockDim.x) {
        int i, j;
        nth_site(k, Tile<4>::W, x0 - 4, y0 - 4, f.pattern, false, i, j);
        s_r[j][i] = 0.0f;
        s_b[j][i] = 0.0f;
    }
    __syncthreads();
    // S6 at G sites over tile + 1 (reads R/B sites only, writes G sites)
    for (int k = threadIdx.x; k < Tile<1>::H * Tile<1>::W / 2; k += blockDim.x) {
        int i, j;
        nth_site(k, Tile<1>::W, x0 - 1, y0 - 1, f.pattern, false, i, j);
        const int x = x0 - 1 + i, y = y0 - 1 + j;
        if (!inside(x, y, w, h, 4)) continue;
        const int rj = j + 3, ri = i + 3;  // chroma planes (apron 4)
        const int gj = j + 5, gi = i + 5;  // s_g (apron 6)
        const int vj = j + 1, vi = i + 1;  // s_vh (apron 2)
        float vc = s_vh[vj][vi];
        float vn = 0.25f * (s_vh[vj - 1][vi - 1] + s_vh[vj - 1][vi + 1] + s_vh[vj + 1][vi - 1] + s_vh[vj + 1][vi + 1]);
        float disc = (fabsf(0.5f - vc) < fabsf(0.5f - vn)) ? vn : vc;
        auto G = [&](int dx, int dy) { return s_g[gj + dy][gi + dx]; };
        const float rv = stage_s6([&](int dx, int dy) { return s_r[rj + dy][ri + dx]; }, G, disc);
        const float bv = stage_s6([&](int dx, int dy) { return s_b[rj + dy][ri + dx]; }, G, disc);
        s_r[rj][ri] = rv;
        s_b[rj][ri] = bv;
    }
    __syncthreads();
    // S7-S9 (PRISM): medians of the colour differences over the tile; R into
    // s_pq, B into s_vh (both dead now), tile coordinates
    float* s_r9 = &s_pq[0][0];
    float* s_b9 = &s_vh[0][0];
    // Each thread takes two vertically adjacent pixels of one column (tile
    // rows j0, j0 + 1): the four row triples of colour differences their
    // windows span are formed and sorted once, the middle two shared.
    if (!RCD) {
        static_assert(TW == 32 && TH == 16, "S7-S9: 256 threads x 2 rows cover the tile");
        const int i = threadIdx.x & (TW - 1), j0 = (threadIdx.x / TW) * 2;
        const int x = x0 + i;
        float tr[4][3], tb[4][3];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int jj = j0 - 1 + q;
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const float g = s_g[jj + 6][i + 6 + dx];
                tr[q][dx + 1] = s_r[jj + 4][i + 4 + dx] - g;
                tb[q][dx + 1] = s_b[jj + 4][i + 4 + dx] - g;
            }
            sort3(tr[q][0], tr[q][1], tr[q][2]);
            sort3(tb[q][0], tb[q][1], tb[q][2]);
        }
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int j = j0 + p, y = y0 + j;
            float r9 = s_r[j + 4][i + 4], b9 = s_b[j + 4][i + 4];
            if (x >= 1 && x < w - 1 && y >= 1 && y < h - 1) {
                float a[9], bb[9];
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        a[3 * q + c] = tr[p + q][c];
                        bb[3 * q + c] = tb[p + q][c];
                    }
                const float lm = median9_of_sorted_rows(a), mm = median9_of_sorted_rows(bb);
                const float g0 = s_g[j + 6][i + 6];
                r9 = clamp01(g0 + lm);
                b9 = clamp01(g0 + mm);
            }
            s_r9[j * TW + i] = r9;
            s_b9[j * TW + i] = b9;
        }
    }
    __syncthreads();
    // S10 + S11: width-4 border fill by source remapping (the source is in
    // this tile), scale, round, clamp, interleave
    const float scale = (float)f.maxv;
    for (int k = threadIdx.x; k < TH * TW; k += blockDim.x) {
        const int j = k / TW, i = k - j * TW;
        const int x = x0 + i, y = y0 + j;
        if (x >= w || y >= h) continue;
        int sx, sy;
        if (y < 4) { sy = 4; sx = x; }
        else if (y >= h - 4) { sy = h - 5; sx = x; }
        else { sy = y; sx = x < 4 ? 4 : (x >= w - 4 ? w - 5 : x); }
        const int q = (sy - y0) * TW + (sx - x0);
        // RCD: S6's chroma planes (apron 4); PRISM: the medians' output
        const float rv = RCD ? s_r[sy - y0 + 4][sx - x0 + 4] : s_r9[q];
        const float bv = RCD ? s_b[sy - y0 + 4][sx - x0 + 4] : s_b9[q];
        int r = (int)fmaf(rv, scale, 0.5f);
        int g = (int)fmaf(s_g[sy - y0 + 6][sx - x0 + 6], scale, 0.5f);
        int b = (int)fmaf(bv, scale, 0.5f);
        r = iclamp(r, 0, f.maxv); g = iclamp(g, 0, f.maxv); b = iclamp(b, 0, f.maxv);
        put_px(f, x, y, r, g, b);
    }
}

}  // namespace

size_t prism_workspace_bytes(int w, int h) { return 4u * plane_stride(w, h) * sizeof(float); }

hipError_t launch_prism(const FrameParams& f, void* ws, hipStream_t st, bool rcd) {
    const dim3 grd(((f.w + TW - 1) / TW) * ((f.h + TH - 1) / TH));
    if (rcd) {
        hipLaunchKernelGGL(k_prism_a<true>, grd, dim3(256), 0, st, f, ws);
        hipLaunchKernelGGL(k_prism_b<true>, grd, dim3(256), 0, st, f, ws);
    } else {
        hipLaunchKernelGGL(k_prism_a<false>, grd, dim3(256), 0, st, f, ws);
        hipLaunchKernelGGL(k_prism_b<false>, grd, dim3(256), 0, st, f, ws);
    }
    return hipGetLastError();
}

}  // namespace imop_hip

// prism_kernels.hip -- PRISM edge-directed demosaic (src/algorithms/prism.cpp:
// 25-316) on gfx950: decode + black level, then the reference's stages as
// six stencil passes over FP32 planes, then the border fill and rounding
// fused into the interleaved-RGB write.
//
// Float reproducibility: every stage keeps the reference's operation order.
// The reference library is compiled by GCC with -mavx2 -mfma, which fuses the
// first product of each a*b + c*d sum into one multiply-add; those sites use
// fmaf() here and every other operation is a separately rounded IEEE op
// (contraction is off for this file; division is correctly rounded).
#include "device_common.hpp"
#include "plain.hpp"

#pragma clang fp contract(off)

namespace imop_hip {

namespace {

constexpr float kEps = 1e-5f;
constexpr float kEpsSq = 1e-10f;

struct PrismPlanes {
    float *cfa, *c0, *c1, *c2, *vh, *pq, *lpf, *r9, *b9;
};

__host__ __device__ inline size_t plane_stride(int w, int h) {
    size_t n = (size_t)w * (size_t)h;
    return (n + 63u) & ~(size_t)63u;  // 256-byte aligned planes
}

__device__ __forceinline__ PrismPlanes planes_of(void* ws, int w, int h) {
    const size_t s = plane_stride(w, h);
    float* b = (float*)ws;
    return PrismPlanes{b, b + s, b + 2 * s, b + 3 * s, b + 4 * s, b + 5 * s, b + 6 * s, b + 7 * s, b + 8 * s};
}

// std::min / std::max argument order: first argument wins ties.
__device__ __forceinline__ float fmin_first(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax_first(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float clamp01(float v) { return fmax_first(0.0f, fmin_first(1.0f, v)); }

__device__ __forceinline__ bool xy_of(int& x, int& y, int w, int h) {
    x = blockIdx.x * blockDim.x + threadIdx.x;
    y = blockIdx.y * blockDim.y + threadIdx.y;
    return x < w && y < h;
}

// S0 + S1 + S2 + S4: normalised CFA, native planes, VH / PQ direction maps and
// the R/B low-pass.  S1/S2/S4 read only the CFA, which S0 computes here from
// the raw samples, so one pass covers them.
__global__ __launch_bounds__(256) void k_prism_s0(FrameParams f, void* ws) {
    int x, y;
    if (!xy_of(x, y, f.w, f.h)) return;
    const int w = f.w, h = f.h;
    PrismPlanes P = planes_of(ws, w, h);
    const size_t i = (size_t)y * w + x;
    const float scale = (float)f.maxv;
    const int s = site_of(f.pattern, y, x);
    const float v = (float)sample_at(f, x, y) / scale;
    P.cfa[i] = v;
    P.c0[i] = s == 0 ? v : 0.0f;
    P.c1[i] = s == 1 ? v : 0.0f;
    P.c2[i] = s == 2 ? v : 0.0f;
}

__global__ __launch_bounds__(256) void k_prism_dirs(FrameParams f, void* ws) {
    int x, y;
    if (!xy_of(x, y, f.w, f.h)) return;
    const int w = f.w, h = f.h;
    PrismPlanes P = planes_of(ws, w, h);
    const size_t i = (size_t)y * w + x;
    const float* cfa = P.cfa;
    const int s = site_of(f.pattern, y, x);
    // S1 (prism.cpp:84-99)
    float vh = 0.5f;
    if (y >= 4 && y < h - 4 && x >= 4 && x < w - 4) {
        float vs = 0.0f, hs = 0.0f;
        for (int k = -4; k <= 4; ++k) {
            vs += fabsf(cfa[i] - cfa[i + (ptrdiff_t)k * w]);
            hs += fabsf(cfa[i] - cfa[i + k]);
        }
        vs = fmax_first(vs, kEpsSq);
        hs = fmax_first(hs, kEpsSq);
        vh = vs / (vs + hs);
    }
    P.vh[i] = vh;
    // S2 (prism.cpp:101-112)
    float lp = 0.0f;
    if (s != 1 && y >= 2 && y < h - 2 && x >= 2 && x < w - 2) {
        float s4 = cfa[i - w] + cfa[i + w] + cfa[i - 1] + cfa[i + 1];
        float d4 = cfa[i - w - 1] + cfa[i - w + 1] + cfa[i + w - 1] + cfa[i + w + 1];
        lp = fmaf(0.0625f, d4, fmaf(0.25f, cfa[i], 0.125f * s4));
    }
    P.lpf[i] = lp;
    // S4 (prism.cpp:165-179)
    float pq = 0.5f;
    if (s != 1 && y >= 4 && y < h - 4 && x >= 4 && x < w - 4) {
        float ps = 0.0f, qs = 0.0f;
        for (int k = -4; k <= 4; ++k) {
            ps += fabsf(cfa[i] - cfa[i + (ptrdiff_t)k * w + k]);
            qs += fabsf(cfa[i] - cfa[i + (ptrdiff_t)k * w - k]);
        }
        ps = fmax_first(ps, kEpsSq);
        qs = fmax_first(qs, kEpsSq);
        pq = ps / (ps + qs);
    }
    P.pq[i] = pq;
}

// S3: green at R/B sites (prism.cpp:114-163)
__global__ __launch_bounds__(256) void k_prism_s3(FrameParams f, void* ws) {
    int x, y;
    if (!xy_of(x, y, f.w, f.h)) return;
    const int w = f.w, h = f.h;
    if (!(y >= 4 && y < h - 4 && x >= 4 && x < w - 4) || site_of(f.pattern, y, x) == 1) return;
    PrismPlanes P = planes_of(ws, w, h);
    const size_t i = (size_t)y * w + x;
    const ptrdiff_t W = w;
    const float* cfa = P.cfa;
    const float* VH = P.vh;
    const float* lpf = P.lpf;
    float vc = VH[i];
    float vn = 0.25f * (VH[i - W - 1] + VH[i - W + 1] + VH[i + W - 1] + VH[i + W + 1]);
    float disc = (fabsf(0.5f - vc) < fabsf(0.5f - vn)) ? vn : vc;
    float c0 = cfa[i];
    float gN = kEps + fabsf(cfa[i - W] - cfa[i + W]) + fabsf(c0 - cfa[i - 2 * W]) + fabsf(cfa[i - W] - cfa[i - 3 * W]);
    float gS = kEps + fabsf(cfa[i + W] - cfa[i - W]) + fabsf(c0 - cfa[i + 2 * W]) + fabsf(cfa[i + W] - cfa[i + 3 * W]);
    float gW = kEps + fabsf(cfa[i - 1] - cfa[i + 1]) + fabsf(c0 - cfa[i - 2]) + fabsf(cfa[i - 1] - cfa[i - 3]);
    float gE = kEps + fabsf(cfa[i + 1] - cfa[i - 1]) + fabsf(c0 - cfa[i + 2]) + fabsf(cfa[i + 1] - cfa[i + 3]);
    float l0 = lpf[i], lN = lpf[i - 2 * W], lS = lpf[i + 2 * W], lW = lpf[i - 2], lE = lpf[i + 2];
    float eN = cfa[i - W] * (1.0f + (l0 - lN) / (kEps + l0 + lN));
    float eS = cfa[i + W] * (1.0f + (l0 - lS) / (kEps + l0 + lS));
    float eW = cfa[i - 1] * (1.0f + (l0 - lW) / (kEps + l0 + lW));
    float eE = cfa[i + 1] * (1.0f + (l0 - lE) / (kEps + l0 + lE));
    float vEst = fmaf(gS, eN, gN * eS) / (gN + gS);
    float hEst = fmaf(gW, eE, gE * eW) / (gE + gW);
    float lapV = (2.0f * c0 - cfa[i - 2 * W] - cfa[i + 2 * W]) * 0.25f;
    float lapH = (2.0f * c0 - cfa[i - 2] - cfa[i + 2]) * 0.25f;
    vEst += lapV;
    hEst += lapH;
    P.c1[i] = clamp01(fmaf(disc, hEst, (1.0f - disc) * vEst));
}

// S5: opposite chroma at R/B sites (prism.cpp:181-225)
__global__ __launch_bounds__(256) void k_prism_s5(FrameParams f, void* ws) {
    int x, y;
    if (!xy_of(x, y, f.w, f.h)) return;
    const int w = f.w, h = f.h;
    const int s = site_of(f.pattern, y, x);
    if (!(y >= 4 && y < h - 4 && x >= 4 && x < w - 4) || s == 1) return;
    PrismPlanes P = planes_of(ws, w, h);
    const size_t i = (size_t)y * w + x;
    const ptrdiff_t W = w;
    float* X = (s == 0) ? P.c2 : P.c0;
    const float* G = P.c1;
    const float* PQ = P.pq;
    float pc = PQ[i];
    float pn = 0.25f * (PQ[i - W - 1] + PQ[i - W + 1] + PQ[i + W - 1] + PQ[i + W + 1]);
    float disc = (fabsf(0.5f - pc) < fabsf(0.5f - pn)) ? pn : pc;
    float g0 = G[i];
    float gNW = kEps + fabsf(X[i - W - 1] - X[i + W + 1]) + fabsf(g0 - G[i - 2 * W - 2]) + fabsf(X[i - W - 1] - X[i - 3 * W - 3]);
    float gNE = kEps + fabsf(X[i - W + 1] - X[i + W - 1]) + fabsf(g0 - G[i - 2 * W + 2]) + fabsf(X[i - W + 1] - X[i - 3 * W + 3]);
    float gSW = kEps + fabsf(X[i + W - 1] - X[i - W + 1]) + fabsf(g0 - G[i + 2 * W - 2]) + fabsf(X[i + W - 1] - X[i + 3 * W - 3]);
    float gSE = kEps + fabsf(X[i + W + 1] - X[i - W - 1]) + fabsf(g0 - G[i + 2 * W + 2]) + fabsf(X[i + W + 1] - X[i + 3 * W + 3]);
    float eNW = X[i - W - 1] - G[i - W - 1];
    float eNE = X[i - W + 1] - G[i - W + 1];
    float eSW = X[i + W - 1] - G[i + W - 1];
    float eSE = X[i + W + 1] - G[i + W + 1];
    float pEst = fmaf(gNW, eSE, gSE * eNW) / (gNW + gSE);
    float qEst = fmaf(gNE, eSW, gSW * eNE) / (gNE + gSW);
    X[i] = clamp01(fmaf(disc, qEst, fmaf(1.0f - disc, pEst, g0)));
}

// S6: R and B at G sites (prism.cpp:227-264)
__global__ __launch_bounds__(256) void k_prism_s6(FrameParams f, void* ws) {
    int x, y;
    if (!xy_of(x, y, f.w, f.h)) return;
    const int w = f.w, h = f.h;
    if (!(y >= 4 && y < h - 4 && x >= 4 && x < w - 4) || site_of(f.pattern, y, x) != 1) return;
    PrismPlanes P = planes_of(ws, w, h);
    const size_t i = (size_t)y * w + x;
    const ptrdiff_t W = w;
    const float* G = P.c1;
    const float* VH = P.vh;
    float vc = VH[i];
    float vn = 0.25f * (VH[i - W - 1] + VH[i - W + 1] + VH[i + W - 1] + VH[i + W + 1]);
    float disc = (fabsf(0.5f - vc) < fabsf(0.5f - vn)) ? vn : vc;
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
        float* X = ch == 0 ? P.c0 : P.c2;
        float g0 = G[i];
        float gN = kEps + fabsf(g0 - G[i - 2 * W]) + fabsf(X[i - W] - X[i + W]) + fabsf(X[i - W] - X[i - 3 * W]);
        float gS = kEps + fabsf(g0 - G[i + 2 * W]) + fabsf(X[i + W] - X[i - W]) + fabsf(X[i + W] - X[i + 3 * W]);
        float gW = kEps + fabsf(g0 - G[i - 2]) + fabsf(X[i - 1] - X[i + 1]) + fabsf(X[i - 1] - X[i - 3]);
        float gE = kEps + fabsf(g0 - G[i + 2]) + fabsf(X[i + 1] - X[i - 1]) + fabsf(X[i + 1] - X[i + 3]);
        float dN = X[i - W] - G[i - W];
        float dS = X[i + W] - G[i + W];
        float dW = X[i - 1] - G[i - 1];
        float dE = X[i + 1] - G[i + 1];
        float vEst = fmaf(gS, dN, gN * dS) / (gN + gS);
        float hEst = fmaf(gW, dE, gE * dW) / (gE + gW);
        X[i] = clamp01(fmaf(1.0f - disc, vEst, fmaf(disc, hEst, g0)));
    }
}

// prism_detail::median_3x3 (prism.cpp:9-21): the same exchange network.
__device__ __forceinline__ float median9(float v[9]) {
#define IMOP_CX(a, b)                        \
    {                                        \
        float ta = v[a], tb = v[b];          \
        bool sw = ta > tb;                   \
        v[a] = sw ? tb : ta;                 \
        v[b] = sw ? ta : tb;                 \
    }
    IMOP_CX(0, 1) IMOP_CX(3, 4) IMOP_CX(6, 7) IMOP_CX(1, 2) IMOP_CX(4, 5) IMOP_CX(7, 8)
    IMOP_CX(0, 1) IMOP_CX(3, 4) IMOP_CX(6, 7) IMOP_CX(0, 3) IMOP_CX(4, 7) IMOP_CX(1, 4)
    IMOP_CX(3, 6) IMOP_CX(2, 5) IMOP_CX(5, 8) IMOP_CX(1, 3) IMOP_CX(4, 6) IMOP_CX(2, 5)
    IMOP_CX(2, 3) IMOP_CX(5, 6) IMOP_CX(2, 4) IMOP_CX(3, 5) IMOP_CX(3, 4)
#undef IMOP_CX
    return v[4];
}

// S7 + S8 + S9: colour differences, their 3x3 medians, reconstruction
// (prism.cpp:266-300).  The outer ring keeps the S6 chroma.
__global__ __launch_bounds__(256) void k_prism_median(FrameParams f, void* ws) {
    int x, y;
    if (!xy_of(x, y, f.w, f.h)) return;
    const int w = f.w, h = f.h;
    PrismPlanes P = planes_of(ws, w, h);
    const size_t i = (size_t)y * w + x;
    if (!(y >= 1 && y < h - 1 && x >= 1 && x < w - 1)) {
        P.r9[i] = P.c0[i];
        P.b9[i] = P.c2[i];
        return;
    }
    float a[9], b[9];
    int k = 0;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
            size_t j = (size_t)(y + dy) * w + (size_t)(x + dx);
            float g = P.c1[j];
            a[k] = P.c0[j] - g;
            b[k] = P.c2[j] - g;
            ++k;
        }
    float lm = median9(a), mm = median9(b);
    float g0 = P.c1[i];
    P.r9[i] = clamp01(g0 + lm);
    P.b9[i] = clamp01(g0 + mm);
}

// S10 + S11: border fill (width 4) by source remapping, scale, round, clamp,
// interleave (prism.cpp:302-313, common.hpp:98-129).
__global__ __launch_bounds__(256) void k_prism_out(FrameParams f, void* ws) {
    int x, y;
    if (!xy_of(x, y, f.w, f.h)) return;
    const int w = f.w, h = f.h;
    PrismPlanes P = planes_of(ws, w, h);
    int sx, sy;
    if (y < 4) { sy = 4; sx = x; }
    else if (y >= h - 4) { sy = h - 5; sx = x; }
    else { sy = y; sx = x < 4 ? 4 : (x >= w - 4 ? w - 5 : x); }
    const size_t j = (size_t)sy * w + sx;
    const float scale = (float)f.maxv;
    int r = (int)fmaf(P.r9[j], scale, 0.5f);
    int g = (int)fmaf(P.c1[j], scale, 0.5f);
    int b = (int)fmaf(P.b9[j], scale, 0.5f);
    r = iclamp(r, 0, f.maxv); g = iclamp(g, 0, f.maxv); b = iclamp(b, 0, f.maxv);
    put_px(f, x, y, r, g, b);
}

}  // namespace

size_t prism_workspace_bytes(int w, int h) { return 9u * plane_stride(w, h) * sizeof(float); }

hipError_t launch_prism(const FrameParams& f, void* ws, hipStream_t st) {
    dim3 blk(64, 4);
    dim3 grd((f.w + 63) / 64, (f.h + 3) / 4);
    hipLaunchKernelGGL(k_prism_s0, grd, blk, 0, st, f, ws);
    hipLaunchKernelGGL(k_prism_dirs, grd, blk, 0, st, f, ws);
    hipLaunchKernelGGL(k_prism_s3, grd, blk, 0, st, f, ws);
    hipLaunchKernelGGL(k_prism_s5, grd, blk, 0, st, f, ws);
    hipLaunchKernelGGL(k_prism_s6, grd, blk, 0, st, f, ws);
    hipLaunchKernelGGL(k_prism_median, grd, blk, 0, st, f, ws);
    hipLaunchKernelGGL(k_prism_out, grd, blk, 0, st, f, ws);
    return hipGetLastError();
}

}  // namespace imop_hip

// chain_kernels.hip -- the differentiable ISP chain in FP32 on gfx950:
// bilinear demosaic -> white balance -> 3x3 CCM -> gamma, its forward (with
// an optional fused MSE) and its hand-written vector-Jacobian product.
//
// Forward, per frame (x = raw / max_value):
//   a = D x                   bilinear demosaic (hqli.cpp:7-44 with exact
//                             averages) + fill_rgb_borders (common.hpp:47-96),
//                             corners 0
//   u = clamp01(g .* a)       white balance (gpu_ops_kernels.cu:119-125)
//   z = clamp01(M u)          CCM (gpu_ops_kernels.cu:92-96, gpu_common.cuh:43-48)
//   y = z^(1/gamma)           gamma (gpu_ops_kernels.cu:57-64)
//
// Backward recomputes the forward from raw and the parameters; nothing of the
// forward is stored.  Every output pixel on the 2-pixel border is a copy of an
// interior pixel (or a constant 0 corner), so its cotangent is folded onto
// that source pixel first; one pass back through gamma, CCM and white balance
// per interior pixel then gives dL/da there, and D^T is a 3x3 gather with the
// bilinear weights (a raw R / B site collects 1 from itself, 1/2 from its 4
// axial neighbours and 1/4 from its 4 diagonal ones in its own colour; a G
// site 1 from itself and 1/4 from its 4 axial neighbours in green).
//
// Derivative conventions: clamp01 passes the gradient where 0 <= v <= 1;
// d/dz z^(1/gamma) is 0 at z = 0.  Parameter gradients are summed per
// workgroup and then over workgroups in a fixed order (deterministic, no
// atomics).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "chain.hpp"
#include "device_common.hpp"

namespace imop_hip {

namespace {

constexpr int kTW = 64;   // backward tile: 64 x 16 raw pixels per iteration
constexpr int kTH = 16;
constexpr int kFW = 256;  // forward: row segments of 256 pixels per workgroup
constexpr int kFR = 4;    // forward: rows per step (at most the 6-row minimum frame height)

__device__ __forceinline__ bool interior(int x, int y, int w, int h) {
    return x >= 2 && x < w - 2 && y >= 2 && y < h - 2;
}

// Source of output pixel (x, y) under fill_rgb_borders(border 2); false for
// the 16 corner pixels.
__device__ __forceinline__ bool src_of(int x, int y, int w, int h, int& sx, int& sy) {
    const bool xin = x >= 2 && x < w - 2;
    if (y < 2 || y >= h - 2) {
        sx = x;
        sy = y < 2 ? 2 : h - 3;
        return xin;
    }
    sy = y;
    sx = x < 2 ? 2 : (x >= w - 2 ? w - 3 : x);
    return true;
}

// Bilinear demosaic at an interior pixel; X(dx, dy) reads the raw
// neighbourhood (already scaled by 1 / max_value).
template <typename Ld>
__device__ __forceinline__ void bilinear_at(int pattern, int x, int y, Ld X, float a[3]) {
    const int s = site_of(pattern, y, x);
    const float c = X(0, 0);
    if (s != 1) {
        const float ax = (X(0, -1) + X(0, 1) + X(1, 0) + X(-1, 0)) * 0.25f;
        const float dg = (X(-1, -1) + X(1, -1) + X(-1, 1) + X(1, 1)) * 0.25f;
        a[1] = ax;
        if (s == 0) { a[0] = c; a[2] = dg; } else { a[2] = c; a[0] = dg; }
    } else {
        const float hz = (X(-1, 0) + X(1, 0)) * 0.5f;
        const float vt = (X(0, -1) + X(0, 1)) * 0.5f;
        a[1] = c;
        if ((y & 1) == r_row_of(pattern)) { a[0] = hz; a[2] = vt; } else { a[0] = vt; a[2] = hz; }
    }
}

__device__ __forceinline__ float clamp01f(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

struct Px {
    float a[3], b[3], u[3], c[3], z[3], y[3];
};

// z^e for z in [0, 1] with the hardware log2 / exp2 (v_log_f32, v_exp_f32,
// about 1 ulp each); 0 at z = 0 (and for denormal z, which v_log flushes).
__device__ __forceinline__ float gamma_pow(float z, float e) {
    return z > 0.0f ? __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(z)) : 0.0f;
}

__device__ __forceinline__ void chain_px(const float a[3], const float g[3], const float M[9], float e, Px& s) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s.a[k] = a[k];
        s.b[k] = g[k] * a[k];
        s.u[k] = clamp01f(s.b[k]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        s.c[i] = M[3 * i] * s.u[0] + M[3 * i + 1] * s.u[1] + M[3 * i + 2] * s.u[2];
        s.z[i] = clamp01f(s.c[i]);
        s.y[i] = gamma_pow(s.z[i], e);
    }
}

// Back through gamma, CCM and white balance: dL/da into ga, this pixel's
// parameter gradients into pg (gains 0..2, CCM 3..11, row-major).
__device__ __forceinline__ void chain_px_bwd(const Px& s, const float gy[3], const float g[3], const float M[9],
                                             float e, float ga[3], float pg[12]) {
    float gc[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        // d/dz z^e = e z^(e-1) = e y / z  (0 where y is 0: z = 0 or flushed)
        const float dz = s.y[i] > 0.0f ? e * s.y[i] * __builtin_amdgcn_rcpf(s.z[i]) : 0.0f;
        const bool pass = s.c[i] >= 0.0f && s.c[i] <= 1.0f;
        gc[i] = pass ? gy[i] * dz : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) pg[3 + 3 * i + j] = gc[i] * s.u[j];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float gu = M[k] * gc[0] + M[3 + k] * gc[1] + M[6 + k] * gc[2];
        const bool pass = s.b[k] >= 0.0f && s.b[k] <= 1.0f;
        const float gb = pass ? gu : 0.0f;
        pg[k] = gb * s.a[k];
        ga[k] = gb * g[k];
    }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum `m` per-thread values over the workgroup into out[0..m) (thread 0..m-1).
template <int M>
__device__ __forceinline__ void block_sum(float (&v)[M], float* out) {
    __shared__ float red[16][12];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const float s = wave_sum(v[k]);
        if (lane == 0) red[wv][k] = s;
    }
    __syncthreads();
    if ((int)threadIdx.x < M) {
        float s = 0.0f;
        for (int q = 0; q < nw; ++q) s += red[q][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

struct Params {
    float g[3], M[9];
};

__device__ __forceinline__ Params load_params(const float* gains, const float* ccm) {
    Params p;
#pragma unroll
    for (int k = 0; k < 3; ++k) p.g[k] = gains[k];
#pragma unroll
    for (int k = 0; k < 9; ++k) p.M[k] = ccm[k];
    return p;
}

}  // namespace

// ------------------------------------------------------------------ forward

__global__ __launch_bounds__(256) void k_chain_fwd(ChainArgs A) {
    // workgroup (bx, by): columns bx * 256 + tid of rows [by * per, (by + 1) * per)
    // of the n * h rows of the batch, a contiguous run, so the 3-row
    // neighbourhoods are re-read from the caches rather than from HBM
    const Params P = load_params(A.gains, A.ccm);
    const int w = A.w, h = A.h;
    const int x = blockIdx.x * kFW + threadIdx.x;
    const float im = A.inv_max;
    float acc[1] = {0.0f};
    const int rows = A.n * h;
    const int per = (rows + gridDim.y - 1) / gridDim.y;
    const int r0 = blockIdx.y * per, r1 = min(rows, r0 + per);
    if (x < w && r0 < r1) {
        // fill_rgb_borders' source column (every frame here is at least 6 x 6)
        const int sx = min(max(x, 2), w - 3);
        const bool xin = x >= 2 && x < w - 2;
        int f = r0 / h, y = r0 - f * h;  // frame and row of row ry
        // kFR rows a step, every load of the step issued before any of its
        // arithmetic, so a wave waits on memory once per step, not per row
        for (int ry = r0; ry < r1; ry += kFR) {
            float nb[kFR][9], tg[kFR][3];
            int ys[kFR];
            bool on[kFR];
#pragma unroll
            for (int u = 0; u < kFR; ++u) {
                int yu = y + u, fu = f;
                if (yu >= h) {  // kFR <= h: at most one frame boundary a step
                    yu -= h;
                    ++fu;
                }
                // rows past the run load the run's last row again and are dropped
                if (ry + u >= r1) {
                    yu = (r1 - 1) - ((r1 - 1) / h) * h;
                    fu = (r1 - 1) / h;
                }
                ys[u] = yu;
                on[u] = xin || (yu >= 2 && yu < h - 2);  // not one of the 16 corner pixels
                const int sy = min(max(yu, 2), h - 3);
                const float* c = A.raw + ((int64_t)fu * h + sy) * w + sx;
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                    for (int dx = -1; dx <= 1; ++dx) nb[u][(dy + 1) * 3 + dx + 1] = c[dy * w + dx];
                if (A.target) {
                    const int64_t i = ((int64_t)fu * h + yu) * w + x;
#pragma unroll
                    for (int k = 0; k < 3; ++k) tg[u][k] = A.target[i * 3 + k];
                }
            }
#pragma unroll
            for (int u = 0; u < kFR; ++u) {
                if (ry + u >= r1) break;
                const int yu = ys[u];
                const int sy = min(max(yu, 2), h - 3);
                float a[3] = {0.0f, 0.0f, 0.0f};
                if (on[u]) bilinear_at(A.pattern, sx, sy, [&](int dx, int dy) { return nb[u][(dy + 1) * 3 + dx + 1] * im; }, a);
                Px s;
                chain_px(a, P.g, P.M, A.inv_gamma, s);
                const int64_t i = (int64_t)(ry + u) * w + x;  // pixel index in the batch
                if (A.out) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) A.out[i * 3 + k] = s.y[k];
                }
                if (A.target) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float d = s.y[k] - tg[u][k];
                        acc[0] += d * d;
                    }
                }
            }
            y += kFR;
            if (y >= h) {
                y -= h;
                ++f;
            }
        }
    }
    if (A.target) block_sum<1>(acc, A.partials + (int64_t)blockIdx.y * gridDim.x + blockIdx.x);
}

// ----------------------------------------------------------------- backward

// (8 waves per SIMD: the recompute + VJP fits in 64 VGPRs without spilling,
// and the extra waves hide the tile loads; measured 80 -> 66 us at 8 x 1006x758)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_chain_bwd(ChainArgs A) {
    __shared__ float s_x[kTH + 4][kTW + 4];
    __shared__ float s_ga[3][kTH + 2][kTW + 2];
    const Params P = load_params(A.gains, A.ccm);
    const int w = A.w, h = A.h;
    const int64_t wh = (int64_t)w * h;
    const int ntx = (w + kTW - 1) / kTW, nty = (h + kTH - 1) / kTH;
    const int ntiles = A.n * ntx * nty;
    float acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = 0.0f;

    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int f = t / (ntx * nty);
        const int r = t - f * ntx * nty;
        const int ty = r / ntx, tx = r - ty * ntx;
        const int x0 = tx * kTW, y0 = ty * kTH;
        const float* fr = A.raw + (int64_t)f * wh;
        const int64_t fo = (int64_t)f * wh;  // pixel offset of the frame
        // raw tile + 2-pixel apron, scaled
        for (int k = threadIdx.x; k < (kTH + 4) * (kTW + 4); k += blockDim.x) {
            const int j = k / (kTW + 4), i = k - j * (kTW + 4);
            const int x = x0 - 2 + i, y = y0 - 2 + j;
            s_x[j][i] = (x >= 0 && x < w && y >= 0 && y < h) ? fr[(int64_t)y * w + x] * A.inv_max : 0.0f;
        }
        __syncthreads();
        // dL/da over the tile + 1-pixel apron
        for (int k = threadIdx.x; k < (kTH + 2) * (kTW + 2); k += blockDim.x) {
            const int j = k / (kTW + 2), i = k - j * (kTW + 2);
            const int x = x0 - 1 + i, y = y0 - 1 + j;
            float ga[3] = {0.0f, 0.0f, 0.0f};
            if (interior(x, y, w, h)) {
                float a[3];
                bilinear_at(A.pattern, x, y, [&](int dx, int dy) { return s_x[j + 1 + dy][i + 1 + dx]; }, a);
                Px s;
                chain_px(a, P.g, P.M, A.inv_gamma, s);
                // cotangent of this pixel plus those of the border pixels copying it
                float cot[3], cnt = 1.0f;
                const float* src = A.grad_out ? A.grad_out : A.target;
                const int64_t p0 = (fo + (int64_t)y * w + x) * 3;
#pragma unroll
                for (int c = 0; c < 3; ++c) cot[c] = src[p0 + c];
                auto add = [&](int bx, int by) {
                    const int64_t q = (fo + (int64_t)by * w + bx) * 3;
#pragma unroll
                    for (int c = 0; c < 3; ++c) cot[c] += src[q + c];
                    cnt += 1.0f;
                };
                if (y == 2) { add(x, 0); add(x, 1); }
                if (y == h - 3) { add(x, h - 2); add(x, h - 1); }
                if (x == 2) { add(0, y); add(1, y); }
                if (x == w - 3) { add(w - 2, y); add(w - 1, y); }
                float gy[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) gy[c] = A.grad_out ? cot[c] : A.mse_scale * (cnt * s.y[c] - cot[c]);
                float pg[12];
                chain_px_bwd(s, gy, P.g, P.M, A.inv_gamma, ga, pg);
                if (i >= 1 && i <= kTW && j >= 1 && j <= kTH) {  // parameters: tile pixels only
#pragma unroll
                    for (int q = 0; q < 12; ++q) acc[q] += pg[q];
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) s_ga[c][j][i] = ga[c];
        }
        __syncthreads();
        // D^T: gather dL/da of the 3x3 neighbourhood with the bilinear weights
        if (A.grad_raw) {
            for (int k = threadIdx.x; k < kTH * kTW; k += blockDim.x) {
                const int j = k / kTW, i = k - j * kTW;
                const int x = x0 + i, y = y0 + j;
                if (x >= w || y >= h) continue;
                const int s = site_of(A.pattern, y, x);
                const int c = s;  // R and B feed their own colour, G feeds green
                const float(*G)[kTW + 2] = s_ga[c];
                const int jj = j + 1, ii = i + 1;
                const float self = G[jj][ii];
                const float ax = G[jj - 1][ii] + G[jj + 1][ii] + G[jj][ii - 1] + G[jj][ii + 1];
                float v;
                if (s != 1) {
                    const float dg = G[jj - 1][ii - 1] + G[jj - 1][ii + 1] + G[jj + 1][ii - 1] + G[jj + 1][ii + 1];
                    v = self + 0.5f * ax + 0.25f * dg;
                } else {
                    v = self + 0.25f * ax;
                }
                A.grad_raw[fo + (int64_t)y * w + x] = v * A.inv_max;
            }
        }
        __syncthreads();
    }
    block_sum<12>(acc, A.partials + (int64_t)blockIdx.x * 12);
}

// Sum G (<= 2048) partial rows of m values in a fixed order: workgroup v
// sums value v, each thread 8 partials loaded together.
__global__ __launch_bounds__(256) void k_chain_reduce(const float* __restrict__ partials, int G, int m, float scale,
                                                      float* __restrict__ out) {
    const int v = blockIdx.x;
    float x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int k = threadIdx.x + u * 256;
        x[u] = k < G ? partials[(int64_t)k * m + v] : 0.0f;
    }
    __shared__ float red[4];
    const float s = wave_sum(((x[0] + x[1]) + (x[2] + x[3])) + ((x[4] + x[5]) + (x[6] + x[7])));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[v] = ((red[0] + red[1]) + (red[2] + red[3])) * scale;
}

// ---------------------------------------------------- the demosaic operator

__global__ __launch_bounds__(256) void k_bilinear_fwd(const float* x, int w, int h, int pattern, float* out) {
    const int64_t total = (int64_t)w * h;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / w), xx = (int)(i - (int64_t)y * w);
        float a[3] = {0.0f, 0.0f, 0.0f};
        int sx, sy;
        if (src_of(xx, y, w, h, sx, sy)) {
            const float* c = x + (int64_t)sy * w + sx;
            bilinear_at(pattern, sx, sy, [&](int dx, int dy) { return c[dy * w + dx]; }, a);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) out[i * 3 + k] = a[k];
    }
}

// D^T v: fold the border samples onto their sources, then gather.
__global__ __launch_bounds__(256) void k_bilinear_vjp(const float* v, int w, int h, int pattern, float* out) {
    const int64_t total = (int64_t)w * h;
    auto folded = [&](int x, int y, int c) {
        float s = v[((int64_t)y * w + x) * 3 + c];
        if (y == 2) s += v[((int64_t)0 * w + x) * 3 + c] + v[((int64_t)1 * w + x) * 3 + c];
        if (y == h - 3) s += v[((int64_t)(h - 2) * w + x) * 3 + c] + v[((int64_t)(h - 1) * w + x) * 3 + c];
        if (x == 2) s += v[((int64_t)y * w + 0) * 3 + c] + v[((int64_t)y * w + 1) * 3 + c];
        if (x == w - 3) s += v[((int64_t)y * w + w - 2) * 3 + c] + v[((int64_t)y * w + w - 1) * 3 + c];
        return s;
    };
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / w), x = (int)(i - (int64_t)y * w);
        const int s = site_of(pattern, y, x);
        float acc = 0.0f;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int px = x + dx, py = y + dy;
                if (!interior(px, py, w, h)) continue;
                const bool self = dx == 0 && dy == 0, diag = dx != 0 && dy != 0;
                float wgt;
                if (s != 1) wgt = self ? 1.0f : (diag ? 0.25f : 0.5f);
                else wgt = self ? 1.0f : (diag ? 0.0f : 0.25f);
                if (wgt != 0.0f) acc += wgt * folded(px, py, s);
            }
        out[i] = acc;
    }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_chain_fwd(const ChainArgs& a, int blocks, hipStream_t st) {
    // blocks = the partial-sum budget; spread over row segments x row groups
    const int gx = (a.w + kFW - 1) / kFW;
    const int rows = a.n * a.h;
    const int gy = std::max(1, std::min(rows, blocks / gx));
    hipLaunchKernelGGL(k_chain_fwd, dim3(gx, gy), dim3(kFW), 0, st, a);
    return hipGetLastError();
}

int chain_fwd_blocks(int n, int w, int h, int budget) {
    const int gx = (w + kFW - 1) / kFW;
    return gx * std::max(1, std::min(n * h, budget / gx));
}

hipError_t launch_chain_bwd(const ChainArgs& a, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_chain_bwd, dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_chain_reduce(const float* partials, int G, int m, float scale, float* out, hipStream_t st) {
    hipLaunchKernelGGL(k_chain_reduce, dim3(m), dim3(256), 0, st, partials, G, m, scale, out);
    return hipGetLastError();
}

hipError_t launch_bilinear_fwd(const float* x, int w, int h, int pattern, float* out, hipStream_t st) {
    const int64_t total = (int64_t)w * h;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(k_bilinear_fwd, dim3(blocks), dim3(256), 0, st, x, w, h, pattern, out);
    return hipGetLastError();
}

hipError_t launch_bilinear_vjp(const float* v, int w, int h, int pattern, float* out, hipStream_t st) {
    const int64_t total = (int64_t)w * h;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(k_bilinear_vjp, dim3(blocks), dim3(256), 0, st, v, w, h, pattern, out);
    return hipGetLastError();
}

int chain_bwd_tiles(int n, int w, int h) { return n * ((w + kTW - 1) / kTW) * ((h + kTH - 1) / kTH); }

}  // namespace imop_hip

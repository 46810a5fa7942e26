// chain.hpp -- launchers of the differentiable ISP chain (chain_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace imop_hip {

// Arguments of the chain kernels.  Parameters (gains[3], ccm[9] row-major)
// are device pointers; raw is [n][h][w] in units of 1 / inv_max; out,
// target and grad_out are [n][h][w][3]; grad_raw is [n][h][w].
struct ChainArgs {
    const float* raw;
    int32_t n, w, h, pattern;
    float inv_max;
    const float* gains;
    const float* ccm;
    float inv_gamma;
    float* out;             // forward output (may be null)
    const float* target;    // MSE target (forward loss; backward when grad_out is null)
    const float* grad_out;  // backward cotangent (null: MSE against target)
    float* grad_raw;        // backward dL/draw (may be null)
    float* partials;        // per-workgroup sums (1 or 12 floats each)
    float mse_scale;        // 2 / (n * h * w * 3)
};

hipError_t launch_chain_fwd(const ChainArgs& a, int blocks, hipStream_t st);
hipError_t launch_chain_bwd(const ChainArgs& a, int blocks, hipStream_t st);
hipError_t launch_chain_reduce(const float* partials, int G, int m, float scale, float* out, hipStream_t st);
hipError_t launch_bilinear_fwd(const float* x, int w, int h, int pattern, float* out, hipStream_t st);
hipError_t launch_bilinear_vjp(const float* v, int w, int h, int pattern, float* out, hipStream_t st);
int chain_bwd_tiles(int n, int w, int h);
// Workgroups (= partial sums) of the forward for a partial-sum budget.
int chain_fwd_blocks(int n, int w, int h, int budget);

}  // namespace imop_hip

// defect_kernels.hip -- adaptive defect correction (src/defect_correct/adaptive.cpp:8-61)
// on the GPU, with the reference's raster-order result.
//
// The reference walks the interior [2, w-2) x [2, h-2) in raster order, in
// place: pixel p becomes
//     F(p) = |orig(p) - med| > T ? med : orig(p),
// where med is element n/2 of the sorted values of p's same-colour
// neighbours at distance 2 that lie in the interior: the four that come
// before p in raster order (x-2..x+2 on row y-2, and x-2 on row y) as
// already corrected, the four after it as in the input.  The corrected frame
// v is therefore the solution of v(p) = F(orig(p), v(before), orig(after)):
// a triangular system with exactly one solution, which any evaluation order
// reaches once every pixel has been evaluated after all of its predecessors
// reached their final value, and which is reached exactly when a full sweep
// changes nothing.
//
// GPU form:
//   k_defect_copy  saves the input (orig) into a scratch frame O.
//   k_defect_pass  (pass k): each 128 x 32 tile loads its pixels and a 2-pixel
//                  halo into LDS (current values from the frame, or from O in
//                  pass 0; originals from O) and iterates F over its pixels in
//                  place until an iteration changes nothing, the halo held
//                  fixed; then writes its changed pixels to the frame.  A tile
//                  records which neighbouring tiles its changes can reach (the
//                  right one through (x+2, y), the left and the three below
//                  through row y+2); pass k+1 re-runs only the tiles reached,
//                  and stops at once when pass k changed nothing.
// Values read while other tiles are writing are older iterates, which the
// fixed point tolerates; the kernel boundary between passes makes every
// write visible.  The host repeats passes until one changes nothing: the
// frame is then exactly the reference's output.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "epilogue.hpp"

namespace imop_hip {
namespace {

constexpr int kTX = 128, kTY = 32;                      // tile (pixels)
constexpr int kLW = kTX + 8, kLH = kTY + 4;             // LDS rows: x0-2 .. x0+TX+1 (+ pad), y0-2 .. y0+TY+1
constexpr uint16_t kPad = 0xFFFF;                       // outside the interior: sorts after every value
constexpr int kLocalCap = 1024;                         // > the longest chain inside a tile (TX/2 + TY + 2)

// Bits a tile records for the tiles its changes can reach.
enum : uint32_t { kRight = 1, kLeft = 2, kBelowLeft = 4, kBelow = 8, kBelowRight = 16, kAny = 32 };

__device__ __forceinline__ void ce(uint32_t& a, uint32_t& b) {
    const uint32_t lo = min(a, b), hi = max(a, b);
    a = lo;
    b = hi;
}

// Sorts four values in place (five compare-exchanges).
__device__ __forceinline__ void sort4(uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3) {
    ce(a0, a1);
    ce(a2, a3);
    ce(a0, a2);
    ce(a1, a3);
    ce(a1, a2);
}

// Element n/2 of the sorted valid neighbours.  a: the four neighbours before
// the pixel, b: the four after it, invalid ones kPad.  With both quartets
// sorted, the k-th smallest (1-based) of the eight is
// min over i + j = k of max(a_i, b_j) (a_0 = b_0 = below every value); the
// valid count n is 8 inside [4, w-4) x [4, h-4), else 1, 2, 3 or 5 (k = 1, 2,
// 2, 3), and the pads sort after the valid values.
__device__ __forceinline__ uint32_t median_n2(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t b0,
                                              uint32_t b1, uint32_t b2, uint32_t b3, int n) {
    sort4(a0, a1, a2, a3);
    sort4(b0, b1, b2, b3);
    if (n == 8) return min(min(max(a0, b3), max(a1, b2)), min(max(a2, b1), max(a3, b0)));
    if (n == 5) return min(min(b2, max(a0, b1)), min(max(a1, b0), a2));
    if (n >= 2) return min(min(b1, max(a0, b0)), a1);
    return min(a0, b0);
}

template <bool U16>
__device__ __forceinline__ uint32_t ld_px(const uint8_t* p, size_t i) {
    return U16 ? (uint32_t)((const uint16_t*)p)[i] : (uint32_t)p[i];
}

template <bool U16>
__device__ __forceinline__ void st_px(uint8_t* p, size_t i, uint32_t v) {
    if (U16) ((uint16_t*)p)[i] = (uint16_t)v;
    else p[i] = (uint8_t)v;
}

struct DefectArgs {
    uint8_t* data;       // the frame, in place
    const uint8_t* orig; // its input values (scratch copy)
    uint8_t* flags;      // 2 x ntiles: reach bits of the previous / this pass
    int* ctl;            // [0..2] "pass changed something" ring, [3] error
    int w, h, thresh, ntx, nty, pass;
};

__global__ void __launch_bounds__(256) k_defect_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     uint64_t n16, const uint8_t* __restrict__ src_b,
                                                     uint8_t* __restrict__ dst_b, uint64_t nbytes, int* ctl) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 4) ctl[t] = 0;
    for (uint64_t i = t; i < n16; i += stride) dst[i] = src[i];
    for (uint64_t i = n16 * 16u + t; i < nbytes; i += stride) dst_b[i] = src_b[i];
}

template <bool U16>
__global__ void __launch_bounds__(256) k_defect_pass(DefectArgs a) {
    __shared__ uint16_t cur[kLH][kLW];
    __shared__ uint16_t org[kLH][kLW];
    __shared__ uint32_t s_reach;

    const int tile = blockIdx.x;
    const int tx = tile % a.ntx, ty = tile / a.ntx;
    const int tid = threadIdx.x;
    uint8_t* flags_now = a.flags + (size_t)(a.pass & 1) * (size_t)(a.ntx * a.nty);
    const uint8_t* flags_prev = a.flags + (size_t)((a.pass + 1) & 1) * (size_t)(a.ntx * a.nty);

    // Pass k clears the ring slot pass k+1 sets; pass k-1's stays readable.
    if (tile == 0 && tid == 0) a.ctl[(a.pass + 1) % 3] = 0;
    bool run = true;
    if (a.pass > 0) {
        if (a.ctl[(a.pass + 2) % 3] == 0) run = false;  // pass k-1 changed nothing: converged
        else {
            uint32_t in = 0;
            if (tx > 0) in |= flags_prev[tile - 1] & kRight;
            if (tx + 1 < a.ntx) in |= flags_prev[tile + 1] & kLeft;
            if (ty > 0) {
                const int up = tile - a.ntx;
                in |= flags_prev[up] & kBelow;
                if (tx > 0) in |= flags_prev[up - 1] & kBelowRight;
                if (tx + 1 < a.ntx) in |= flags_prev[up + 1] & kBelowLeft;
            }
            run = in != 0;
        }
    }
    if (!run) {
        if (tid == 0) flags_now[tile] = 0;
        return;
    }

    const int x0 = tx * kTX, y0 = ty * kTY;
    const int w = a.w, h = a.h;
    // Load tile + halo; positions outside the interior [2, w-2) x [2, h-2) hold kPad.
    const uint8_t* cur_src = a.pass == 0 ? a.orig : a.data;
    for (int i = tid; i < kLH * (kTX + 4); i += 256) {
        const int r = i / (kTX + 4), c = i - r * (kTX + 4);
        const int x = x0 - 2 + c, y = y0 - 2 + r;
        uint32_t o = kPad, v = kPad;
        if (x >= 2 && x < w - 2 && y >= 2 && y < h - 2) {
            const size_t gi = (size_t)y * (size_t)w + (size_t)x;
            o = ld_px<U16>(a.orig, gi);
            v = a.pass == 0 ? o : ld_px<U16>(cur_src, gi);
        }
        org[r][c] = (uint16_t)o;
        cur[r][c] = (uint16_t)v;
    }
    if (tid == 0) s_reach = 0;
    __syncthreads();

    // This thread: columns x0 + 4g .. +3, rows y0 + 8 * wave + par + {0, 2, 4, 6}.
    const int lane = tid & 63, wave = tid >> 6;
    const int g = lane & 31, par = lane >> 5;
    const int c0 = 4 * g + 2;  // LDS column of x0 + 4g
    int nvalid[4][4];          // valid neighbour count per pixel, 0 = not in the interior
    uint16_t start[4][4];      // values at load time
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int y = y0 + 8 * wave + par + 2 * s;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int x = x0 + 4 * g + j;
            int n = 0;
            if (x >= 2 && x < w - 2 && y >= 2 && y < h - 2) {
                const int nx = 1 + (x - 2 >= 2) + (x + 2 < w - 2), ny = 1 + (y - 2 >= 2) + (y + 2 < h - 2);
                n = nx * ny - 1;
            }
            nvalid[s][j] = n;
            start[s][j] = cur[8 * wave + par + 2 * s + 2][c0 + j];
        }
    }

    int iters = 0;
    bool again;
    do {
        bool changed = false;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int r = 8 * wave + par + 2 * s + 2;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n = nvalid[s][j];
                if (n == 0) continue;
                const int c = c0 + j;
                const uint32_t med =
                    median_n2(cur[r - 2][c - 2], cur[r - 2][c], cur[r - 2][c + 2], cur[r][c - 2], org[r][c + 2],
                              org[r + 2][c - 2], org[r + 2][c], org[r + 2][c + 2], n);
                const uint32_t center = org[r][c];
                const uint32_t diff = center > med ? center - med : med - center;
                const uint32_t nv = (int)diff > a.thresh ? med : center;
                if (nv != cur[r][c]) {
                    cur[r][c] = (uint16_t)nv;
                    changed = true;
                }
            }
        }
        ++iters;
        again = __syncthreads_or(changed) != 0;
    } while (again && iters < kLocalCap);
    if (again && tid == 0) a.ctl[3] = 1;  // cannot happen: the chain bound is below the cap

    uint32_t reach = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int r = 8 * wave + par + 2 * s + 2;
        const int y = y0 + r - 2;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = c0 + j;
            const uint32_t v = cur[r][c];
            if (nvalid[s][j] == 0 || v == start[s][j]) continue;
            const int x = x0 + c - 2;
            st_px<U16>(a.data, (size_t)y * (size_t)w + (size_t)x, v);
            // successors: (x+2, y) and (x-2 .. x+2, y+2)
            const bool right = x + 2 >= x0 + kTX, left = x - 2 < x0, down = y + 2 >= y0 + kTY;
            reach |= kAny;
            if (right) reach |= down ? (kRight | kBelowRight) : kRight;
            if (left) reach |= down ? kBelowLeft : kLeft;
            if (down) reach |= kBelow;
        }
    }
    if (reach) atomicOr(&s_reach, reach);
    __syncthreads();
    if (tid == 0) {
        flags_now[tile] = (uint8_t)s_reach;
        if (s_reach) a.ctl[a.pass % 3] = 1;
    }
}

}  // namespace

hipError_t launch_defect_copy(const uint8_t* src, uint8_t* dst, uint64_t nbytes, int* ctl, hipStream_t st) {
    const bool vec = (((uintptr_t)src | (uintptr_t)dst) & 15u) == 0;
    const uint64_t n16 = vec ? nbytes / 16u : 0u;
    const uint64_t work = n16 ? n16 : nbytes;
    const uint64_t blocks = (work + 255u) / 256u;
    const dim3 grid((unsigned)(blocks < 8192u ? (blocks ? blocks : 1u) : 8192u)), block(256);
    hipLaunchKernelGGL(k_defect_copy, grid, block, 0, st, (const uint4*)src, (uint4*)dst, n16, src, dst, nbytes,
                       ctl);
    return hipGetLastError();
}

int defect_tiles_x(int w) { return (w + kTX - 1) / kTX; }
int defect_tiles_y(int h) { return (h + kTY - 1) / kTY; }

hipError_t launch_defect_pass(uint8_t* data, const uint8_t* orig, uint8_t* flags, int* ctl, int w, int h, int bd,
                              int thresh, int pass, hipStream_t st) {
    DefectArgs a;
    a.data = data;
    a.orig = orig;
    a.flags = flags;
    a.ctl = ctl;
    a.w = w;
    a.h = h;
    a.thresh = thresh;
    a.ntx = defect_tiles_x(w);
    a.nty = defect_tiles_y(h);
    a.pass = pass;
    const dim3 grid((unsigned)(a.ntx * a.nty)), block(256);
    if (bd > 8) hipLaunchKernelGGL((k_defect_pass<true>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_defect_pass<false>), grid, block, 0, st, a);
    return hipGetLastError();
}

}  // namespace imop_hip

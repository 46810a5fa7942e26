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
//                  fixed; then writes its changed pixels to the frame (a tile
//                  run again in a later pass starts from the pairs next to its
//                  halo: the rest were consistent when it last ran).  A tile
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

constexpr int kTX = 128, kTY = 32;      // tile (pixels)
constexpr int kLH = kTY + 4;            // LDS rows: y0-2 .. y0+TY+1
constexpr int kQ = kTX / 2 + 2;         // pixel pairs per LDS row: x0-2 .. x0+TX+1
constexpr int kQW = kQ + 1;             // row pitch (pairs; odd: the two half-waves on different banks)
constexpr uint32_t kPad = 0xFFFFu;      // outside the interior: sorts after every value
constexpr uint32_t kPad2 = 0xFFFFFFFFu;
constexpr int kLocalCap = 1024;         // > the longest chain inside a tile (TX/2 + TY + 2)

// Bits a tile records for the tiles its changes can reach.
enum : uint32_t { kRight = 1, kLeft = 2, kBelowLeft = 4, kBelow = 8, kBelowRight = 16, kAny = 32 };

// Two pixels of one row, an even and an odd column (two colour planes, never
// each other's neighbours), as the halves of a 32-bit word.
typedef uint16_t px2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ px2 as2(uint32_t v) { return __builtin_bit_cast(px2, v); }
__device__ __forceinline__ uint32_t as32(px2 v) { return __builtin_bit_cast(uint32_t, v); }

template <typename T>
__device__ __forceinline__ void ce(T& a, T& b) {
    const T lo = __builtin_elementwise_min(a, b), hi = __builtin_elementwise_max(a, b);
    a = lo;
    b = hi;
}

// Sorts four values in place (five compare-exchanges).
template <typename T>
__device__ __forceinline__ void sort4(T& a0, T& a1, T& a2, T& a3) {
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

// F on a pair whose two pixels both have eight valid neighbours, in packed
// 16-bit arithmetic: the median (k = 5), then med where |c - med| > t, else c.
__device__ __forceinline__ uint32_t f_pair8(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t b0,
                                            uint32_t b1, uint32_t b2, uint32_t b3, uint32_t c, px2 t) {
    px2 A0 = as2(a0), A1 = as2(a1), A2 = as2(a2), A3 = as2(a3);
    px2 B0 = as2(b0), B1 = as2(b1), B2 = as2(b2), B3 = as2(b3);
    sort4(A0, A1, A2, A3);
    sort4(B0, B1, B2, B3);
    const px2 med = __builtin_elementwise_min(
        __builtin_elementwise_min(__builtin_elementwise_max(A0, B3), __builtin_elementwise_max(A1, B2)),
        __builtin_elementwise_min(__builtin_elementwise_max(A2, B1), __builtin_elementwise_max(A3, B0)));
    const px2 C = as2(c);
    const px2 diff = __builtin_elementwise_max(C, med) - __builtin_elementwise_min(C, med);
    const px2 over = __builtin_elementwise_sub_sat(diff, t);                    // > 0 where diff > t
    const uint32_t m = as32((px2){0, 0} - __builtin_elementwise_min(over, (px2){1, 1}));  // 0xFFFF there
    return (as32(med) & m) | (c & ~m);
}

// F on a pair near the interior's edge, pixel by pixel (n from the position;
// pixels outside the interior keep their value).
__device__ __noinline__ uint32_t f_pair_edge(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t b0,
                                             uint32_t b1, uint32_t b2, uint32_t b3, uint32_t c, uint32_t cur,
                                             int x, int y, int w, int h, int t) {
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int sh = 16 * k, px = x + k;
        uint32_t v = (cur >> sh) & 0xFFFFu;
        if (px >= 2 && px < w - 2 && y >= 2 && y < h - 2) {
            const int nx = 1 + (px - 2 >= 2) + (px + 2 < w - 2), ny = 1 + (y - 2 >= 2) + (y + 2 < h - 2);
            const int n = nx * ny - 1;
            if (n > 0) {
                const uint32_t med = median_n2((a0 >> sh) & 0xFFFFu, (a1 >> sh) & 0xFFFFu, (a2 >> sh) & 0xFFFFu,
                                               (a3 >> sh) & 0xFFFFu, (b0 >> sh) & 0xFFFFu, (b1 >> sh) & 0xFFFFu,
                                               (b2 >> sh) & 0xFFFFu, (b3 >> sh) & 0xFFFFu, n);
                const uint32_t cc = (c >> sh) & 0xFFFFu;
                const uint32_t diff = cc > med ? cc - med : med - cc;
                v = (int)diff > t ? med : cc;
            }
        }
        out |= v << sh;
    }
    return out;
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

// The pair at (x, y) (x even) of a frame, kPad for pixels outside the
// interior.  aligned: w even and the frame's base aligned to a pair, so the
// pair is one load (the interior then holds whole pairs).
template <bool U16>
__device__ __forceinline__ uint32_t ld_pair(const uint8_t* p, int x, int y, int w, int h, bool aligned) {
    if (y < 2 || y >= h - 2 || x < 2 || x >= w - 2) return kPad2;
    const size_t i = (size_t)y * (size_t)w + (size_t)x;
    if (aligned) {
        if (U16) return *(const uint32_t*)((const uint16_t*)p + i);
        const uint32_t v = *(const uint16_t*)(p + i);
        return (v & 0xFFu) | ((v >> 8) << 16);
    }
    const uint32_t lo = ld_px<U16>(p, i);
    const uint32_t hi = x + 1 < w - 2 ? ld_px<U16>(p, i + 1) : kPad;
    return lo | (hi << 16);
}

struct DefectArgs {
    uint8_t* data;       // the frame, in place
    const uint8_t* orig; // its input values (scratch copy)
    uint8_t* flags;      // 2 x ntiles: reach bits of the previous / this pass
    int* ctl;            // [0..2] "pass changed something" ring, [3] error
    int w, h, thresh, ntx, nty, pass, aligned;
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

// One tile to its local fixed point, then its changed pixels to the frame.
// EDGE: the tile holds pixels with fewer than eight valid neighbours (within
// 4 of the frame's edge) or outside the interior; otherwise every pair takes
// the packed path.
template <bool U16, bool EDGE>
__device__ __forceinline__ void defect_tile(const DefectArgs& a, int tile, int x0, int y0, uint32_t (*cur)[kQW],
                                            uint32_t (*org)[kQW], uint16_t (*seen)[kQW], uint32_t* s_reach) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int w = a.w, h = a.h;
    const bool aligned = a.aligned != 0;
    // Load tile + halo: wave v takes rows v, v+4, ..., lane l the pair l (and
    // lanes 0, 1 the pairs 64, 65); all loads issued before the LDS stores.
    // Without EDGE the halo lies inside the interior: no bounds to check.
    uint32_t vo[9][2], vc[9][2];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int r = wave + 4 * i, y = y0 - 2 + r;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int q = lane + 64 * e, x = x0 - 2 + 2 * q;
            vo[i][e] = vc[i][e] = kPad2;
            if (e == 0 || lane < kQ - 64) {
                if (!EDGE && aligned) {
                    const size_t gi = (size_t)y * (size_t)w + (size_t)x;
                    if (U16) {
                        vo[i][e] = *(const uint32_t*)((const uint16_t*)a.orig + gi);
                        if (a.pass > 0) vc[i][e] = *(const uint32_t*)((const uint16_t*)a.data + gi);
                    } else {
                        const uint32_t o = *(const uint16_t*)(a.orig + gi);
                        vo[i][e] = (o & 0xFFu) | ((o >> 8) << 16);
                        if (a.pass > 0) {
                            const uint32_t v = *(const uint16_t*)(a.data + gi);
                            vc[i][e] = (v & 0xFFu) | ((v >> 8) << 16);
                        }
                    }
                } else {
                    vo[i][e] = ld_pair<U16>(a.orig, x, y, w, h, aligned);
                    if (a.pass > 0) vc[i][e] = ld_pair<U16>(a.data, x, y, w, h, aligned);
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int r = wave + 4 * i;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int q = lane + 64 * e;
            if (e == 0 || lane < kQ - 64) {
                org[r][q] = vo[i][e];
                cur[r][q] = a.pass > 0 ? vc[i][e] : vo[i][e];
                seen[r][q] = 0;
            }
        }
    }
    __syncthreads();

    // This thread: pairs q = 2g+1, 2g+2 (pixels x0 + 4g .. +3) on rows
    // y0 + 8 * wave + par + {0, 2, 4, 6}.
    const int g = lane & 31, par = lane >> 5;
    uint32_t kinds = 0;  // EDGE: 2 bits per pair: 0 outside the interior, 1 edge, 2 eight neighbours
    uint32_t start[4][2];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int r = 8 * wave + par + 2 * s + 2, y = y0 + r - 2;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int q = 2 * g + 1 + k, x = x0 - 2 + 2 * q;
            if (EDGE) {
                uint32_t kind = 0;
                if (x >= 4 && x + 1 < w - 4 && y >= 4 && y < h - 4) kind = 2;
                else if (y >= 2 && y < h - 2 && x + 1 >= 2 && x < w - 2) kind = 1;
                kinds |= kind << (2 * (2 * s + k));
            }
            start[s][k] = cur[r][q];
        }
    }
    const px2 t2 = (px2){(uint16_t)a.thresh, (uint16_t)a.thresh};

    int it = 0;
    bool again;
    do {
        ++it;
        bool changed = false;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int r = 8 * wave + par + 2 * s + 2;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const uint32_t kind = EDGE ? (kinds >> (2 * (2 * s + k))) & 3u : 2u;
                if (EDGE && kind == 0) continue;
                const int q = 2 * g + 1 + k;
                // A tile re-run in a later pass sat at its fixed point with the halo it
                // loaded then, and only it writes its pixels: only pairs with an input
                // in the halo (top two rows; first and last pair columns, through
                // (x-2, y) and (x+2, y-2)) can change first.
                if (it == 1 && a.pass > 0 && r >= 4 && q != 1 && q != kQ - 2) continue;
                if (it > 1) {  // unchanged inputs since the last evaluation: same value
                    const int last = max(max((int)seen[r - 2][q - 1], (int)seen[r - 2][q]),
                                         max((int)seen[r - 2][q + 1], (int)seen[r][q - 1]));
                    if (last < it - 1) continue;
                }
                const uint32_t a0 = cur[r - 2][q - 1], a1 = cur[r - 2][q], a2 = cur[r - 2][q + 1], a3 = cur[r][q - 1];
                const uint32_t b0 = org[r][q + 1], b1 = org[r + 2][q - 1], b2 = org[r + 2][q], b3 = org[r + 2][q + 1];
                const uint32_t c = org[r][q], old = cur[r][q];
                uint32_t nv;
                if (!EDGE || kind == 2) nv = f_pair8(a0, a1, a2, a3, b0, b1, b2, b3, c, t2);
                else nv = f_pair_edge(a0, a1, a2, a3, b0, b1, b2, b3, c, old, x0 - 2 + 2 * q, y0 + r - 2, w, h,
                                      a.thresh);
                if (nv != old) {
                    cur[r][q] = nv;
                    seen[r][q] = (uint16_t)it;
                    changed = true;
                }
            }
        }
        again = __syncthreads_or(changed) != 0;
    } while (again && it < kLocalCap);
    if (again && tid == 0) a.ctl[3] = 1;  // cannot happen: the chain bound is below the cap

    uint32_t reach = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int r = 8 * wave + par + 2 * s + 2;
        const int y = y0 + r - 2;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int q = 2 * g + 1 + k;
            const uint32_t v = cur[r][q], v0 = start[s][k];
            if ((EDGE && ((kinds >> (2 * (2 * s + k))) & 3u) == 0) || v == v0) continue;
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const uint32_t nvh = (v >> (16 * hh)) & 0xFFFFu;
                if (nvh == ((v0 >> (16 * hh)) & 0xFFFFu)) continue;
                const int x = x0 - 2 + 2 * q + hh;
                st_px<U16>(a.data, (size_t)y * (size_t)w + (size_t)x, nvh);
                // successors: (x+2, y) and (x-2 .. x+2, y+2)
                const bool right = x + 2 >= x0 + kTX, left = x - 2 < x0, down = y + 2 >= y0 + kTY;
                reach |= kAny;
                if (right) reach |= down ? (kRight | kBelowRight) : kRight;
                if (left) reach |= down ? kBelowLeft : kLeft;
                if (down) reach |= kBelow;
            }
        }
    }
    if (reach) atomicOr(s_reach, reach);
}

template <bool U16>
__global__ void __launch_bounds__(256) k_defect_pass(DefectArgs a) {
    __shared__ uint32_t cur[kLH][kQW];   // current values (pairs)
    __shared__ uint32_t org[kLH][kQW];   // input values (pairs)
    __shared__ uint16_t seen[kLH][kQW];  // local iteration in which the pair last changed (0: not yet)
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
    if (tid == 0) s_reach = 0;
    const int x0 = tx * kTX, y0 = ty * kTY;
    // every pixel of the tile with eight valid neighbours: [4, w-4) x [4, h-4)
    if (x0 >= 4 && x0 + kTX <= a.w - 4 && y0 >= 4 && y0 + kTY <= a.h - 4)
        defect_tile<U16, false>(a, tile, x0, y0, cur, org, seen, &s_reach);
    else
        defect_tile<U16, true>(a, tile, x0, y0, cur, org, seen, &s_reach);
    __syncthreads();
    if (tid == 0) {
        flags_now[tile] = (uint8_t)s_reach;
        if (s_reach) a.ctl[a.pass % 3] = 1;
    }
}

// After `passes` passes: *status = 0 when the last pass changed nothing (the
// frame is the fixed point), 1 when it still changed pixels, 2 when a tile
// reached its local iteration cap.
__global__ void k_defect_status(const int* ctl, int passes, int* status) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *status = ctl[3] ? 2 : (ctl[(passes - 1) % 3] == 0 ? 0 : 1);
}

}  // namespace

hipError_t launch_defect_status(const int* ctl, int passes, int* status, hipStream_t st) {
    hipLaunchKernelGGL(k_defect_status, dim3(1), dim3(64), 0, st, ctl, passes, status);
    return hipGetLastError();
}

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
    // a pixel pair is one aligned load in both frames (the scratch copy is 256-byte aligned)
    a.aligned = (w % 2 == 0) && ((uintptr_t)data % (bd > 8 ? 4u : 2u) == 0);
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

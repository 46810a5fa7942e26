/*
 * imop_oracle.c -- CPU restatement of im-operator's Bayer demosaic hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP path is compared
 * against; it is never linked into, loaded by, or called from the product
 * library (im-operator_amd/).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * What it restates (reference files are cited per function):
 *   - pixel decode: unpacked u8 / u16 (MSB-aligned fix-up), 12-bit packed
 *     (3 B / 2 px), 10-bit packed (5 B / 4 px), generic LSB-first bitstream,
 *     GigE Vision 10-bit packing (3 B / 2 px; is_packed = 2)
 *                                      include/imop/pixel_utils.hpp:58-189
 *   - global / per-channel black level  src/black_level/global.cpp:6-24,
 *                                       src/black_level/per_channel.cpp:6-33
 *   - HQLI ("bilinear")                 src/algorithms/hqli.cpp:7-44,
 *                                       src/algorithms/interp_core.hpp:28-68,
 *                                       src/algorithms/avx2.cpp:257-352
 *   - MG ("Malvar-He-Cutler")           src/algorithms/mg.cpp:8-217,
 *                                       src/algorithms/avx2.cpp:422-687
 *   - PRISM (edge-directed)             src/algorithms/prism.cpp:25-316
 *   - SUPER_FAST (nearest / 2-4 average) src/algorithms/super_fast.cpp:9-43,
 *                                       avx2.cpp:99-253
 *   - L7 (7x7 distance-weighted)        src/algorithms/l7.cpp:6-128, avx2.cpp:362-420
 *   - AHD (adaptive homogeneity)        src/algorithms/ahd.cpp:7-187
 *   - DFPD (directional green, chroma)  src/algorithms/dfpd.cpp:7-263
 *   - AMAZE (gradient-weighted green)   src/algorithms/amaze.cpp:7-203,
 *                                       common.hpp:143-214
 *   - RCD (ratio-corrected directional) src/algorithms/rcd.cpp:7-248 (PRISM's
 *                                       stages S0-S6 without the Laplacian
 *                                       term, then borders and output)
 *   - border replication                src/algorithms/common.hpp:47-129
 *
 * The reference CPU library is built with -mavx2 (CMakeLists.txt:222-229), so
 * has_avx2() is true on every host this runs on and the 8-bit MG path is the
 * AVX2 one (G plane initialised to 0, clamps at 255).  The scalar variant
 * (G initialised to -1, clamps at max_val; mg.cpp:23-161) is selectable with
 * orc_set_mg8_scalar(1).
 *
 * Floating-point paths (MG > 8-bit, PRISM) keep the reference's operation
 * order.  This file is compiled with -ffp-contract=off and the multiply-adds
 * that GCC fuses in the reference build (-O3 -mavx2 -mfma; GCC contracts
 * C++ by default even under -std=c++17) are written as explicit fmaf()
 * calls, see mad() below.
 *
 * Pinning: every function here is checked in tests/test_oracle.py against
 * the reference's own known-answer tests, closed forms, the golden digests in
 * tests/golden/ (produced from the compiled reference) and, where
 * oracle/_ref/ is built, against the compiled reference directly.
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#ifdef _OPENMP
#include <omp.h>
#define ORC_PAR _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PAR
#endif

/* DemosaicError values, include/imop/types.hpp:31-39 */
enum { E_OK = 0, E_NULL = 1, E_DIMS = 2, E_BITDEPTH = 3, E_PATTERN = 4, E_SMALL = 5, E_INTERNAL = 6 };
/* DemosaicAlgorithm values, include/imop/types.hpp:19-29 */
enum { A_SUPER_FAST = 0, A_HQLI = 1, A_MG = 2, A_L7 = 3, A_DFPD = 4, A_AHD = 5, A_AMAZE = 6, A_RCD = 7, A_PRISM = 8 };

/* Colour-site roles used throughout: 0 = R, 1 = G, 2 = B. */
typedef struct { int rr, rc, br, bc; } phase_t;

/* PatternOffsets::from_pattern, include/imop/types.hpp:88-101 */
static phase_t phase_of(int pattern) {
    phase_t p;
    switch (pattern) {
    case 1:  p.rr = 1; p.rc = 1; p.br = 0; p.bc = 0; break; /* BGGR */
    case 2:  p.rr = 0; p.rc = 1; p.br = 1; p.bc = 0; break; /* GRBG */
    case 3:  p.rr = 1; p.rc = 0; p.br = 0; p.bc = 1; break; /* GBRG */
    default: p.rr = 0; p.rc = 0; p.br = 1; p.bc = 1; break; /* RGGB */
    }
    return p;
}

static inline int site(const phase_t *p, int y, int x) {
    if ((y & 1) == p->rr && (x & 1) == p->rc) return 0;
    if ((y & 1) == p->br && (x & 1) == p->bc) return 2;
    return 1;
}

/* pixel::safe_max_val, pixel_utils.hpp:15-19 */
int orc_max_val(int bd) {
    if (bd <= 0) return 0;
    if (bd > 16) return 65535;
    return (int)((1u << bd) - 1u);
}

/* compute_packed_byte_size / compute_bayer_byte_size / compute_rgb_byte_size,
 * pixel_utils.hpp:21-34 */
size_t orc_packed_bytes(int w, int h, int bd) { return ((size_t)w * (size_t)h * (size_t)bd + 7u) / 8u; }
size_t orc_bayer_bytes(int w, int h, int bd, int packed) {
    if (packed == 2) return ((size_t)w * (size_t)h + 1u) / 2u * 3u;
    if (packed) return orc_packed_bytes(w, h, bd);
    return (size_t)w * (size_t)h * (bd <= 8 ? 1u : 2u);
}
size_t orc_rgb_bytes(int w, int h, int bd) { return (size_t)w * (size_t)h * 3u * (bd > 8 ? 2u : 1u); }

/* Decode pixel number `idx` (linear, no row padding) from a packed stream of
 * `nbytes` bytes.  get_packed_raw, pixel_utils.hpp:58-118: 12-bit groups are
 * [P0 hi8][P1 lo4 | P0 lo4][P1 hi8]; 10-bit groups are four high bytes then
 * one byte of 2-bit tails (P0 in bits 1:0); anything else is an LSB-first
 * bitstream.  A group that would touch the last byte of the stream decodes
 * as 0 (the reference's ">=" bound). */
int orc_packed_pixel(const uint8_t *d, size_t idx, int bd, size_t nbytes) {
    if (bd < 1 || bd > 16) return 0;
    const int mx = orc_max_val(bd);
    if (bd == 12) {
        size_t o = idx / 2u * 3u;
        if (o + 2u >= nbytes) return 0;
        if ((idx & 1u) == 0) return ((d[o] << 4) | (d[o + 1] & 0x0F)) & mx;
        return ((d[o + 2] << 4) | (d[o + 1] >> 4)) & mx;
    }
    if (bd == 10) {
        size_t o = idx / 4u * 5u;
        unsigned s = (unsigned)(idx & 3u);
        if (o + 4u >= nbytes) return 0;
        return ((d[o + s] << 2) | ((d[o + 4] >> (2u * s)) & 3)) & mx;
    }
    size_t bit = idx * (size_t)bd;
    size_t o = bit / 8u;
    int sh = (int)(bit % 8u);
    int nb = (bd + sh + 7) / 8;
    if (nb < 1) nb = 1;
    if (o >= nbytes) return 0;
    if ((size_t)nb > nbytes - o) nb = (int)(nbytes - o);
    uint32_t v = 0;
    for (int i = 0; i < nb; ++i) v |= (uint32_t)d[o + (size_t)i] << (8 * i);
    return (int)((v >> sh) & (uint32_t)mx);
}

/* GigE Vision 10-bit packing (BayerXX10Packed), the layout of the
 * reference's test_data/<capture>_pBayerRG10_Packed.raw, which its own 10-bit rule
 * (4 px / 5 B, pixel_utils.hpp:82-98) does not decode: 3-byte groups
 * [P0 9:2][P1 1:0 << 4 | P0 1:0][P1 9:2], as the GigE Vision pixel format
 * defines them.  A group past the stream decodes as 0. */
int orc_gige10_pixel(const uint8_t *d, size_t idx, size_t nbytes) {
    size_t o = idx / 2u * 3u;
    if (o + 2u >= nbytes) return 0;
    if ((idx & 1u) == 0) return (d[o] << 2) | (d[o + 1] & 3);
    return (d[o + 2] << 2) | ((d[o + 1] >> 4) & 3);
}

/* align_raw_value, pixel_utils.hpp:129-137: MSB-aligned 9..15-bit samples
 * (value above max) are shifted down. */
static inline int align_u16(unsigned v, int bd) {
    if (bd > 8 && bd < 16) {
        unsigned mx = (unsigned)orc_max_val(bd);
        if (v > mx) v >>= (16 - bd);
    }
    return (int)v;
}

/* Decode the whole frame to a plane of ints the way pixel::get_raw /
 * get_clamped read it (pixel_utils.hpp:139-175): bytes for bd <= 8 (packed or
 * not), the packed decoder for bd > 8 packed, aligned little-endian u16
 * otherwise. */
void orc_decode_plane(const uint8_t *in, int w, int h, int bd, int packed, int *plane) {
    const size_t n = (size_t)w * (size_t)h;
    if (bd <= 8) {
        for (size_t i = 0; i < n; ++i) plane[i] = in[i];
    } else if (packed == 2) {
        const size_t nb = orc_bayer_bytes(w, h, bd, 2);
        for (size_t i = 0; i < n; ++i) plane[i] = orc_gige10_pixel(in, i, nb);
    } else if (packed) {
        const size_t nb = orc_packed_bytes(w, h, bd);
        for (size_t i = 0; i < n; ++i) plane[i] = orc_packed_pixel(in, i, bd, nb);
    } else {
        for (size_t i = 0; i < n; ++i) {
            unsigned v = (unsigned)in[2 * i] | ((unsigned)in[2 * i + 1] << 8);
            plane[i] = align_u16(v, bd);
        }
    }
}

/* Unpack a packed (or unpacked) frame into a plain u16 / u8 buffer, the
 * caller-side step that precedes black level in the reference composition:
 * value = pixel::get_raw(in, x, y, w, bd, is_packed, 0, h). */
void orc_unpack(const uint8_t *in, int w, int h, int bd, int packed, uint8_t *out) {
    const size_t n = (size_t)w * (size_t)h;
    int *pl = (int *)malloc(n * sizeof(int));
    orc_decode_plane(in, w, h, bd, packed, pl);
    if (bd <= 8) {
        for (size_t i = 0; i < n; ++i) out[i] = (uint8_t)pl[i];
    } else {
        for (size_t i = 0; i < n; ++i) { out[2 * i] = (uint8_t)(pl[i] & 0xFF); out[2 * i + 1] = (uint8_t)((unsigned)pl[i] >> 8); }
    }
    free(pl);
}

/* Black level in place on an unpacked u8/u16 Bayer buffer.
 * process_global (global.cpp:6-24) uses offs[0] for every site;
 * process_per_channel (per_channel.cpp:6-33) picks R / Gr / Gb / B.
 * Values are read raw (no MSB alignment, black_level/common.hpp:143-147),
 * new = trunc(float(v) - max(0, off)), clamped to [0, max]. */
void orc_black_level(uint8_t *data, int w, int h, int pattern, int bd, const float offs[4], int per_channel) {
    const phase_t p = phase_of(pattern);
    const int mx = bd <= 0 ? 255 : (bd > 16 ? 65535 : (int)((1u << bd) - 1u));
    float o[4];
    for (int i = 0; i < 4; ++i) o[i] = offs[per_channel ? i : 0] > 0.0f ? offs[per_channel ? i : 0] : 0.0f;
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * (size_t)w + (size_t)x;
            int v = bd <= 8 ? data[i] : (int)((unsigned)data[2 * i] | ((unsigned)data[2 * i + 1] << 8));
            int c;
            if ((y & 1) == p.rr && (x & 1) == p.rc) c = 0;
            else if ((y & 1) == p.br && (x & 1) == p.bc) c = 3;
            else if ((y & 1) == p.rr) c = 1;
            else c = 2;
            int nv = (int)((float)v - o[c]);
            if (nv < 0) nv = 0;
            if (nv > mx) nv = mx;
            if (bd <= 8) data[i] = (uint8_t)nv;
            else { data[2 * i] = (uint8_t)(nv & 0xFF); data[2 * i + 1] = (uint8_t)((unsigned)nv >> 8); }
        }
    }
}

static inline void put_rgb(uint8_t *rgb, int w, int x, int y, int bd, int r, int g, int b) {
    size_t i = ((size_t)y * (size_t)w + (size_t)x) * 3u;
    if (bd <= 8) { rgb[i] = (uint8_t)r; rgb[i + 1] = (uint8_t)g; rgb[i + 2] = (uint8_t)b; }
    else {
        uint16_t *q = (uint16_t *)rgb;
        q[i] = (uint16_t)r; q[i + 1] = (uint16_t)g; q[i + 2] = (uint16_t)b;
    }
}

/* fill_rgb_borders, common.hpp:47-96: top rows copy row `b` whole, bottom
 * rows copy row h-1-b whole, then columns of the middle rows copy column b /
 * w-1-b.  The four corner blocks therefore keep row b's (or h-1-b's) edge
 * pixels as they were before the column step. */
static void rgb_borders(uint8_t *rgb, int w, int h, int bd, int b) {
    if (b <= 0 || w <= 2 * b || h <= 2 * b) return;
    const size_t px = bd <= 8 ? 3u : 6u;
    const size_t stride = (size_t)w * px;
    for (int y = 0; y < b; ++y) memcpy(rgb + (size_t)y * stride, rgb + (size_t)b * stride, stride);
    for (int y = h - b; y < h; ++y) memcpy(rgb + (size_t)y * stride, rgb + (size_t)(h - 1 - b) * stride, stride);
    for (int y = b; y < h - b; ++y) {
        uint8_t *row = rgb + (size_t)y * stride;
        for (int x = 0; x < b; ++x) memcpy(row + (size_t)x * px, row + (size_t)b * px, px);
        for (int x = w - b; x < w; ++x) memcpy(row + (size_t)x * px, row + (size_t)(w - 1 - b) * px, px);
    }
}

/* fill_intermediate_borders_t, common.hpp:98-129, on three planes. */
#define DEFINE_PLANE_BORDERS(NAME, T)                                                    \
    static void NAME(T *R, T *G, T *B, int w, int h, int b) {                            \
        if (b <= 0 || w <= 2 * b || h <= 2 * b) return;                                 \
        for (int y = 0; y < b; ++y) {                                                    \
            size_t d = (size_t)y * w, s = (size_t)b * w;                                 \
            for (int x = 0; x < w; ++x) { R[d + x] = R[s + x]; G[d + x] = G[s + x]; B[d + x] = B[s + x]; } \
        }                                                                                \
        for (int y = h - b; y < h; ++y) {                                                \
            size_t d = (size_t)y * w, s = (size_t)(h - 1 - b) * w;                       \
            for (int x = 0; x < w; ++x) { R[d + x] = R[s + x]; G[d + x] = G[s + x]; B[d + x] = B[s + x]; } \
        }                                                                                \
        for (int y = b; y < h - b; ++y) {                                                \
            size_t r0 = (size_t)y * w, l = r0 + (size_t)b, r = r0 + (size_t)(w - 1 - b); \
            for (int x = 0; x < b; ++x) { R[r0 + x] = R[l]; G[r0 + x] = G[l]; B[r0 + x] = B[l]; } \
            for (int x = w - b; x < w; ++x) { R[r0 + x] = R[r]; G[r0 + x] = G[r]; B[r0 + x] = B[r]; } \
        }                                                                                \
    }
DEFINE_PLANE_BORDERS(plane_borders_i, int)
DEFINE_PLANE_BORDERS(plane_borders_f, float)

static int g_mg8_scalar = 0;
void orc_set_mg8_scalar(int on) { g_mg8_scalar = on; }

/* ------------------------------------------------------------ SUPER_FAST */
/* process_super_fast (super_fast.cpp:9-43) on an AVX2 host, which is where
 * the reference takes process_super_fast_optimized (avx2.cpp:209-253):
 *   R/B site: own value, green = (l + r + u + d + 2) >> 2, the other chroma
 *             from the diagonal (x +- 1 toward the pair, y +- 1), clamped;
 *   G site:   own value, the row's chroma = (l + r + 1) >> 1, the column's
 *             chroma = (u + d + 1) >> 1;
 * every neighbour clamped into the frame.  8-bit unpacked frames take
 * super_fast_avx2_8bit (avx2.cpp:99-207): 32-pixel vectors whose "left"
 * operand is loaded from x - 1 except in a row's first vector (loaded from
 * 0: pixels 1..31 see themselves as their left neighbour), and whose "right"
 * operand comes from x + 1 except in a vector starting at w - 32 (when w is
 * a multiple of 32: pixels w-32..w-1 see themselves); the scalar tail is
 * exact.  g_sf_scalar = 1
 * gives the scalar core (interp_core.hpp:73-118) instead: the exact
 * neighbours everywhere. */
static int g_sf_scalar = 0;
void orc_set_super_fast_scalar(int on) { g_sf_scalar = on; }

static void super_fast(const int *P, uint8_t *rgb, int w, int h, phase_t ph, int bd, int byte_path) {
    const int mx = orc_max_val(bd);
    const int quirks = byte_path && !g_sf_scalar;
    const int full = (w / 32) * 32;  /* pixels covered by 32-pixel vectors */
    for (int y = 0; y < h; ++y) {
        const int yu = y > 0 ? y - 1 : 0, yd = y < h - 1 ? y + 1 : h - 1;
        for (int x = 0; x < w; ++x) {
            const int xl = x > 0 ? x - 1 : 0, xr = x < w - 1 ? x + 1 : w - 1;
            int l = P[(size_t)y * w + xl], r = P[(size_t)y * w + xr];
            if (quirks && x < full) {
                if (x < 32) l = P[(size_t)y * w + x];
                if ((x & ~31) == w - 32) r = P[(size_t)y * w + x];
            }
            const int c = P[(size_t)y * w + x], u = P[(size_t)yu * w + x], d = P[(size_t)yd * w + x];
            const int s = site(&ph, y, x);
            int R, G, B;
            if (s != 1) {
                int nx = (x & 1) ? x - 1 : x + 1, ny = (y & 1) ? y - 1 : y + 1;
                nx = nx < 0 ? 0 : (nx >= w ? w - 1 : nx);
                ny = ny < 0 ? 0 : (ny >= h ? h - 1 : ny);
                const int diag = P[(size_t)ny * w + nx];
                G = (l + r + u + d + 2) >> 2;
                R = s == 0 ? c : diag;
                B = s == 0 ? diag : c;
            } else {
                G = c;
                const int hz = (l + r + 1) >> 1, vt = (u + d + 1) >> 1;
                const int gr_row = (y & 1) == ph.rr;
                R = gr_row ? hz : vt;
                B = gr_row ? vt : hz;
            }
            if (!byte_path) { /* set_rgb_clamped; the 8-bit routine stores the bytes as computed */
                R = R > mx ? mx : R; G = G > mx ? mx : G; B = B > mx ? mx : B;
            }
            put_rgb(rgb, w, x, y, bd, R, G, B);
        }
    }
}

/* -------------------------------------------------------------------- L7 */
/* process_l7 (l7.cpp:65-128): inside [3, w-3) x [3, h-3), each colour the
 * pixel lacks is the rounded integer mean (sum + wsum / 2) / wsum of that
 * colour's samples in the 7x7 window, weighted by kL7Weights7
 * (l7.cpp:6-14: (4 - max(|dx|, |dy|))^2 ... the table itself; centre 0); the
 * pixel's own colour is its sample.  8-bit unpacked frames (l7_avx2_8bit /
 * l7_core_8bit, the same arithmetic) clamp to [0, 255]; the generic path to
 * [0, max].  Then fill_rgb_borders with width 3. */
static const int l7_weights[7][7] = {
    {1, 4, 9, 16, 9, 4, 1},   {4, 9, 16, 25, 16, 9, 4},   {9, 16, 25, 36, 25, 16, 9}, {16, 25, 36, 0, 36, 25, 16},
    {9, 16, 25, 36, 25, 16, 9}, {4, 9, 16, 25, 16, 9, 4}, {1, 4, 9, 16, 9, 4, 1}};

static void l7(const int *P, uint8_t *rgb, int w, int h, phase_t ph, int bd, int byte_path) {
    const int hi = byte_path ? 255 : orc_max_val(bd);
    for (int y = 3; y < h - 3; ++y)
        for (int x = 3; x < w - 3; ++x) {
            int sum[3] = {0, 0, 0}, ws[3] = {0, 0, 0};
            for (int dy = -3; dy <= 3; ++dy)
                for (int dx = -3; dx <= 3; ++dx) {
                    const int wv = l7_weights[dy + 3][dx + 3];
                    if (wv == 0) continue;
                    const int c = site(&ph, y + dy, x + dx);
                    sum[c] += P[(size_t)(y + dy) * w + (x + dx)] * wv;
                    ws[c] += wv;
                }
            const int own = site(&ph, y, x), v = P[(size_t)y * w + x];
            int o[3];
            for (int c = 0; c < 3; ++c) {
                o[c] = c == own ? v : (ws[c] > 0 ? (sum[c] + ws[c] / 2) / ws[c] : 0);
                o[c] = o[c] < 0 ? 0 : (o[c] > hi ? hi : o[c]);
            }
            put_rgb(rgb, w, x, y, bd, o[0], o[1], o[2]);
        }
    rgb_borders(rgb, w, h, bd, 3);
}

/* ------------------------------------------------------------------ HQLI */
/* process_hqli (hqli.cpp:7-44): 3x3 bilinear averages on [2,w-2)x[2,h-2)
 * with integer rounding, then fill_rgb_borders(2).  The AVX2 8-bit path
 * (avx2.cpp:257-352) computes the same averages without the final clamp;
 * the generic core clamps to max_val (interp_core.hpp:28-68). */
static void hqli(const int *P, uint8_t *rgb, int w, int h, phase_t ph, int bd, int packed) {
    const int mx = orc_max_val(bd);
    const int clamp_on = !(bd <= 8 && !packed);
    ORC_PAR
    for (int y = 2; y < h - 2; ++y) {
        const int in_r_row = ((y & 1) == ph.rr);
        for (int x = 2; x < w - 2; ++x) {
            const int *c = P + (size_t)y * w + x;
            int r, g, b;
            int s = site(&ph, y, x);
            int cross = c[-1] + c[1] + c[-w] + c[w];
            int diag = c[-w - 1] + c[-w + 1] + c[w - 1] + c[w + 1];
            if (s == 0) { r = c[0]; g = (cross + 2) / 4; b = (diag + 2) / 4; }
            else if (s == 2) { r = (diag + 2) / 4; g = (cross + 2) / 4; b = c[0]; }
            else {
                int hz = (c[-1] + c[1] + 1) / 2, vt = (c[-w] + c[w] + 1) / 2;
                g = c[0];
                if (in_r_row) { r = hz; b = vt; } else { r = vt; b = hz; }
            }
            if (clamp_on) {
                r = r < 0 ? 0 : (r > mx ? mx : r);
                g = g < 0 ? 0 : (g > mx ? mx : g);
                b = b < 0 ? 0 : (b > mx ? mx : b);
            }
            put_rgb(rgb, w, x, y, bd, r, g, b);
        }
    }
    rgb_borders(rgb, w, h, bd, 2);
}

/* ---------------------------------------------------------- MG, 8-bit int */
/* mg_avx2_8bit (avx2.cpp:422-595) / scalar mg.cpp:23-161.
 * P1 G = raw at G sites (sentinel elsewhere); P2 G at interior R/B sites by
 * the 5x5 MHC green kernel with +4 >> 3 rounding and a clamp; P3 R/B on
 * [1,w-1)x[1,h-1) by colour-difference averages rounded with C's truncating
 * division (sum + cnt/2) / cnt; plane borders 2; clamp and store u8. */
static void mg_int8(const int *P, uint8_t *rgb, int w, int h, phase_t ph, int bd) {
    const size_t n = (size_t)w * h;
    const int scalar = g_mg8_scalar;
    const int gmax = scalar ? orc_max_val(bd) : 255;
    int *G = (int *)malloc(n * sizeof(int));
    int *R = (int *)calloc(n, sizeof(int));
    int *B = (int *)calloc(n, sizeof(int));
    for (size_t i = 0; i < n; ++i) G[i] = scalar ? -1 : 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            if (site(&ph, y, x) == 1) G[(size_t)y * w + x] = P[(size_t)y * w + x];
    ORC_PAR
    for (int y = 2; y < h - 2; ++y)
        for (int x = 2; x < w - 2; ++x) {
            if (site(&ph, y, x) == 1) continue;
            const int *c = P + (size_t)y * w + x;
            int near = c[-w] + c[w] + c[1] + c[-1];
            int far = c[-2 * w] + c[2 * w] + c[2] + c[-2];
            int g = (4 * c[0] + 2 * near - far + 4) >> 3;
            G[(size_t)y * w + x] = g < 0 ? 0 : (g > gmax ? gmax : g);
        }
    ORC_PAR
    for (int y = 1; y < h - 1; ++y)
        for (int x = 1; x < w - 1; ++x) {
            size_t i = (size_t)y * w + x;
            int s = site(&ph, y, x);
            if (s != 1) {
                /* opposite colour sits on the four diagonals */
                int sum = 0, cnt = 0;
                const int dy[4] = {-1, -1, 1, 1}, dx[4] = {-1, 1, -1, 1};
                for (int k = 0; k < 4; ++k) {
                    int yy = y + dy[k], xx = x + dx[k];
                    if (site(&ph, yy, xx) == 2 - s) { size_t j = (size_t)yy * w + xx; sum += P[j] - G[j]; ++cnt; }
                }
                int other = G[i] + (cnt > 0 ? (sum + cnt / 2) / cnt : 0);
                if (s == 0) { R[i] = P[i]; B[i] = other; } else { B[i] = P[i]; R[i] = other; }
            } else {
                int rs = 0, rn = 0, bs = 0, bn = 0;
                const int dy[4] = {-1, 1, 0, 0}, dx[4] = {0, 0, -1, 1};
                for (int k = 0; k < 4; ++k) {
                    int yy = y + dy[k], xx = x + dx[k];
                    size_t j = (size_t)yy * w + xx;
                    int t = site(&ph, yy, xx);
                    if (t == 0) { rs += P[j] - G[j]; ++rn; }
                    else if (t == 2) { bs += P[j] - G[j]; ++bn; }
                }
                R[i] = G[i] + (rn > 0 ? (rs + rn / 2) / rn : 0);
                B[i] = G[i] + (bn > 0 ? (bs + bn / 2) / bn : 0);
            }
        }
    plane_borders_i(R, G, B, w, h, 2);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            int r = R[i], g = G[i], b = B[i];
            r = r < 0 ? 0 : (r > gmax ? gmax : r);
            g = g < 0 ? 0 : (g > gmax ? gmax : g);
            b = b < 0 ? 0 : (b > gmax ? gmax : b);
            put_rgb(rgb, w, x, y, 8, r, g, b);
        }
    free(G); free(R); free(B);
}

/* ------------------------------------------------------- MG, float path */
/* process_mg_optimized float branch (avx2.cpp:606-686) == mg.cpp:163-216:
 * G = -1 sentinel, raw at G sites, (4c + 2*near - far) / 8 at interior R/B
 * sites; R/B on [1,w-1)x[1,h-1) = G + mean(raw - G) over the opposite
 * colour's diagonals (R/B sites) or the two same-colour 4-neighbours
 * (G sites); plane borders 2; int(v + 0.5f) clamped. */
static void mg_float(const int *P, uint8_t *rgb, int w, int h, phase_t ph, int bd) {
    const size_t n = (size_t)w * h;
    const int mx = orc_max_val(bd);
    float *G = (float *)malloc(n * sizeof(float));
    float *R = (float *)calloc(n, sizeof(float));
    float *B = (float *)calloc(n, sizeof(float));
    for (size_t i = 0; i < n; ++i) G[i] = -1.0f;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            if (site(&ph, y, x) == 1) G[(size_t)y * w + x] = (float)P[(size_t)y * w + x];
    ORC_PAR
    for (int y = 2; y < h - 2; ++y)
        for (int x = 2; x < w - 2; ++x) {
            if (site(&ph, y, x) == 1) continue;
            const int *c = P + (size_t)y * w + x;
            float nn = (float)c[-w], ss = (float)c[w], ee = (float)c[1], ww = (float)c[-1];
            float n2 = (float)c[-2 * w], s2 = (float)c[2 * w], e2 = (float)c[2], w2 = (float)c[-2];
            float cc = (float)c[0];
            G[(size_t)y * w + x] = (4.0f * cc + 2.0f * (nn + ss + ee + ww) - (n2 + s2 + e2 + w2)) / 8.0f;
        }
    ORC_PAR
    for (int y = 1; y < h - 1; ++y)
        for (int x = 1; x < w - 1; ++x) {
            size_t i = (size_t)y * w + x;
            float g = G[i];
            int s = site(&ph, y, x);
            if (s != 1) {
                float sum = 0.0f; int cnt = 0;
                const int dy[4] = {-1, -1, 1, 1}, dx[4] = {-1, 1, -1, 1};
                for (int k = 0; k < 4; ++k) {
                    int yy = y + dy[k], xx = x + dx[k];
                    if (site(&ph, yy, xx) == 2 - s) { size_t j = (size_t)yy * w + xx; sum += (float)P[j] - G[j]; ++cnt; }
                }
                float other = g + (cnt > 0 ? sum / (float)cnt : 0.0f);
                if (s == 0) { R[i] = (float)P[i]; B[i] = other; } else { B[i] = (float)P[i]; R[i] = other; }
            } else {
                float rs = 0.0f, bs = 0.0f; int rn = 0, bn = 0;
                const int dy[4] = {-1, 1, 0, 0}, dx[4] = {0, 0, -1, 1};
                for (int k = 0; k < 4; ++k) {
                    int yy = y + dy[k], xx = x + dx[k];
                    size_t j = (size_t)yy * w + xx;
                    int t = site(&ph, yy, xx);
                    if (t == 0) { rs += (float)P[j] - G[j]; ++rn; }
                    else if (t == 2) { bs += (float)P[j] - G[j]; ++bn; }
                }
                R[i] = g + (rn > 0 ? rs / (float)rn : 0.0f);
                B[i] = g + (bn > 0 ? bs / (float)bn : 0.0f);
            }
        }
    plane_borders_f(R, G, B, w, h, 2);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            int r = (int)(R[i] + 0.5f), g = (int)(G[i] + 0.5f), b = (int)(B[i] + 0.5f);
            r = r > mx ? mx : r; g = g > mx ? mx : g; b = b > mx ? mx : b;
            r = r < 0 ? 0 : r; g = g < 0 ? 0 : g; b = b < 0 ? 0 : b;
            put_rgb(rgb, w, x, y, bd, r, g, b);
        }
    free(G); free(R); free(B);
}

/* ------------------------------------------------------------------ PRISM */
/* The reference library is compiled by GCC with -O3 -mavx2 -mfma
 * (CMakeLists.txt:45,222-229).  GCC contracts a*b + c into one fused
 * multiply-add in C++ even under -std=c++17, always taking the first
 * product of a sum (a*b + c*d -> fma(a, b, c*d)).  mad() reproduces that for
 * the float PRISM stages; orc_set_prism_fma(0) gives the uncontracted
 * arithmetic of a build without FMA instead.  (Power-of-two products, where
 * contraction cannot change a result, are written plainly.) */
static int g_prism_fma = 1;
void orc_set_prism_fma(int on) { g_prism_fma = on; }
/* A planted mistake for the tests: RCD with PRISM's Laplacian term kept. */
static int g_rcd_laplacian = 0;
void orc_set_rcd_laplacian(int on) { g_rcd_laplacian = on; }
static inline float mad(float a, float b, float c) { return g_prism_fma ? fmaf(a, b, c) : a * b + c; }

/* std::min / std::max semantics (first argument wins ties). */
static inline float fmin_first(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_first(float a, float b) { return (a < b) ? b : a; }
static inline float clamp01(float v) { return fmax_first(0.0f, fmin_first(1.0f, v)); }

/* prism_detail::median_3x3 (prism.cpp:9-21): the same 19-exchange network,
 * element 4 of the partially sorted window is returned. */
static float median9(float *v) {
    static const unsigned char net[23][2] = {
        {0, 1}, {3, 4}, {6, 7}, {1, 2}, {4, 5}, {7, 8}, {0, 1}, {3, 4}, {6, 7}, {0, 3}, {4, 7}, {1, 4},
        {3, 6}, {2, 5}, {5, 8}, {1, 3}, {4, 6}, {2, 5}, {2, 3}, {5, 6}, {2, 4}, {3, 5}, {3, 4}};
    for (int k = 0; k < 23; ++k) {
        float t = v[net[k][0]];
        if (t > v[net[k][1]]) { v[net[k][0]] = v[net[k][1]]; v[net[k][1]] = t; }
    }
    return v[4];
}

/* process_prism (prism.cpp:25-316), stage by stage.  Planes: cfa, the three
 * colour planes, VH and PQ direction maps, the R/B low-pass, and the two
 * colour-difference planes with their median-filtered copies.
 * rcd: process_rcd (rcd.cpp:7-248), whose stages are the same statements
 * (rcd.cpp:43-224 = prism.cpp:68-264) except that its green estimate has no
 * Laplacian term (prism.cpp:132-136) and it has no colour-difference median
 * (prism.cpp:266-300): S6's planes go straight to the border fill. */
static void prism(const int *P, uint8_t *rgb, int w, int h, phase_t ph, int bd, int rcd) {
    const size_t n = (size_t)w * h;
    const int mx = orc_max_val(bd);
    const float scale = (float)mx;
    const float eps = 1e-5f, epssq = 1e-10f;
    float *cfa = (float *)calloc(n, sizeof(float));
    float *C[3];
    for (int k = 0; k < 3; ++k) C[k] = (float *)calloc(n, sizeof(float));
    float *VH = (float *)malloc(n * sizeof(float));
    float *PQ = (float *)malloc(n * sizeof(float));
    float *lpf = (float *)calloc(n, sizeof(float));
    float *L = (float *)calloc(n, sizeof(float));
    float *M = (float *)calloc(n, sizeof(float));
    float *Lm = (float *)calloc(n, sizeof(float));
    float *Mm = (float *)calloc(n, sizeof(float));
    for (size_t i = 0; i < n; ++i) { VH[i] = 0.5f; PQ[i] = 0.5f; }
    float *G = C[1];

    /* S0: normalised CFA and native-colour planes (prism.cpp:68-78) */
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            float v = (float)P[i] / scale;
            cfa[i] = v;
            C[site(&ph, y, x)][i] = v;
        }

    /* S1: vertical / horizontal discrimination over a 9-tap cross (:84-99) */
    ORC_PAR
    for (int y = 4; y < h - 4; ++y)
        for (int x = 4; x < w - 4; ++x) {
            size_t i = (size_t)y * w + x;
            float vs = 0.0f, hs = 0.0f;
            for (int k = -4; k <= 4; ++k) {
                vs += fabsf(cfa[i] - cfa[i + (ptrdiff_t)k * w]);
                hs += fabsf(cfa[i] - cfa[i + k]);
            }
            vs = fmax_first(vs, epssq);
            hs = fmax_first(hs, epssq);
            VH[i] = vs / (vs + hs);
        }

    /* S2: 3x3 low-pass at R/B sites (:101-112) */
    ORC_PAR
    for (int y = 2; y < h - 2; ++y) {
        int x0 = 2 + (site(&ph, y, 0) & 1);
        for (int x = x0; x < w - 2; x += 2) {
            size_t i = (size_t)y * w + x;
            float s4 = cfa[i - w] + cfa[i + w] + cfa[i - 1] + cfa[i + 1];
            float d4 = cfa[i - w - 1] + cfa[i - w + 1] + cfa[i + w - 1] + cfa[i + w + 1];
            lpf[i] = mad(0.0625f, d4, mad(0.25f, cfa[i], 0.125f * s4));
        }
    }

    /* S3: ratio-corrected directional green at R/B with Laplacian term (:114-163) */
    ORC_PAR
    for (int y = 4; y < h - 4; ++y) {
        int x0 = 4 + (site(&ph, y, 0) & 1);
        for (int x = x0; x < w - 4; x += 2) {
            size_t i = (size_t)y * w + x;
            const ptrdiff_t W = w;
            float vc = VH[i];
            float vn = 0.25f * (VH[i - W - 1] + VH[i - W + 1] + VH[i + W - 1] + VH[i + W + 1]);
            float disc = (fabsf(0.5f - vc) < fabsf(0.5f - vn)) ? vn : vc;
            float c0 = cfa[i];
            float gN = eps + fabsf(cfa[i - W] - cfa[i + W]) + fabsf(c0 - cfa[i - 2 * W]) + fabsf(cfa[i - W] - cfa[i - 3 * W]);
            float gS = eps + fabsf(cfa[i + W] - cfa[i - W]) + fabsf(c0 - cfa[i + 2 * W]) + fabsf(cfa[i + W] - cfa[i + 3 * W]);
            float gW = eps + fabsf(cfa[i - 1] - cfa[i + 1]) + fabsf(c0 - cfa[i - 2]) + fabsf(cfa[i - 1] - cfa[i - 3]);
            float gE = eps + fabsf(cfa[i + 1] - cfa[i - 1]) + fabsf(c0 - cfa[i + 2]) + fabsf(cfa[i + 1] - cfa[i + 3]);
            /* lpf taps at +-2 along the axes; the reference clamps their
             * coordinates to [2, dim-3], which never binds on [4, dim-4). */
            float l0 = lpf[i], lN = lpf[i - 2 * W], lS = lpf[i + 2 * W], lW = lpf[i - 2], lE = lpf[i + 2];
            float eN = cfa[i - W] * (1.0f + (l0 - lN) / (eps + l0 + lN));
            float eS = cfa[i + W] * (1.0f + (l0 - lS) / (eps + l0 + lS));
            float eW = cfa[i - 1] * (1.0f + (l0 - lW) / (eps + l0 + lW));
            float eE = cfa[i + 1] * (1.0f + (l0 - lE) / (eps + l0 + lE));
            float vEst = mad(gS, eN, gN * eS) / (gN + gS);
            float hEst = mad(gW, eE, gE * eW) / (gE + gW);
            if (!rcd || g_rcd_laplacian) {
                float lapV = (2.0f * c0 - cfa[i - 2 * W] - cfa[i + 2 * W]) * 0.25f;
                float lapH = (2.0f * c0 - cfa[i - 2] - cfa[i + 2]) * 0.25f;
                vEst += lapV;
                hEst += lapH;
            }
            G[i] = clamp01(mad(disc, hEst, (1.0f - disc) * vEst));
        }
    }

    /* S4: diagonal discrimination at R/B (:165-179) */
    ORC_PAR
    for (int y = 4; y < h - 4; ++y) {
        int x0 = 4 + (site(&ph, y, 0) & 1);
        for (int x = x0; x < w - 4; x += 2) {
            size_t i = (size_t)y * w + x;
            float ps = 0.0f, qs = 0.0f;
            for (int k = -4; k <= 4; ++k) {
                ps += fabsf(cfa[i] - cfa[i + (ptrdiff_t)k * w + k]);
                qs += fabsf(cfa[i] - cfa[i + (ptrdiff_t)k * w - k]);
            }
            ps = fmax_first(ps, epssq);
            qs = fmax_first(qs, epssq);
            PQ[i] = ps / (ps + qs);
        }
    }

    /* S5: the opposite chroma at R/B sites along the diagonals (:181-225) */
    ORC_PAR
    for (int y = 4; y < h - 4; ++y) {
        int x0 = 4 + (site(&ph, y, 0) & 1);
        for (int x = x0; x < w - 4; x += 2) {
            size_t i = (size_t)y * w + x;
            const ptrdiff_t W = w;
            float *X = C[2 - site(&ph, y, x)];
            float pc = PQ[i];
            float pn = 0.25f * (PQ[i - W - 1] + PQ[i - W + 1] + PQ[i + W - 1] + PQ[i + W + 1]);
            float disc = (fabsf(0.5f - pc) < fabsf(0.5f - pn)) ? pn : pc;
            float g0 = G[i];
            float gNW = eps + fabsf(X[i - W - 1] - X[i + W + 1]) + fabsf(g0 - G[i - 2 * W - 2]) + fabsf(X[i - W - 1] - X[i - 3 * W - 3]);
            float gNE = eps + fabsf(X[i - W + 1] - X[i + W - 1]) + fabsf(g0 - G[i - 2 * W + 2]) + fabsf(X[i - W + 1] - X[i - 3 * W + 3]);
            float gSW = eps + fabsf(X[i + W - 1] - X[i - W + 1]) + fabsf(g0 - G[i + 2 * W - 2]) + fabsf(X[i + W - 1] - X[i + 3 * W - 3]);
            float gSE = eps + fabsf(X[i + W + 1] - X[i - W - 1]) + fabsf(g0 - G[i + 2 * W + 2]) + fabsf(X[i + W + 1] - X[i + 3 * W + 3]);
            float eNW = X[i - W - 1] - G[i - W - 1];
            float eNE = X[i - W + 1] - G[i - W + 1];
            float eSW = X[i + W - 1] - G[i + W - 1];
            float eSE = X[i + W + 1] - G[i + W + 1];
            float pEst = mad(gNW, eSE, gSE * eNW) / (gNW + gSE);
            float qEst = mad(gNE, eSW, gSW * eNE) / (gNE + gSW);
            X[i] = clamp01(mad(disc, qEst, mad(1.0f - disc, pEst, g0)));
        }
    }

    /* S6: R and B at G sites, directional colour differences (:227-264) */
    ORC_PAR
    for (int y = 4; y < h - 4; ++y) {
        int x0 = 4 + (site(&ph, y, 1) & 1);
        for (int x = x0; x < w - 4; x += 2) {
            size_t i = (size_t)y * w + x;
            const ptrdiff_t W = w;
            float vc = VH[i];
            float vn = 0.25f * (VH[i - W - 1] + VH[i - W + 1] + VH[i + W - 1] + VH[i + W + 1]);
            float disc = (fabsf(0.5f - vc) < fabsf(0.5f - vn)) ? vn : vc;
            for (int ch = 0; ch <= 2; ch += 2) {
                float *X = C[ch];
                float g0 = G[i];
                float gN = eps + fabsf(g0 - G[i - 2 * W]) + fabsf(X[i - W] - X[i + W]) + fabsf(X[i - W] - X[i - 3 * W]);
                float gS = eps + fabsf(g0 - G[i + 2 * W]) + fabsf(X[i + W] - X[i - W]) + fabsf(X[i + W] - X[i + 3 * W]);
                float gW = eps + fabsf(g0 - G[i - 2]) + fabsf(X[i - 1] - X[i + 1]) + fabsf(X[i - 1] - X[i - 3]);
                float gE = eps + fabsf(g0 - G[i + 2]) + fabsf(X[i + 1] - X[i - 1]) + fabsf(X[i + 1] - X[i + 3]);
                float dN = X[i - W] - G[i - W];
                float dS = X[i + W] - G[i + W];
                float dW = X[i - 1] - G[i - 1];
                float dE = X[i + 1] - G[i + 1];
                float vEst = mad(gS, dN, gN * dS) / (gN + gS);
                float hEst = mad(gW, dE, gE * dW) / (gE + gW);
                X[i] = clamp01(mad(1.0f - disc, vEst, mad(disc, hEst, g0)));
            }
        }
    }

    /* S7: colour differences (:266-269) */
    if (!rcd)
    for (size_t i = 0; i < n; ++i) { L[i] = C[0][i] - G[i]; M[i] = C[2][i] - G[i]; }

    /* S8: 3x3 median of both differences into separate planes; the outer
     * ring stays 0 and replaces L/M wholesale (:273-292) */
    if (!rcd)
    for (int y = 1; y < h - 1; ++y)
        for (int x = 1; x < w - 1; ++x) {
            float a[9], b[9];
            int k = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    size_t j = (size_t)(y + dy) * w + (size_t)(x + dx);
                    a[k] = L[j]; b[k] = M[j]; ++k;
                }
            size_t i = (size_t)y * w + x;
            Lm[i] = median9(a);
            Mm[i] = median9(b);
        }

    /* S9: reconstruct R and B from the filtered differences (:294-300) */
    if (!rcd)
    for (int y = 1; y < h - 1; ++y)
        for (int x = 1; x < w - 1; ++x) {
            size_t i = (size_t)y * w + x;
            C[0][i] = clamp01(G[i] + Lm[i]);
            C[2][i] = clamp01(G[i] + Mm[i]);
        }

    /* S10 / S11: plane borders 4, scale, round, clamp (:302-313; rcd.cpp:226-245) */
    plane_borders_f(C[0], C[1], C[2], w, h, 4);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            int r = (int)mad(C[0][i], scale, 0.5f);
            int g = (int)mad(C[1][i], scale, 0.5f);
            int b = (int)mad(C[2][i], scale, 0.5f);
            r = r > mx ? mx : r; g = g > mx ? mx : g; b = b > mx ? mx : b;
            r = r < 0 ? 0 : r; g = g < 0 ? 0 : g; b = b < 0 ? 0 : b;
            put_rgb(rgb, w, x, y, bd, r, g, b);
        }

    free(cfa); free(C[0]); free(C[1]); free(C[2]); free(VH); free(PQ); free(lpf);
    free(L); free(M); free(Lm); free(Mm);
}

/* ------------------------------------------------------------------- AHD */
/* process_ahd (ahd.cpp:43-187): horizontal and vertical bilinear estimates
 * (samples clamped into the frame; halves and quarters of integer sums,
 * exact in float), their colour differences R - G and B - G, each filtered
 * by the 3x3 exchange network of apply_median_filter_3x3 (ahd.cpp:7-40, the
 * same network as prism_detail::median_3x3) on [1, w-1) x [1, h-1), then per
 * pixel the direction whose filtered differences have the smaller sum of 3x3
 * variances (coordinates clamped; mean = sum / 9, var = sum of diff^2 with
 * GCC's contraction fmaf(diff, diff, var), / 9); strict < picks horizontal.
 * Output (int)(v + 0.5f) clamped to [0, max]. */
/* Planted mistakes for the tests: 1 = ties go horizontal (<=), 2 = no median. */
static int g_ahd_mistake = 0;
void orc_set_ahd_mistake(int k) { g_ahd_mistake = k; }

static void ahd(const int *P, uint8_t *rgb, int w, int h, phase_t ph, int bd) {
    const size_t n = (size_t)w * h;
    const int mx = orc_max_val(bd);
    float *pl[10];
    for (int k = 0; k < 10; ++k) pl[k] = (float *)calloc(n, sizeof(float));
    float *Rh = pl[0], *Gh = pl[1], *Bh = pl[2], *Rv = pl[3], *Gv = pl[4], *Bv = pl[5];
    float *Lm[4] = {pl[6], pl[7], pl[8], pl[9]};  /* Lh, Mh, Lv, Mv */
#define RAW(X, Y) ((float)P[(size_t)((Y) < 0 ? 0 : (Y) >= h ? h - 1 : (Y)) * w + (size_t)((X) < 0 ? 0 : (X) >= w ? w - 1 : (X))])
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            const int s = site(&ph, y, x);
            for (int dir = 0; dir < 2; ++dir) {
                float r, g, b;
                if (s == 1) {
                    g = RAW(x, y);
                    const float hz = (RAW(x - 1, y) + RAW(x + 1, y)) * 0.5f, vt = (RAW(x, y - 1) + RAW(x, y + 1)) * 0.5f;
                    /* the chroma whose samples sit beside the pixel along dir */
                    const int r_beside = dir == 0 ? (site(&ph, y, x - 1) == 0 || site(&ph, y, x + 1) == 0)
                                                  : (site(&ph, y - 1, x) == 0 || site(&ph, y + 1, x) == 0);
                    const float along = dir == 0 ? hz : vt, across = dir == 0 ? vt : hz;
                    r = r_beside ? along : across;
                    b = r_beside ? across : along;
                } else {
                    const float own = RAW(x, y);
                    g = dir == 0 ? (RAW(x - 1, y) + RAW(x + 1, y)) * 0.5f : (RAW(x, y - 1) + RAW(x, y + 1)) * 0.5f;
                    const float diag = dir == 0
                        ? (RAW(x - 1, y - 1) + RAW(x + 1, y - 1) + RAW(x - 1, y + 1) + RAW(x + 1, y + 1)) * 0.25f
                        : (RAW(x - 1, y - 1) + RAW(x - 1, y + 1) + RAW(x + 1, y - 1) + RAW(x + 1, y + 1)) * 0.25f;
                    r = s == 0 ? own : diag;
                    b = s == 0 ? diag : own;
                }
                if (dir == 0) { Rh[i] = r; Gh[i] = g; Bh[i] = b; }
                else { Rv[i] = r; Gv[i] = g; Bv[i] = b; }
            }
        }
#undef RAW
    for (size_t i = 0; i < n; ++i) {
        Lm[0][i] = Rh[i] - Gh[i]; Lm[1][i] = Bh[i] - Gh[i];
        Lm[2][i] = Rv[i] - Gv[i]; Lm[3][i] = Bv[i] - Gv[i];
    }
    float *filt = (float *)malloc(n * sizeof(float));
    for (int k = 0; k < 4; ++k) {
        float *p = Lm[k];
        if (w >= 3 && h >= 3 && g_ahd_mistake != 2) {
            for (int y = 1; y < h - 1; ++y)
                for (int x = 1; x < w - 1; ++x) {
                    float v[9];
                    int q = 0;
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) v[q++] = p[(size_t)(y + dy) * w + (x + dx)];
                    filt[(size_t)y * w + x] = median9(v);
                }
            for (int y = 1; y < h - 1; ++y)
                for (int x = 1; x < w - 1; ++x) p[(size_t)y * w + x] = filt[(size_t)y * w + x];
        }
    }
    free(filt);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float var[4];
            for (int k = 0; k < 4; ++k) {
                float mean = 0.0f, v = 0.0f;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int nx = x + dx, ny = y + dy;
                        nx = nx < 0 ? 0 : (nx > w - 1 ? w - 1 : nx);
                        ny = ny < 0 ? 0 : (ny > h - 1 ? h - 1 : ny);
                        mean += Lm[k][(size_t)ny * w + nx];
                    }
                mean /= 9.0f;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int nx = x + dx, ny = y + dy;
                        nx = nx < 0 ? 0 : (nx > w - 1 ? w - 1 : nx);
                        ny = ny < 0 ? 0 : (ny > h - 1 ? h - 1 : ny);
                        const float d = Lm[k][(size_t)ny * w + nx] - mean;
                        v = mad(d, d, v);
                    }
                var[k] = v / 9.0f;
            }
            const size_t i = (size_t)y * w + x;
            const int horiz = g_ahd_mistake == 1 ? (var[0] + var[1]) <= (var[2] + var[3])
                                                 : (var[0] + var[1]) < (var[2] + var[3]);
            const float fr = horiz ? Rh[i] : Rv[i], fg = horiz ? Gh[i] : Gv[i], fb = horiz ? Bh[i] : Bv[i];
            int o[3] = {(int)(fr + 0.5f), (int)(fg + 0.5f), (int)(fb + 0.5f)};
            for (int c = 0; c < 3; ++c) o[c] = o[c] < 0 ? 0 : (o[c] > mx ? mx : o[c]);
            put_rgb(rgb, w, x, y, bd, o[0], o[1], o[2]);
        }
    for (int k = 0; k < 10; ++k) free(pl[k]);
}

/* ------------------------------------------------------------------ DFPD */
/* process_dfpd (dfpd.cpp:146-263) and its 8-bit core (dfpd.cpp:7-144): green
 * on [2, w-2) x [2, h-2) along the direction of smaller gradient (the
 * horizontal / vertical average plus half the Laplacian correction; the mean
 * of both on a tie), then on [3, w-3) x [3, h-3) the chroma from the mean
 * diagonal (R/B sites, 4 terms) or cross (G sites, 2 terms) colour
 * differences.  The 8-bit core works in integers: (a + b + 1) >> 1 halves,
 * (sum + cnt / 2) / cnt means (C division), clamp to [0, 255].  The generic
 * path works in float, where every value is a multiple of 1/16 and exact:
 * (int)(v + 0.5f) and set_rgb_clamped.  Green is only ever needed where it
 * was computed, so interp_g's recomputation of a sentinel never applies.
 * Then fill_rgb_borders with width 5. */
static void dfpd(const int *P, uint8_t *rgb, int w, int h, phase_t ph, int bd, int byte_path) {
    const size_t n = (size_t)w * h;
    const int mx = byte_path ? 255 : orc_max_val(bd);
    float *G = (float *)calloc(n, sizeof(float)); /* 8-bit: integer values */
#define RV(X, Y) (P[(size_t)(Y) * w + (size_t)(X)])
    for (int y = 2; y < h - 2; ++y)
        for (int x = 2; x < w - 2; ++x) {
            const size_t i = (size_t)y * w + x;
            const int c = RV(x, y);
            if (site(&ph, y, x) == 1) { G[i] = (float)c; continue; }
            const int xm1 = RV(x - 1, y), xp1 = RV(x + 1, y), xm2 = RV(x - 2, y), xp2 = RV(x + 2, y);
            const int ym1 = RV(x, y - 1), yp1 = RV(x, y + 1), ym2 = RV(x, y - 2), yp2 = RV(x, y + 2);
            const int gh = abs(xm1 - xp1) + abs(2 * c - xm2 - xp2), gv = abs(ym1 - yp1) + abs(2 * c - ym2 - yp2);
            if (byte_path) {
                const int g_h = (xm1 + xp1 + 1) >> 1, g_v = (ym1 + yp1 + 1) >> 1;
                int g;
                if (gh < gv) g = g_h + ((c - ((xm2 + xp2 + 1) >> 1) + 1) >> 1);
                else if (gv < gh) g = g_v + ((c - ((ym2 + yp2 + 1) >> 1) + 1) >> 1);
                else g = (g_h + g_v + 1) >> 1;
                G[i] = (float)g;
            } else {
                const float g_h = (float)(xm1 + xp1) * 0.5f, g_v = (float)(ym1 + yp1) * 0.5f;
                if (gh < gv) G[i] = g_h + ((float)c - (float)(xm2 + xp2) * 0.5f) * 0.5f;
                else if (gv < gh) G[i] = g_v + ((float)c - (float)(ym2 + yp2) * 0.5f) * 0.5f;
                else G[i] = (g_h + g_v) * 0.5f;
            }
        }
    for (int y = 3; y < h - 3; ++y)
        for (int x = 3; x < w - 3; ++x) {
            const size_t i = (size_t)y * w + x;
            const int s = site(&ph, y, x);
            const float g = G[i];
            int R, Gc, B;
            static const int dg[4][2] = {{-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
            static const int cr[4][2] = {{0, -1}, {0, 1}, {-1, 0}, {1, 0}};
            if (byte_path) {
                const int gi = (int)g;
                if (s != 1) {
                    int sum = 0, cnt = 0;
                    for (int k = 0; k < 4; ++k) {
                        const int nx = x + dg[k][0], ny = y + dg[k][1];
                        if (site(&ph, ny, nx) == 2 - s) { sum += RV(nx, ny) - (int)G[(size_t)ny * w + nx]; cnt++; }
                    }
                    const int other = cnt > 0 ? gi + (sum + cnt / 2) / cnt : gi;
                    R = s == 0 ? RV(x, y) : other;
                    B = s == 0 ? other : RV(x, y);
                } else {
                    int rs = 0, rc = 0, bs = 0, bc = 0;
                    for (int k = 0; k < 4; ++k) {
                        const int nx = x + cr[k][0], ny = y + cr[k][1];
                        if (site(&ph, ny, nx) == 0) { rs += RV(nx, ny) - (int)G[(size_t)ny * w + nx]; rc++; }
                    }
                    for (int k = 0; k < 4; ++k) {
                        const int nx = x + cr[k][0], ny = y + cr[k][1];
                        if (site(&ph, ny, nx) == 2) { bs += RV(nx, ny) - (int)G[(size_t)ny * w + nx]; bc++; }
                    }
                    R = rc > 0 ? gi + (rs + rc / 2) / rc : gi;
                    B = bc > 0 ? gi + (bs + bc / 2) / bc : gi;
                }
                Gc = gi;
            } else {
                if (s != 1) {
                    float sum = 0.0f;
                    int cnt = 0;
                    for (int k = 0; k < 4; ++k) {
                        const int nx = x + dg[k][0], ny = y + dg[k][1];
                        if (site(&ph, ny, nx) == 2 - s) { sum += (float)RV(nx, ny) - G[(size_t)ny * w + nx]; cnt++; }
                    }
                    const int own = (int)((float)RV(x, y) + 0.5f);
                    const int other = (int)(g + (cnt > 0 ? sum / (float)cnt : 0.0f) + 0.5f);
                    R = s == 0 ? own : other;
                    B = s == 0 ? other : own;
                } else {
                    float rs = 0.0f, bs = 0.0f;
                    int rc = 0, bc = 0;
                    for (int k = 0; k < 4; ++k) {
                        const int nx = x + cr[k][0], ny = y + cr[k][1];
                        if (site(&ph, ny, nx) == 0) { rs += (float)RV(nx, ny) - G[(size_t)ny * w + nx]; rc++; }
                    }
                    for (int k = 0; k < 4; ++k) {
                        const int nx = x + cr[k][0], ny = y + cr[k][1];
                        if (site(&ph, ny, nx) == 2) { bs += (float)RV(nx, ny) - G[(size_t)ny * w + nx]; bc++; }
                    }
                    R = (int)(g + (rc > 0 ? rs / (float)rc : 0.0f) + 0.5f);
                    B = (int)(g + (bc > 0 ? bs / (float)bc : 0.0f) + 0.5f);
                }
                Gc = (int)(g + 0.5f);
            }
            R = R < 0 ? 0 : (R > mx ? mx : R);
            Gc = Gc < 0 ? 0 : (Gc > mx ? mx : Gc);
            B = B < 0 ? 0 : (B > mx ? mx : B);
            put_rgb(rgb, w, x, y, bd, R, Gc, B);
        }
#undef RV
    rgb_borders(rgb, w, h, bd, 5);
    free(G);
}

/* ----------------------------------------------------------------- AMAZE */
/* process_amaze (amaze.cpp:158-200) and its 8-bit core (amaze.cpp:7-156):
 *   G plane starts at -1 everywhere; inside [2, w-2) x [2, h-2) it is the
 *   sample at G sites, else the inverse-gradient blend of the horizontal and
 *   vertical estimates (each with a quarter Laplacian), fmaf(wh, gh, wv * gv)
 *   as GCC contracts `wh * gh + wv * gv`;
 *   inside [1, w-1) x [1, h-1): the chroma at R/B sites from the diagonal
 *   opposite-colour differences, at G sites from the cross neighbours
 *   (interpolate_rb_at_*_positions, common.hpp:143-199), every difference
 *   against the G plane as it stands, sentinel included; R = B = 0 elsewhere;
 *   fill_intermediate_borders(2), then (int)(v + 0.5f) clamped to [0, max]
 *   (the 8-bit core: [0, 255], then fill_rgb_borders(2) on the output). */
static void amaze(const int *P, uint8_t *rgb, int w, int h, phase_t ph, int bd, int byte_path) {
    const size_t n = (size_t)w * h;
    float *G = (float *)malloc(n * sizeof(float)), *R = (float *)calloc(n, sizeof(float)),
          *B = (float *)calloc(n, sizeof(float));
    for (size_t i = 0; i < n; ++i) G[i] = -1.0f;
#define RW(X, Y) ((float)P[(size_t)(Y) * w + (size_t)(X)])
    for (int y = 2; y < h - 2; ++y)
        for (int x = 2; x < w - 2; ++x) {
            const size_t i = (size_t)y * w + x;
            if (site(&ph, y, x) == 1) { G[i] = RW(x, y); continue; }
            const float c = RW(x, y);
            const float gH = fabsf(RW(x - 1, y) - RW(x + 1, y)) + fabsf(2.0f * c - RW(x - 2, y) - RW(x + 2, y));
            const float gV = fabsf(RW(x, y - 1) - RW(x, y + 1)) + fabsf(2.0f * c - RW(x, y - 2) - RW(x, y + 2));
            float wh = 1.0f / (1.0f + gH), wv = 1.0f / (1.0f + gV);
            const float sw = wh + wv;
            wh /= sw;
            wv /= sw;
            float gh = (RW(x - 1, y) + RW(x + 1, y)) * 0.5f, gv = (RW(x, y - 1) + RW(x, y + 1)) * 0.5f;
            const float lH = 2.0f * c - RW(x - 2, y) - RW(x + 2, y), lV = 2.0f * c - RW(x, y - 2) - RW(x, y + 2);
            gh += lH * 0.25f;
            gv += lV * 0.25f;
            G[i] = mad(wh, gh, wv * gv);
        }
    for (int y = 1; y < h - 1; ++y)
        for (int x = 1; x < w - 1; ++x) {
            const size_t i = (size_t)y * w + x;
            const float g = G[i];
            const int s = site(&ph, y, x);
            if (s != 1) {
                float sum = 0.0f;
                int cnt = 0;
                static const int dg[4][2] = {{-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
                for (int k = 0; k < 4; ++k) {
                    const int nx = x + dg[k][0], ny = y + dg[k][1];
                    if (nx >= 0 && nx < w && ny >= 0 && ny < h && site(&ph, ny, nx) == 2 - s) {
                        sum += RW(nx, ny) - G[(size_t)ny * w + nx];
                        cnt++;
                    }
                }
                const float other = g + (cnt > 0 ? sum / (float)cnt : 0.0f);
                if (s == 0) { R[i] = RW(x, y); B[i] = other; }
                else { B[i] = RW(x, y); R[i] = other; }
            } else {
                float rs = 0.0f, bs = 0.0f;
                int rc = 0, bc = 0;
                static const int cr[4][2] = {{0, -1}, {0, 1}, {-1, 0}, {1, 0}};
                for (int k = 0; k < 4; ++k) {
                    const int nx = x + cr[k][0], ny = y + cr[k][1];
                    if (site(&ph, ny, nx) == 0) { rs += RW(nx, ny) - G[(size_t)ny * w + nx]; rc++; }
                }
                for (int k = 0; k < 4; ++k) {
                    const int nx = x + cr[k][0], ny = y + cr[k][1];
                    if (site(&ph, ny, nx) == 2) { bs += RW(nx, ny) - G[(size_t)ny * w + nx]; bc++; }
                }
                R[i] = g + (rc > 0 ? rs / (float)rc : 0.0f);
                B[i] = g + (bc > 0 ? bs / (float)bc : 0.0f);
            }
        }
#undef RW
    plane_borders_f(R, G, B, w, h, 2);
    const int mx = byte_path ? 255 : orc_max_val(bd);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            int o[3] = {(int)(R[i] + 0.5f), (int)(G[i] + 0.5f), (int)(B[i] + 0.5f)};
            for (int c = 0; c < 3; ++c) o[c] = o[c] < 0 ? 0 : (o[c] > mx ? mx : o[c]);
            put_rgb(rgb, w, x, y, bd, o[0], o[1], o[2]);
        }
    if (byte_path) rgb_borders(rgb, w, h, bd, 2);
    free(G); free(R); free(B);
}

/* demosaic_cpu (dispatch.cpp:115-128) restricted to SUPER_FAST / HQLI / MG /
 * L7 / DFPD / AHD / AMAZE / PRISM / RCD (every registry entry):
 * validation order and error codes of validate_demosaic_inputs
 * (dispatch.cpp:14-20), then each algorithm's own minimum size
 * (super_fast.cpp:13, hqli.cpp:11, mg.cpp:12, l7.cpp:69, dfpd.cpp:149,
 * ahd.cpp:46, amaze.cpp:162, prism.cpp:29, rcd.cpp:11).  Other algorithms
 * return InternalError: they are not restated. */
int orc_demosaic(const uint8_t *in, uint8_t *rgb, int w, int h, int pattern, int algo, int bd, int packed) {
    if (!in || !rgb) return E_NULL;
    if (w <= 0 || h <= 0) return E_DIMS;
    if (bd < 1 || bd > 16) return E_BITDEPTH;
    if (pattern < 0 || pattern > 3) return E_PATTERN;
    int min_dim;
    if (algo == A_HQLI || algo == A_MG) min_dim = 6;
    else if (algo == A_SUPER_FAST) min_dim = 2;
    else if (algo == A_L7) min_dim = 8;
    else if (algo == A_AHD || algo == A_AMAZE) min_dim = 6;
    else if (algo == A_DFPD) min_dim = 12;
    else if (algo == A_PRISM || algo == A_RCD) min_dim = 10;
    else return E_INTERNAL;
    if (w < min_dim || h < min_dim) return E_SMALL;
    if (packed < 0 || packed > 2) return E_INTERNAL;
    if (packed == 2 && bd != 10) return E_BITDEPTH;  /* GigE packing is defined for 10 bits */
    const phase_t ph = phase_of(pattern);
    int *P = (int *)malloc((size_t)w * h * sizeof(int));
    if (!P) return E_INTERNAL;
    orc_decode_plane(in, w, h, bd, packed, P);
    if (algo == A_SUPER_FAST) super_fast(P, rgb, w, h, ph, bd, bd <= 8 && !packed);
    else if (algo == A_L7) l7(P, rgb, w, h, ph, bd, bd <= 8 && !packed);
    else if (algo == A_AHD) ahd(P, rgb, w, h, ph, bd);
    else if (algo == A_AMAZE) amaze(P, rgb, w, h, ph, bd, bd <= 8 && !packed);
    else if (algo == A_DFPD) dfpd(P, rgb, w, h, ph, bd, bd <= 8 && !packed);
    else if (algo == A_HQLI) hqli(P, rgb, w, h, ph, bd, packed);
    else if (algo == A_MG) {
        if (bd <= 8 && !packed) mg_int8(P, rgb, w, h, ph, bd);
        else mg_float(P, rgb, w, h, ph, bd);
    } else prism(P, rgb, w, h, ph, bd, algo == A_RCD);
    free(P);
    return E_OK;
}

/* The composed path the fused kernels implement:
 *   black_level == NULL: demosaic(in, packed) directly;
 *   otherwise: packed input is first unpacked with pixel::get_raw (unpacked
 *   input is used as is), then black level runs in place on the unpacked
 *   buffer (per_channel selects process_per_channel, else process_global
 *   with black_level[0]), then the unpacked buffer is demosaiced. */
int orc_pipeline(const uint8_t *in, uint8_t *rgb, int w, int h, int pattern, int algo, int bd, int packed,
                 const float *black_level, int per_channel) {
    if (!black_level) return orc_demosaic(in, rgb, w, h, pattern, algo, bd, packed);
    if (!in || !rgb) return E_NULL;
    if (w <= 0 || h <= 0) return E_DIMS;
    if (bd < 1 || bd > 16) return E_BITDEPTH;
    if (packed == 2 && bd != 10) return E_BITDEPTH;
    size_t ub = orc_bayer_bytes(w, h, bd, 0);
    uint8_t *u = (uint8_t *)malloc(ub);
    if (!u) return E_INTERNAL;
    if (packed) orc_unpack(in, w, h, bd, packed, u);
    else memcpy(u, in, ub);
    orc_black_level(u, w, h, pattern, bd, black_level, per_channel);
    int e = orc_demosaic(u, rgb, w, h, pattern, algo, bd, 0);
    free(u);
    return e;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

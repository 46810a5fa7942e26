/*
 * chain_oracle.c -- FP64 restatement of the differentiable ISP chain
 * (bilinear demosaic -> white balance -> 3x3 CCM -> gamma) and of its
 * vector-Jacobian product.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP chain kernels
 * are compared against; it is never linked into, loaded by, or called from
 * the product library (im-operator_amd/).  FP64 is used here and nowhere in
 * the product path.
 *
 * Forward, per frame (x = raw / max_value):
 *   a = D x                  bilinear demosaic, as a linear operator
 *   u = clamp01(g .* a)      white balance, gpu_ops_kernels.cu:119-125 (op 7)
 *   z = clamp01(M u)         CCM, gpu_ops_kernels.cu:92-96 + dev_ccm_3x3
 *                            (gpu_common.cuh:43-48)
 *   y = clamp01(z)^(1/gamma) gamma, gpu_ops_kernels.cu:57-64 (op 0)
 * D is the bilinear ("HQLI") demosaic of src/algorithms/hqli.cpp:7-44 /
 * interp_core.hpp:28-68 with exact averages instead of the integer rounding
 * (R site: g = mean of the 4 axial neighbours, b = mean of the 4 diagonal
 * ones; B site mirrored; G site: the two same-row / same-column neighbours
 * of each colour), over [2, w-2) x [2, h-2), and the border rule of
 * fill_rgb_borders (common.hpp:47-96): rows 0-1 copy row 2 and rows h-2,
 * h-1 copy row h-3 over the interior columns, the middle rows' columns 0-1
 * copy column 2 and columns w-2, w-1 copy column w-3, and the 16 corner
 * pixels keep the output buffer's prior value, here 0.
 *
 * D is assembled explicitly as a sparse matrix (CSR, one row per output
 * sample (pixel, channel), one column per raw pixel, FP64 weights); the
 * forward applies it and the VJP applies its transpose by walking the rows,
 * so the adjoint is the transposed assembled operator, independent of any
 * closed-form gather.
 *
 * Conventions of the derivative at the clamps and the power: clamp01 passes
 * the gradient where 0 <= v <= 1 (as torch.clamp does); d/dz z^(1/gamma) is
 * taken as 0 at z = 0.  Loss = mean over every output sample of (y - t)^2.
 *
 * The forward colour stages are pinned by closed forms (identity
 * parameters reduce to clamp01(D x); a permutation matrix permutes channels;
 * gamma 2 maps 0.25 to 0.5; g a > 1 clamps before the matrix) and by the
 * integer-domain stages of the reference (manual.cpp, linear_3x3.cpp,
 * gamma.cpp) on inputs where every product is exact; planted mistakes
 * (orc_chain_set_plant) show each of those checks can fail.
 *
 * Backward has no counterpart in the reference: it is pinned by central
 * finite differences of this forward, by the adjoint identity of the
 * assembled operator and by the HQLI equivalence of D (tests/test_chain_oracle.py).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* Colour role of site (x, y): 0 R, 1 G, 2 B (PatternOffsets::from_pattern,
 * include/imop/types.hpp:88-101). */
static int chain_site(int pattern, int x, int y) {
    const int rr = (pattern == 1 || pattern == 3) ? 1 : 0;
    const int rc = (pattern == 1 || pattern == 2) ? 1 : 0;
    if ((y & 1) == rr && (x & 1) == rc) return 0;
    if ((y & 1) != rr && (x & 1) != rc) return 2;
    return 1;
}

/* Source pixel of output pixel (x, y) under fill_rgb_borders(border 2);
 * returns 0 for the corner pixels, which no interior pixel feeds. */
static int chain_src(int w, int h, int x, int y, int* sx, int* sy) {
    const int xin = x >= 2 && x < w - 2;
    if (y < 2 || y >= h - 2) {
        if (!xin) return 0;
        *sx = x;
        *sy = y < 2 ? 2 : h - 3;
        return 1;
    }
    *sy = y;
    *sx = x < 2 ? 2 : (x >= w - 2 ? w - 3 : x);
    return 1;
}

/* Entries of the interior row for output sample (x, y, c). */
static int chain_row_entries(int w, int pattern, int x, int y, int c, int* col, double* val) {
    const int s = chain_site(pattern, x, y);
    const int rr = (pattern == 1 || pattern == 3) ? 1 : 0;
    const int p = y * w + x;
    const int N = p - w, S = p + w, E = p + 1, W = p - 1;
    int n = 0;
#define PUT(q, v) do { col[n] = (q); val[n] = (v); ++n; } while (0)
    if (s == c) {
        PUT(p, 1.0);
    } else if (s != 1 && c == 1) {
        PUT(N, 0.25); PUT(S, 0.25); PUT(E, 0.25); PUT(W, 0.25);
    } else if (s != 1) {  /* opposite colour at R / B */
        PUT(N - 1, 0.25); PUT(N + 1, 0.25); PUT(S - 1, 0.25); PUT(S + 1, 0.25);
    } else {              /* R or B at a G site */
        const int r_row = (y & 1) == rr;
        const int horizontal = (c == 0) == r_row;
        if (horizontal) { PUT(W, 0.5); PUT(E, 0.5); }
        else { PUT(N, 0.5); PUT(S, 0.5); }
    }
#undef PUT
    return n;
}

/* CSR of D for one w x h frame: rows 3 * w * h (pixel-major, channel-minor),
 * columns w * h.  With rowptr == NULL only the number of entries is returned. */
long orc_bilinear_csr(int w, int h, int pattern, long* rowptr, int* col, double* val) {
    long nnz = 0;
    int cbuf[4];
    double vbuf[4];
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) {
                const long r = ((long)y * w + x) * 3 + c;
                if (rowptr) rowptr[r] = nnz;
                int sx, sy;
                if (!chain_src(w, h, x, y, &sx, &sy)) continue;
                const int k = chain_row_entries(w, pattern, sx, sy, c, cbuf, vbuf);
                if (col)
                    for (int j = 0; j < k; ++j) {
                        col[nnz + j] = cbuf[j];
                        val[nnz + j] = vbuf[j];
                    }
                nnz += k;
            }
    if (rowptr) rowptr[(long)w * h * 3] = nnz;
    return nnz;
}

typedef struct {
    long* rowptr;
    int* col;
    double* val;
} csr_t;

static int csr_build(int w, int h, int pattern, csr_t* m) {
    const long nnz = orc_bilinear_csr(w, h, pattern, NULL, NULL, NULL);
    m->rowptr = (long*)malloc(sizeof(long) * ((size_t)w * h * 3 + 1));
    m->col = (int*)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    m->val = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    if (!m->rowptr || !m->col || !m->val) return -1;
    orc_bilinear_csr(w, h, pattern, m->rowptr, m->col, m->val);
    return 0;
}

static void csr_free(csr_t* m) {
    free(m->rowptr);
    free(m->col);
    free(m->val);
}

static double clamp01d(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

/* Planted mistakes, for the tests that show the forward's closed-form checks
 * can fail (0 = the chain as specified; bits may be combined):
 *   1  the colour matrix applied transposed (M^T u)
 *   2  gamma used as the exponent instead of 1 / gamma
 *   4  white balance without its clamp (u = g .* a, clamped only after the CCM)
 *   8  the colour matrix before the white balance (u = clamp01(M a), then g) */
static int g_chain_plant = 0;
void orc_chain_set_plant(int mask) { g_chain_plant = mask; }

/* Forward of one pixel from its demosaiced a[3]; keeps the intermediates. */
typedef struct {
    double a[3], b[3], u[3], c[3], z[3], y[3];
} chain_px;

static void chain_pixel(const double a[3], const double* g, const double* M, double inv_gamma, chain_px* s) {
    const int tr = g_chain_plant & 1;
    const double e = (g_chain_plant & 2) ? 1.0 / inv_gamma : inv_gamma;
    if (g_chain_plant & 8) {
        double t[3];
        for (int i = 0; i < 3; ++i) t[i] = clamp01d(M[3 * i] * a[0] + M[3 * i + 1] * a[1] + M[3 * i + 2] * a[2]);
        for (int k = 0; k < 3; ++k) {
            s->a[k] = a[k];
            s->b[k] = s->u[k] = t[k];
            s->c[k] = g[k] * t[k];
            s->z[k] = clamp01d(s->c[k]);
            s->y[k] = s->z[k] > 0.0 ? pow(s->z[k], e) : 0.0;
        }
        return;
    }
    for (int k = 0; k < 3; ++k) {
        s->a[k] = a[k];
        s->b[k] = g[k] * a[k];
        s->u[k] = (g_chain_plant & 4) ? s->b[k] : clamp01d(s->b[k]);
    }
    for (int i = 0; i < 3; ++i) {
        const double m0 = tr ? M[i] : M[3 * i], m1 = tr ? M[3 + i] : M[3 * i + 1], m2 = tr ? M[6 + i] : M[3 * i + 2];
        s->c[i] = m0 * s->u[0] + m1 * s->u[1] + m2 * s->u[2];
        s->z[i] = clamp01d(s->c[i]);
        s->y[i] = s->z[i] > 0.0 ? pow(s->z[i], e) : 0.0;
    }
}

/* Forward over n frames of raw values (float, [n][h][w], in units of
 * max_value).  out (may be NULL): [n][h][w][3] FP64.  With target (float,
 * [n][h][w][3]) the MSE goes to *loss. */
int orc_chain_fwd64(const float* raw, int n, int w, int h, int pattern, double max_value, const double* gains,
                    const double* ccm, double inv_gamma, double* out, const float* target, double* loss) {
    csr_t D;
    if (csr_build(w, h, pattern, &D)) return -1;
    const long px = (long)w * h;
    double acc = 0.0;
    for (int f = 0; f < n; ++f) {
        const float* x = raw + (size_t)f * px;
        for (long p = 0; p < px; ++p) {
            double a[3];
            for (int c = 0; c < 3; ++c) {
                const long r = p * 3 + c;
                double s = 0.0;
                for (long k = D.rowptr[r]; k < D.rowptr[r + 1]; ++k) s += D.val[k] * ((double)x[D.col[k]] / max_value);
                a[c] = s;
            }
            chain_px st;
            chain_pixel(a, gains, ccm, inv_gamma, &st);
            const size_t o = ((size_t)f * px + p) * 3;
            for (int c = 0; c < 3; ++c) {
                if (out) out[o + c] = st.y[c];
                if (target) {
                    const double d = st.y[c] - (double)target[o + c];
                    acc += d * d;
                }
            }
        }
    }
    if (loss) *loss = target ? acc / ((double)n * px * 3.0) : 0.0;
    csr_free(&D);
    return 0;
}

/* Vector-Jacobian product.  The output cotangent is grad_out ([n][h][w][3])
 * or, when grad_out is NULL, the MSE's 2 (y - t) / N against target.
 * grad_raw ([n][h][w], may be NULL) is with respect to raw (so includes the
 * 1 / max_value); grad_gains[3] and grad_ccm[9] sum over every frame. */
int orc_chain_vjp64(const float* raw, int n, int w, int h, int pattern, double max_value, const double* gains,
                    const double* ccm, double inv_gamma, const float* grad_out, const float* target,
                    double* grad_raw, double* grad_gains, double* grad_ccm) {
    csr_t D;
    if (csr_build(w, h, pattern, &D)) return -1;
    const long px = (long)w * h;
    const double inv_n = 1.0 / ((double)n * px * 3.0);
    double* ga = (double*)malloc(sizeof(double) * (size_t)px * 3);
    if (!ga) {
        csr_free(&D);
        return -1;
    }
    for (int k = 0; k < 3; ++k) grad_gains[k] = 0.0;
    for (int k = 0; k < 9; ++k) grad_ccm[k] = 0.0;
    for (int f = 0; f < n; ++f) {
        const float* x = raw + (size_t)f * px;
        for (long p = 0; p < px; ++p) {
            double a[3];
            for (int c = 0; c < 3; ++c) {
                const long r = p * 3 + c;
                double s = 0.0;
                for (long k = D.rowptr[r]; k < D.rowptr[r + 1]; ++k) s += D.val[k] * ((double)x[D.col[k]] / max_value);
                a[c] = s;
            }
            chain_px st;
            chain_pixel(a, gains, ccm, inv_gamma, &st);
            const size_t o = ((size_t)f * px + p) * 3;
            double gy[3], gc[3], gu[3];
            for (int c = 0; c < 3; ++c)
                gy[c] = grad_out ? (double)grad_out[o + c] : 2.0 * (st.y[c] - (double)target[o + c]) * inv_n;
            for (int i = 0; i < 3; ++i) {
                const double dz = st.z[i] > 0.0 ? inv_gamma * pow(st.z[i], inv_gamma - 1.0) : 0.0;
                const double pass = (st.c[i] >= 0.0 && st.c[i] <= 1.0) ? 1.0 : 0.0;
                gc[i] = gy[i] * dz * pass;
            }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) grad_ccm[3 * i + j] += gc[i] * st.u[j];
            for (int j = 0; j < 3; ++j) gu[j] = ccm[j] * gc[0] + ccm[3 + j] * gc[1] + ccm[6 + j] * gc[2];
            for (int k = 0; k < 3; ++k) {
                const double pass = (st.b[k] >= 0.0 && st.b[k] <= 1.0) ? 1.0 : 0.0;
                const double gb = gu[k] * pass;
                grad_gains[k] += gb * st.a[k];
                ga[p * 3 + k] = gb * gains[k];
            }
        }
        if (grad_raw) {
            /* transposed assembled operator: walk the rows of D */
            double* gx = grad_raw + (size_t)f * px;
            for (long q = 0; q < px; ++q) gx[q] = 0.0;
            for (long r = 0; r < px * 3; ++r)
                for (long k = D.rowptr[r]; k < D.rowptr[r + 1]; ++k) gx[D.col[k]] += D.val[k] * ga[r];
            for (long q = 0; q < px; ++q) gx[q] /= max_value;
        }
    }
    free(ga);
    csr_free(&D);
    return 0;
}

/* D x and D^T v of the assembled operator alone, FP64 (one frame). */
int orc_bilinear_apply64(const double* x, int w, int h, int pattern, double* out) {
    csr_t D;
    if (csr_build(w, h, pattern, &D)) return -1;
    for (long r = 0; r < (long)w * h * 3; ++r) {
        double s = 0.0;
        for (long k = D.rowptr[r]; k < D.rowptr[r + 1]; ++k) s += D.val[k] * x[D.col[k]];
        out[r] = s;
    }
    csr_free(&D);
    return 0;
}

int orc_bilinear_apply_t64(const double* v, int w, int h, int pattern, double* out) {
    csr_t D;
    if (csr_build(w, h, pattern, &D)) return -1;
    for (long q = 0; q < (long)w * h; ++q) out[q] = 0.0;
    for (long r = 0; r < (long)w * h * 3; ++r)
        for (long k = D.rowptr[r]; k < D.rowptr[r + 1]; ++k) out[D.col[k]] += D.val[k] * v[r];
    csr_free(&D);
    return 0;
}

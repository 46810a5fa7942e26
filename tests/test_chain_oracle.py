"""Pins the FP64 oracle of the differentiable ISP chain (oracle/chain_oracle.c).

* D, the bilinear demosaic operator, equals the reference's HQLI ("bilinear",
  hqli.cpp:7-44 with fill_rgb_borders, common.hpp:47-96) wherever the
  integer averages are exact (raw values that are multiples of 4), border
  rows, columns and zero corners included: against the compiled reference
  when it is built, and against the golden-pinned restatement.
* The assembled operator's transpose satisfies the adjoint identity and has
  the closed-form column sums of a bilinear stencil.
* The VJP matches central finite differences of the FP64 forward for every
  parameter and for raw pixels in the interior, next to each border and in
  the corners.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O

SIZES = [(10, 8), (13, 11), (37, 23), (64, 48)]


def _bayer_mult4(w, h, seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 64, size=(h, w)) * 4).astype(np.uint8)


@pytest.mark.parametrize("w,h", SIZES)
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
def test_bilinear_operator_equals_reference_hqli(w, h, pattern):
    raw = _bayer_mult4(w, h, w * 100 + h + pattern)
    d = O.bilinear_apply64(raw.astype(np.float64), pattern)
    e, hq = O.demosaic(raw, w, h, pattern, O.HQLI, 8)
    assert e == 0
    np.testing.assert_array_equal(d, hq.astype(np.float64))
    if O.reference_available():
        e, rf = O.ref_demosaic(raw, w, h, pattern, O.HQLI, 8)
        assert e == 0
        np.testing.assert_array_equal(d, rf.astype(np.float64))


@pytest.mark.parametrize("pattern", [0, 3])
def test_adjoint_identity_and_csr_transpose(pattern):
    w, h = 41, 29
    rng = np.random.default_rng(3)
    x = rng.standard_normal((h, w))
    v = rng.standard_normal((h, w, 3))
    dx = O.bilinear_apply64(x, pattern)
    dtv = O.bilinear_apply_t64(v, pattern)
    lhs, rhs = float(np.sum(dx * v)), float(np.sum(x * dtv))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))
    D = O.bilinear_csr(w, h, pattern)
    np.testing.assert_allclose(D @ x.ravel(), dx.ravel(), rtol=0, atol=1e-13)
    np.testing.assert_allclose(D.T @ v.ravel(), dtv.ravel(), rtol=0, atol=1e-13)


@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
def test_transposed_operator_column_sums(pattern):
    """Deep inside the frame a raw R (B) site feeds its own red (blue) output
    with weight 1, its 4 axial G neighbours with 1/2 and its 4 diagonal
    neighbours with 1/4 (total 4, red / blue channel only); a G site feeds
    its own green with 1 and its 4 axial neighbours with 1/4 (total 2)."""
    w, h = 24, 20
    ones = np.ones((h, w, 3))
    for c in range(3):
        v = np.zeros((h, w, 3))
        v[..., c] = 1.0
        cs = O.bilinear_apply_t64(v, pattern)
        rr = 1 if pattern in (1, 3) else 0
        rc = 1 if pattern in (1, 2) else 0
        for y in range(5, h - 5):
            for x in range(5, w - 5):
                site = 0 if ((y & 1) == rr and (x & 1) == rc) else (2 if ((y & 1) != rr and (x & 1) != rc) else 1)
                want = (4.0 if site == c else 0.0) if site != 1 else (2.0 if c == 1 else 0.0)
                assert cs[y, x] == want, (c, x, y, site)
    assert np.all(O.bilinear_apply_t64(ones, pattern) >= 0.0)


def _problem(n, w, h, seed):
    rng = np.random.default_rng(seed)
    # raw on a 2^-20 grid in [0, 1] so raw +- 2^-20 is exact in float32
    raw = (rng.integers(0, (1 << 20) + 1, size=(n, h, w)) / float(1 << 20)).astype(np.float32)
    gains = rng.uniform(1.0, 2.5, 3)
    ccm = np.eye(3) + rng.normal(0.0, 0.1, (3, 3))
    target = rng.uniform(0.0, 1.0, (n, h, w, 3)).astype(np.float32)
    return raw, gains, ccm, target


def test_vjp_matches_central_differences():
    n, w, h, pat, gamma = 2, 14, 12, 1, 2.2
    raw, gains, ccm, target = _problem(n, w, h, 5)

    def loss(r=raw, g=gains, m=ccm):
        return O.chain_fwd64(r, pat, 1.0, g, m, gamma, target)[1]

    gr, gg, gm = O.chain_vjp64(raw, pat, 1.0, gains, ccm, gamma, target=target)
    eps = 1e-6
    for k in range(3):
        gp, gn = gains.copy(), gains.copy()
        gp[k] += eps
        gn[k] -= eps
        fd = (loss(g=gp) - loss(g=gn)) / (2 * eps)
        assert abs(fd - gg[k]) <= 1e-6 * max(abs(fd), 1e-3), (k, fd, gg[k])
    for k in range(9):
        mp, mn = ccm.copy(), ccm.copy()
        mp.flat[k] += eps
        mn.flat[k] -= eps
        fd = (loss(m=mp) - loss(m=mn)) / (2 * eps)
        assert abs(fd - gm.flat[k]) <= 1e-6 * max(abs(fd), 1e-3), (k, fd, gm.flat[k])
    step = 2.0 ** -20
    for (f, y, x) in [(0, 5, 6), (1, 6, 7), (0, 2, 2), (0, 2, 9), (1, 9, 2), (0, 9, 11), (1, 0, 0), (0, 1, 5),
                      (1, 11, 13), (0, 3, 3), (1, 10, 4), (0, 6, 12)]:
        rp, rn = raw.copy(), raw.copy()
        rp[f, y, x] += step
        rn[f, y, x] -= step
        fd = (loss(r=rp) - loss(r=rn)) / (2 * step)
        assert abs(fd - gr[f, y, x]) <= 1e-5 * max(abs(fd), 1e-4) + 1e-9, ((f, y, x), fd, gr[f, y, x])


def test_vjp_with_cotangent_equals_mse_form():
    """The MSE mode is the cotangent 2 (y - t) / N fed to the general VJP."""
    n, w, h, pat, gamma = 1, 12, 10, 2, 2.2
    raw, gains, ccm, target = _problem(n, w, h, 9)
    out, _ = O.chain_fwd64(raw, pat, 1.0, gains, ccm, gamma)
    cot = (2.0 * (out - target) / out.size).astype(np.float32)
    a = O.chain_vjp64(raw, pat, 1.0, gains, ccm, gamma, grad_out=cot)
    b = O.chain_vjp64(raw, pat, 1.0, gains, ccm, gamma, target=target)
    for u, v in zip(a, b):
        np.testing.assert_allclose(u, v, rtol=1e-6, atol=1e-12)


# ------------------------------------------------ forward colour stages
#
# The closed forms and the integer-domain comparisons below pin the order and
# orientation of white balance (gpu_ops_kernels.cu:119-125), the colour matrix
# (gpu_ops_kernels.cu:92-96, gpu_common.cuh:43-48) and gamma
# (gpu_ops_kernels.cu:57-64).  Each planted mistake of orc_chain_set_plant
# must fail at least one of them (test_planted_mistakes_fail).

PERM = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]])  # out r <- g, g <- b, b <- r


def _frame(w, h, seed, hi=1.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, hi, size=(1, h, w)).astype(np.float32)


def _fwd(raw, pattern, gains, ccm, gamma, max_value=1.0):
    return O.chain_fwd64(raw, pattern, max_value, np.asarray(gains, float), np.asarray(ccm, float), gamma)[0]


def check_identity_parameters():
    """Gains 1, M = I, gamma 1: the chain is clamp01(D x) (x above 1 included)."""
    raw = _frame(23, 17, 1, hi=1.5)
    out = _fwd(raw, 0, [1, 1, 1], np.eye(3), 1.0)
    want = np.clip(O.bilinear_apply64(raw[0].astype(np.float64), 0), 0.0, 1.0)
    return np.array_equal(out[0], want)


def check_permutation_matrix():
    """A permutation matrix permutes the channels of clamp01(D x) exactly."""
    raw = _frame(21, 19, 2)
    a = np.clip(O.bilinear_apply64(raw[0].astype(np.float64), 1), 0.0, 1.0)
    out = _fwd(raw, 1, [1, 1, 1], PERM, 1.0)
    return np.array_equal(out[0], a[..., [1, 2, 0]])


def check_diagonal_matrix_equals_gains():
    """With D x in [0, 1] and gains >= 1, M = diag(g) with unit gains equals
    the gains with M = I."""
    raw = _frame(19, 15, 3)
    g = np.array([1.25, 1.5, 2.0])
    a = _fwd(raw, 2, g, np.eye(3), 1.0)
    b = _fwd(raw, 2, [1, 1, 1], np.diag(g), 1.0)
    return np.array_equal(a, b)


def check_gamma_two_at_quarter():
    """A constant 0.25 frame with gamma 2 gives 0.5 exactly wherever D feeds
    the pixel (the four 2 x 2 corners are 0)."""
    raw = np.full((1, 12, 14), 0.25, dtype=np.float32)
    out = _fwd(raw, 0, [1, 1, 1], np.eye(3), 2.0)[0]
    inner = out[2:-2, :]
    return np.all(inner == 0.5) and np.all(out[:2, 2:-2] == 0.5) and np.all(out[:2, :2] == 0.0)


def check_white_balance_clamps_before_matrix():
    """Gains (3, 1, 1) on a constant 0.5 frame: u = (1, .5, .5), then
    M = [[.5, 0, 0], [0, 1, 0], [0, 0, 1]] gives red 0.5, not 0.75."""
    raw = np.full((1, 12, 12), 0.5, dtype=np.float32)
    out = _fwd(raw, 0, [3, 1, 1], np.diag([0.5, 1.0, 1.0]), 1.0)[0]
    return np.all(out[2:-2, 2:-2] == [0.5, 0.5, 0.5])


def check_white_balance_before_matrix():
    """Gains (2, 1, 1) then the R <-> G swap on a constant 0.5 frame gives
    (0.5, 1, 0.5); the matrix first would give (1, 0.5, 0.5)."""
    raw = np.full((1, 12, 12), 0.5, dtype=np.float32)
    swap = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    out = _fwd(raw, 0, [2, 1, 1], swap, 1.0)[0]
    return np.all(out[2:-2, 2:-2] == [0.5, 1.0, 0.5])


def _integer_scene(w, h, seed, step=8):
    """u16 12-bit raw whose values are multiples of `step`, so the reference's
    integer HQLI averages are exact and equal D x."""
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 4096 // step, size=(h, w)) * step).astype(np.uint16)


def _hqli12(raw):
    h, w = raw.shape
    e, rgb = O.demosaic(raw, w, h, 0, O.HQLI, 12)
    assert e == 0
    return rgb


MAT_HALVES = np.array([[1.5, -0.5, 0.0], [0.0, 0.5, 0.5], [-0.5, 1.0, 0.5]])


def check_matrix_vs_reference_linear_3x3():
    """The chain's colour matrix (gains 1, gamma 1) on D x equals the
    reference's integer-domain linear_3x3 (ccm/linear_3x3.cpp:32-41, out =
    M * in, row-major) run on HQLI's output, where every product is an exact
    integer: a non-symmetric matrix with half-integer entries on even values."""
    raw = _integer_scene(26, 18, 7)
    h, w = raw.shape
    y = _fwd(raw[None].astype(np.float32), 0, [1, 1, 1], MAT_HALVES, 1.0, max_value=4095.0)[0] * 4095.0
    ok = np.max(np.abs(y - np.rint(y))) < 1e-6
    want = O.isp_ccm(_hqli12(raw), 12, MAT_HALVES.ravel())
    ok = ok and np.array_equal(np.rint(y).astype(np.int64), want.astype(np.int64))
    if O.reference_isp_available():
        ok = ok and np.array_equal(want, O.ref_ccm(_hqli12(raw), 12, MAT_HALVES.ravel()))
    return ok


def check_gains_vs_reference_manual_wb():
    """The chain's white balance (M = I, gamma 1) equals the reference's
    integer-domain manual white balance (white_balance/manual.cpp:29-31) on
    HQLI's output, with gains whose products with even values are exact."""
    raw = _integer_scene(24, 20, 8)
    g = np.array([0.5, 2.0, 1.5])
    y = _fwd(raw[None].astype(np.float32), 0, g, np.eye(3), 1.0, max_value=4095.0)[0] * 4095.0
    want = O.isp_wb(_hqli12(raw), 12, g)
    ok = np.array_equal(np.rint(y).astype(np.int64), want.astype(np.int64))
    if O.reference_isp_available():
        ok = ok and np.array_equal(want, O.ref_wb(_hqli12(raw), 12, g))
    return ok


def check_gamma_vs_reference():
    """The chain's gamma (gains 1, M = I) against the reference's
    tone::process_gamma (tone/gamma.cpp:26-29, powf(v / max, 1 / gamma)) on
    HQLI's output: within 1 step after rounding (FP64 pow vs float powf),
    equal at most pixels."""
    raw = _integer_scene(30, 22, 9)
    y = _fwd(raw[None].astype(np.float32), 3, [1, 1, 1], np.eye(3), 2.2, max_value=4095.0)[0] * 4095.0
    h, w = raw.shape
    e, rgb = O.demosaic(raw, w, h, 3, O.HQLI, 12)
    want = O.isp_gamma(rgb, 12, 2.2).astype(np.int64)
    d = np.abs(np.floor(y + 0.5).astype(np.int64) - want)
    return d.max() <= 1 and np.mean(d == 0) > 0.95


FORWARD_CHECKS = [check_identity_parameters, check_permutation_matrix, check_diagonal_matrix_equals_gains,
                  check_gamma_two_at_quarter, check_white_balance_clamps_before_matrix,
                  check_white_balance_before_matrix, check_matrix_vs_reference_linear_3x3,
                  check_gains_vs_reference_manual_wb, check_gamma_vs_reference]


@pytest.mark.parametrize("check", FORWARD_CHECKS, ids=lambda f: f.__name__)
def test_forward_colour_stages(check):
    assert check()


@pytest.mark.parametrize("mask,name", [(1, "ccm transposed"), (2, "gamma for 1/gamma"),
                                       (4, "white balance unclamped"), (8, "ccm before white balance")])
def test_planted_mistakes_fail(mask, name):
    """Each planted mistake fails at least one forward check (so the checks
    can fail), and the oracle is restored afterwards."""
    with O.chain_plant(mask):
        failed = [f.__name__ for f in FORWARD_CHECKS if not f()]
    assert failed, f"no forward check detects: {name}"
    assert all(f() for f in FORWARD_CHECKS)

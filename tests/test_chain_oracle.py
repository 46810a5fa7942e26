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

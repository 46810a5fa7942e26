"""The differentiable ISP chain on the GPU (imop_hip_isp_chain_fwd / _bwd,
imop_hip_bilinear_fwd / _vjp through the C ABI), against the FP64 oracle
(oracle/chain_oracle.c: the assembled bilinear operator and its transpose)
and against central finite differences of the FP64 forward.

Tolerances (FP32 kernels vs FP64): forward samples within 5e-5 absolute
and 99.99 % of them within 2e-6 (the gamma amplifies FP32 rounding where its
input is near 0), loss within 1e-5 relative; parameter gradients within 1e-4 relative error
(||g - g64|| / ||g64||, per parameter group) and the raw gradient within
1e-4 away from outputs whose gamma input is below 1.4e-3 (the error over
every pixel is printed and bounded by 3e-2); finite differences within 1e-4
relative.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from imop_amd import chain as C
from imop_amd import scene

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
GAMMA, MAXV = 2.2, 4095.0


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def problem(n, w, h, seed, scene_raw=False):
    rng = np.random.default_rng(seed)
    if scene_raw:
        raw = np.stack([scene.make_scene(h, w, 12, 0, seed=seed * 10 + k) for k in range(n)]).astype(np.float32)
    else:
        raw = rng.integers(0, 4096, size=(n, h, w)).astype(np.float32)
    gains = rng.uniform(1.0, 2.5, 3)
    ccm = np.eye(3) + rng.normal(0.0, 0.1, (3, 3))
    target = rng.uniform(0.0, 1.0, (n, h, w, 3)).astype(np.float32)
    return raw, gains, ccm, target


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)


@pytest.mark.parametrize("w,h", [(10, 8), (13, 11), (130, 98), (1006, 758)])
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
def test_bilinear_operator_and_transpose(w, h, pattern):
    rng = np.random.default_rng(w + h + pattern)
    x = rng.uniform(0.0, 1.0, (h, w)).astype(np.float32)
    v = rng.standard_normal((h, w, 3)).astype(np.float32)
    dx = C.bilinear(dev(x), pattern)
    dtv = C.bilinear_vjp(dev(v), pattern)
    torch.cuda.synchronize()
    want_dx = O.bilinear_apply64(x.astype(np.float64), pattern)
    want_dtv = O.bilinear_apply_t64(v.astype(np.float64), pattern)
    np.testing.assert_allclose(dx.cpu().numpy(), want_dx, rtol=0, atol=1e-6)
    assert rel(dtv.cpu().numpy(), want_dtv) <= 1e-6


@pytest.mark.parametrize("n,w,h,pattern", [(2, 130, 98, 0), (1, 37, 29, 3), (1, 1006, 758, 1)])
def test_forward_and_loss_match_fp64(n, w, h, pattern):
    raw, gains, ccm, target = problem(n, w, h, 3 + w)
    out, loss = C.chain_forward(dev(raw), dev(gains), dev(ccm.ravel()), pattern, MAXV, GAMMA, target=dev(target))
    torch.cuda.synchronize()
    want, want_loss = O.chain_fwd64(raw, pattern, MAXV, gains, ccm, GAMMA, target)
    err = np.abs(out.cpu().numpy().astype(np.float64) - want)
    # z^(1/gamma) amplifies the FP32 rounding of M u (~6e-8) by y / (gamma z)
    # where z is tiny, so a few samples near z = 0 carry more error
    assert err.max() <= 5e-5, err.max()
    assert np.mean(err <= 2e-6) >= 0.9999, np.mean(err <= 2e-6)
    assert abs(float(loss.item()) - want_loss) <= 1e-5 * want_loss


@pytest.mark.parametrize("mode", ["mse", "cotangent"])
@pytest.mark.parametrize("n,w,h,pattern", [(2, 130, 98, 0), (1, 14, 12, 2), (8, 1006, 758, 0)])
def test_gradients_match_fp64_transposed_operator(mode, n, w, h, pattern):
    raw, gains, ccm, target = problem(n, w, h, 7 + n + w, scene_raw=(w > 500))
    gout = None
    if mode == "cotangent":
        # not zero-mean: a zero-mean cotangent makes the 12 parameter sums
        # cancel to ~1/sqrt(pixels) of their terms, which measures the
        # summation's rounding rather than the VJP
        gout = (np.random.default_rng(1).uniform(-0.3, 1.0, (n, h, w, 3)) * 1e-3).astype(np.float32)
    gr, gg, gm = C.chain_backward(dev(raw), dev(gains), dev(ccm.ravel()), pattern, MAXV, GAMMA,
                                  grad_out=None if gout is None else dev(gout), target=dev(target))
    torch.cuda.synchronize()
    # the oracle sees the same (FP32-rounded) parameters as the kernels
    g32, m32 = gains.astype(np.float32).astype(np.float64), ccm.astype(np.float32).astype(np.float64)
    wr, wg, wm = O.chain_vjp64(raw, pattern, MAXV, g32, m32, GAMMA, grad_out=gout,
                               target=None if gout is not None else target)
    gr = gr.cpu().numpy()
    e_raw, e_g, e_m = rel(gr, wr), rel(gg.cpu().numpy(), wg), rel(gm.cpu().numpy(), wm)
    # Raw pixels next to an output with z = y^gamma < 1.4e-3: there the slope
    # of z^(1/gamma) multiplies FP32 rounding of M u (~1e-7) by up to 1e3, so
    # FP32 cannot match FP64 to 1e-4 on those few pixels.
    y64, _ = O.chain_fwd64(raw, pattern, MAXV, g32, m32, GAMMA)
    low = (y64 < 0.05).any(axis=-1)
    near = np.zeros_like(low)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            near |= np.roll(np.roll(low, dy, axis=1), dx, axis=2)
    e_ok = rel(gr[~near], wr[~near])
    print(f"relative error vs FP64 transposed operator: raw {e_raw:.2e} (pixels away from z < 1.4e-3: {e_ok:.2e}, "
          f"{near.mean() * 100:.3f} % near) gains {e_g:.2e} ccm {e_m:.2e}", flush=True)
    assert e_ok <= 1e-4 and e_g <= 1e-4 and e_m <= 1e-4
    assert e_raw <= 3e-2


def test_gradients_match_central_differences():
    n, w, h, pat = 2, 16, 14, 3
    rng = np.random.default_rng(21)
    raw = (rng.integers(0, (1 << 20) + 1, size=(n, h, w)) / float(1 << 20)).astype(np.float32)
    gains = rng.uniform(1.0, 2.5, 3)
    ccm = np.eye(3) + rng.normal(0.0, 0.1, (3, 3))
    target = rng.uniform(0.0, 1.0, (n, h, w, 3)).astype(np.float32)
    gr, gg, gm = C.chain_backward(dev(raw), dev(gains), dev(ccm.ravel()), pat, 1.0, GAMMA, target=dev(target))
    torch.cuda.synchronize()
    gr, gg, gm = gr.cpu().numpy(), gg.cpu().numpy(), gm.cpu().numpy()

    def loss(r=raw, g=gains, m=ccm):
        return O.chain_fwd64(r, pat, 1.0, g, m, GAMMA, target)[1]

    eps = 1e-6
    fd_g = np.array([(loss(g=gains + eps * np.eye(3)[k]) - loss(g=gains - eps * np.eye(3)[k])) / (2 * eps)
                     for k in range(3)])
    fd_m = np.zeros(9)
    for k in range(9):
        d = np.zeros(9)
        d[k] = eps
        fd_m[k] = (loss(m=ccm + d.reshape(3, 3)) - loss(m=ccm - d.reshape(3, 3))) / (2 * eps)
    assert rel(gg, fd_g) <= 1e-4, (gg, fd_g)
    assert rel(gm.ravel(), fd_m) <= 1e-4, (gm, fd_m)
    step = 2.0 ** -20
    pts = [(0, 5, 6), (1, 6, 7), (0, 2, 2), (0, 2, 9), (1, 9, 2), (0, 11, 13), (1, 0, 0), (0, 1, 5), (1, 13, 15)]
    fd_r, got_r = [], []
    for (f, y, x) in pts:
        rp, rn = raw.copy(), raw.copy()
        rp[f, y, x] += step
        rn[f, y, x] -= step
        fd_r.append((loss(r=rp) - loss(r=rn)) / (2 * step))
        got_r.append(gr[f, y, x])
    assert rel(got_r, fd_r) <= 1e-4, (got_r, fd_r)


def fd_problem(n, w, h, seed):
    """Scene frames on a pedestal (raw = 0.1 * 4095 + 0.8 * scene) and a
    colour matrix with positive entries (0.8 I + 0.1 + N(0, 0.02)), so that
    every gamma input z stays above 0.1.  Near z = 0 the slope of
    z^(1/gamma) is unbounded, and a pixel within a step of it moves a
    central difference by ~1e3 / N: on plain scene frames (zeros, and a CCM
    whose negative entries cross z = 0) the FP64 oracle's own central
    differences then miss its exact VJP by 1e-3..1e-2 at step 1e-6, while
    here they agree to ~1e-6."""
    rng = np.random.default_rng(seed)
    raw = np.stack([scene.make_scene(h, w, 12, 0, seed=seed * 10 + k) for k in range(n)]).astype(np.float32)
    raw = (409.5 + 0.8 * raw).astype(np.float32)
    gains = rng.uniform(1.0, 2.5, 3)
    ccm = 0.8 * np.eye(3) + 0.1 + rng.normal(0.0, 0.02, (3, 3))
    target = rng.uniform(0.0, 1.0, (n, h, w, 3)).astype(np.float32)
    return raw, gains, ccm, target


def test_parameter_gradients_match_central_differences_full_frame():
    """One 1006x758 12-bit scene frame (fd_problem): the 12 parameter
    gradients against central differences (step 1e-6) of the FP64 loss, to
    1e-4 relative."""
    n, w, h, pat = 1, 1006, 758, 0
    raw, gains, ccm, target = fd_problem(n, w, h, 41)
    g32, m32 = gains.astype(np.float32).astype(np.float64), ccm.astype(np.float32).astype(np.float64)
    _, gg, gm = C.chain_backward(dev(raw), dev(g32), dev(m32.ravel()), pat, MAXV, GAMMA, target=dev(target),
                                 want_raw=False)
    torch.cuda.synchronize()

    def loss(g=g32, m=m32):
        return O.chain_fwd64(raw, pat, MAXV, g, m, GAMMA, target)[1]

    eps = 1e-6
    fd = []
    for k in range(12):
        d = np.zeros(12)
        d[k] = eps
        fd.append((loss(g32 + d[:3], m32 + d[3:].reshape(3, 3)) - loss(g32 - d[:3], m32 - d[3:].reshape(3, 3)))
                  / (2 * eps))
    fd = np.array(fd)
    e_g, e_m = rel(gg.cpu().numpy(), fd[:3]), rel(gm.cpu().numpy().ravel(), fd[3:])
    print(f"1006x758 vs central differences: gains {e_g:.2e} ccm {e_m:.2e}", flush=True)
    assert e_g <= 1e-4 and e_m <= 1e-4


def test_autograd_functions():
    """ISPChain's backward equals the VJP with the incoming dL/drgb, and
    ISPChainMSE's equals the fused MSE backward."""
    ISPChain, ISPChainMSE = C.functions()
    n, w, h = 2, 64, 40
    raw, gains, ccm, target = problem(n, w, h, 31)
    r = dev(raw).requires_grad_(True)
    g = dev(gains).requires_grad_(True)
    m = dev(ccm).requires_grad_(True)
    t = dev(target)
    out = ISPChain.apply(r, g, m, 0, MAXV, GAMMA)
    loss = ((out - t) ** 2).mean()
    loss.backward()
    wr, wg, wm = C.chain_backward(dev(raw), dev(gains), dev(ccm.ravel()), 0, MAXV, GAMMA, target=t)
    assert rel(r.grad.cpu().numpy(), wr.cpu().numpy()) <= 1e-5
    assert rel(g.grad.cpu().numpy(), wg.cpu().numpy()) <= 1e-5
    assert rel(m.grad.cpu().numpy(), wm.cpu().numpy()) <= 1e-5
    r2 = dev(raw).requires_grad_(True)
    g2 = dev(gains).requires_grad_(True)
    m2 = dev(ccm).requires_grad_(True)
    l2 = ISPChainMSE.apply(r2, g2, m2, t, 0, MAXV, GAMMA)
    assert abs(float(l2.item()) - float(loss.item())) <= 1e-6 * float(loss.item())
    l2.backward()
    assert rel(r2.grad.cpu().numpy(), wr.cpu().numpy()) <= 1e-6
    assert rel(g2.grad.cpu().numpy(), wg.cpu().numpy()) <= 1e-6


def test_backward_within_three_forwards():
    """At the configuration size (8 x 1006x758) the fused backward costs at
    most 3x the fused forward + loss."""
    n, w, h = 8, 1006, 758
    raw, gains, ccm, target = problem(n, w, h, 5)
    r, g, m, t = dev(raw), dev(gains), dev(ccm.ravel()), dev(target)

    def med(fn, k=15):
        ts = []
        for _ in range(k):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        return sorted(ts)[k // 2]

    fwd = med(lambda: C.chain_forward(r, g, m, 0, MAXV, GAMMA, target=t, with_output=False))
    bwd = med(lambda: C.chain_backward(r, g, m, 0, MAXV, GAMMA, target=t))
    print(f"forward+loss {fwd * 1e3:.1f} us, backward {bwd * 1e3:.1f} us, ratio {bwd / fwd:.2f}", flush=True)
    assert bwd <= 3.0 * fwd


@pytest.fixture(scope="module")
def scene_batch():
    n, w, h = 8, 1006, 758
    raw = np.stack([scene.make_scene(h, w, 12, 0, seed=500 + k) for k in range(n)]).astype(np.float32)
    return dev(raw)


@pytest.mark.parametrize("seed", [6, 7, 8, 9, 10, 11])
def test_adam_recovers_parameters(scene_batch, seed):
    """8 x 1006x758 12-bit scene frames; target rendered with seeded gains in
    [1, 2.5] and CCM = I + N(0, 0.1); 200 Adam steps on the fused chain + MSE
    recover every parameter to < 1e-3 (these seeds played no part in choosing
    the optimiser settings)."""
    rng = np.random.default_rng(seed)
    g_t = rng.uniform(1.0, 2.5, 3)
    m_t = np.eye(3) + rng.normal(0.0, 0.1, (3, 3))
    target, _ = C.chain_forward(scene_batch, dev(g_t), dev(m_t.ravel()), 0, MAXV, GAMMA)
    g, m, loss = C.fit_gains_ccm(scene_batch, target, 0, MAXV, GAMMA, steps=200)
    err = max(float(np.abs(g.cpu().numpy() - g_t.astype(np.float32)).max()),
              float(np.abs(m.cpu().numpy() - m_t.astype(np.float32)).max()))
    print(f"seed {seed}: max |param error| {err:.2e}, final loss {float(loss):.3e}", flush=True)
    assert err < 1e-3

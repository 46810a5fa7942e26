"""imop_hip_isp_chain_bwd_noalloc (the chain backward with its partial sums
in a fixed device array instead of an allocation) against
imop_hip_isp_chain_bwd, both through the C ABI: the same kernels and the
same fixed-order reduction, so dL/draw and the 12 parameter gradients are
equal bit for bit (the bar for this FP32 path: within 1e-6 relative; equal
is asserted).  Two calls in a row, the second on a smaller input, each in
both the MSE and the cotangent form, and a call that passes every argument
check but with no cotangent or target, which both refuse."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import imop_amd

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
vp, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float


def _fns():
    L = imop_amd.lib()
    out = []
    for name in ("imop_hip_isp_chain_bwd", "imop_hip_isp_chain_bwd_noalloc"):
        fn = getattr(L, name)
        fn.argtypes = [vp, i, i, i, i, f, vp, vp, f, vp, vp, vp, vp, vp]
        fn.restype = i
        out.append(fn)
    return out


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _problem(n, w, h, seed):
    rng = np.random.default_rng(seed)
    raw = torch.from_numpy(rng.integers(0, 4096, size=(n, h, w)).astype(np.float32)).to(DEV)
    gains = torch.tensor(rng.uniform(1.0, 2.5, 3), dtype=torch.float32, device=DEV)
    ccm = torch.tensor((np.eye(3) + rng.normal(0.0, 0.1, (3, 3))).ravel(), dtype=torch.float32, device=DEV)
    target = torch.from_numpy(rng.uniform(0.0, 1.0, (n, h, w, 3)).astype(np.float32)).to(DEV)
    cot = torch.from_numpy((rng.uniform(-0.3, 1.0, (n, h, w, 3)) * 1e-3).astype(np.float32)).to(DEV)
    return raw, gains, ccm, target, cot


def _run(fn, raw, gains, ccm, target, cot, pattern):
    n, h, w = raw.shape
    gr = torch.full_like(raw, float("nan"))
    gp = torch.full((12,), float("nan"), dtype=torch.float32, device=DEV)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    e = fn(_p(raw), n, w, h, pattern, 4095.0, _p(gains), _p(ccm), 2.2, _p(cot), _p(target), _p(gr), _p(gp), st)
    torch.cuda.synchronize()
    return e, gr.cpu(), gp.cpu()


@pytest.mark.parametrize("mode", ["mse", "cotangent"])
def test_noalloc_backward_equals_allocating_backward(mode):
    bwd, bwd_noalloc = _fns()
    # first a larger input, then a smaller one on the same device array
    for k, (n, w, h, pattern) in enumerate([(3, 517, 389, 1), (1, 37, 29, 2)]):
        print(f"noalloc {mode} call {k}: {n} x {w}x{h} pattern {pattern}", flush=True)
        raw, gains, ccm, target, cot = _problem(n, w, h, 50 + k)
        if mode == "mse":
            cot = None
        e0, gr0, gp0 = _run(bwd, raw, gains, ccm, target, cot, pattern)
        e1, gr1, gp1 = _run(bwd_noalloc, raw, gains, ccm, target, cot, pattern)
        assert e0 == e1 == 0
        assert torch.isfinite(gp1).all() and torch.isfinite(gr1).all()
        assert torch.equal(gr0, gr1)
        assert torch.equal(gp0, gp1)


def test_noalloc_backward_argument_checks():
    bwd, bwd_noalloc = _fns()
    raw, gains, ccm, target, cot = _problem(1, 16, 12, 7)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    gp = torch.zeros(12, dtype=torch.float32, device=DEV)
    for fn in (bwd, bwd_noalloc):
        # neither cotangent nor target: NullInput; too small: ImageTooSmall
        assert fn(_p(raw), 1, 16, 12, 0, 4095.0, _p(gains), _p(ccm), 2.2, None, None, None, _p(gp), st) == 1
        assert fn(_p(raw), 1, 5, 12, 0, 4095.0, _p(gains), _p(ccm), 2.2, None, _p(target), None, _p(gp), st) == 5
    torch.cuda.synchronize()

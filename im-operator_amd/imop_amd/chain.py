"""The differentiable ISP chain (bilinear demosaic -> white balance -> 3x3
CCM -> gamma, FP32) on the gfx950 kernels, as torch.autograd.Functions.

The kernels are imop_hip_isp_chain_fwd / imop_hip_isp_chain_bwd of
libimop_hip.so (include/imop_hip.h); this module only binds them:

  * ISPChain.apply(raw, gains, ccm, pattern, max_value, gamma) -> rgb
      forward keeps raw and the parameters only; backward calls the
      hand-written VJP with dL/drgb
  * ISPChainMSE.apply(raw, gains, ccm, target, pattern, max_value, gamma) -> loss
      the same chain with the mean squared error fused into both kernels

raw is [n, h, w] float32 (units of max_value), gains [3], ccm [3, 3],
target [n, h, w, 3], all on the GPU.
"""
from __future__ import annotations

import ctypes

from . import lib as _lib


def _bind():
    L = _lib()
    if not getattr(L, "_chain_bound", False):
        vp, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        L.imop_hip_isp_chain_fwd.argtypes = [vp, i, i, i, i, f, vp, vp, f, vp, vp, vp, vp]
        L.imop_hip_isp_chain_fwd.restype = i
        L.imop_hip_isp_chain_bwd.argtypes = [vp, i, i, i, i, f, vp, vp, f, vp, vp, vp, vp, vp]
        L.imop_hip_isp_chain_bwd.restype = i
        L.imop_hip_bilinear_fwd.argtypes = [vp, i, i, i, vp, vp]
        L.imop_hip_bilinear_fwd.restype = i
        L.imop_hip_bilinear_vjp.argtypes = [vp, i, i, i, vp, vp]
        L.imop_hip_bilinear_vjp.restype = i
        L._chain_bound = True
    return L


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(t, name):
    import torch
    if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float32 CUDA tensor")


def chain_forward(raw, gains, ccm, pattern: int = 0, max_value: float = 4095.0, gamma: float = 2.2,
                  out=None, target=None, with_output: bool = True):
    """Forward: returns (rgb or None, loss or None); loss is a 1-element
    tensor (mean squared error against target)."""
    import torch
    n, h, w = raw.shape
    for t, nm in ((raw, "raw"), (gains, "gains"), (ccm, "ccm")):
        _check(t, nm)
    if with_output and out is None:
        out = torch.empty((n, h, w, 3), dtype=torch.float32, device=raw.device)
    if not with_output:
        out = None
    loss = None
    if target is not None:
        _check(target, "target")
        loss = torch.empty(1, dtype=torch.float32, device=raw.device)
    e = _bind().imop_hip_isp_chain_fwd(_p(raw), n, w, h, int(pattern), float(max_value), _p(gains), _p(ccm),
                                       float(gamma), _p(out), _p(target), _p(loss), _stream())
    if e != 0:
        raise RuntimeError(f"imop_hip_isp_chain_fwd returned {e}")
    return out, loss


def chain_backward(raw, gains, ccm, pattern: int = 0, max_value: float = 4095.0, gamma: float = 2.2,
                   grad_out=None, target=None, want_raw: bool = True):
    """VJP: returns (dL/draw or None, dL/dgains [3], dL/dccm [3, 3]); the
    cotangent is grad_out, or the MSE gradient against target."""
    import torch
    n, h, w = raw.shape
    if grad_out is not None:
        _check(grad_out, "grad_out")
    if target is not None:
        _check(target, "target")
    grad_raw = torch.empty_like(raw) if want_raw else None
    gp = torch.empty(12, dtype=torch.float32, device=raw.device)
    e = _bind().imop_hip_isp_chain_bwd(_p(raw), n, w, h, int(pattern), float(max_value), _p(gains), _p(ccm),
                                       float(gamma), _p(grad_out), _p(None if grad_out is not None else target),
                                       _p(grad_raw), _p(gp), _stream())
    if e != 0:
        raise RuntimeError(f"imop_hip_isp_chain_bwd returned {e}")
    return grad_raw, gp[:3], gp[3:].view(3, 3)


def bilinear(x, pattern: int = 0):
    """D x for one frame x [h, w] -> [h, w, 3] (the chain's demosaic operator)."""
    import torch
    _check(x, "x")
    h, w = x.shape
    out = torch.empty((h, w, 3), dtype=torch.float32, device=x.device)
    e = _bind().imop_hip_bilinear_fwd(_p(x), w, h, int(pattern), _p(out), _stream())
    if e != 0:
        raise RuntimeError(f"imop_hip_bilinear_fwd returned {e}")
    return out


def bilinear_vjp(v, pattern: int = 0):
    """D^T v for one frame v [h, w, 3] -> [h, w]."""
    import torch
    _check(v, "v")
    h, w, _ = v.shape
    out = torch.empty((h, w), dtype=torch.float32, device=v.device)
    e = _bind().imop_hip_bilinear_vjp(_p(v), w, h, int(pattern), _p(out), _stream())
    if e != 0:
        raise RuntimeError(f"imop_hip_bilinear_vjp returned {e}")
    return out


def _autograd():
    import torch

    class ISPChain(torch.autograd.Function):
        @staticmethod
        def forward(ctx, raw, gains, ccm, pattern=0, max_value=4095.0, gamma=2.2):
            raw, gains, ccm = raw.contiguous(), gains.contiguous(), ccm.contiguous()
            out, _ = chain_forward(raw, gains, ccm.view(9), pattern, max_value, gamma)
            ctx.save_for_backward(raw, gains, ccm)
            ctx.meta = (pattern, max_value, gamma)
            return out

        @staticmethod
        def backward(ctx, grad_out):
            raw, gains, ccm = ctx.saved_tensors
            pattern, max_value, gamma = ctx.meta
            gr, gg, gm = chain_backward(raw, gains, ccm.view(9), pattern, max_value, gamma,
                                        grad_out=grad_out.contiguous(), want_raw=ctx.needs_input_grad[0])
            return gr, gg, gm.view_as(ccm), None, None, None

    class ISPChainMSE(torch.autograd.Function):
        @staticmethod
        def forward(ctx, raw, gains, ccm, target, pattern=0, max_value=4095.0, gamma=2.2):
            raw, gains, ccm, target = raw.contiguous(), gains.contiguous(), ccm.contiguous(), target.contiguous()
            _, loss = chain_forward(raw, gains, ccm.view(9), pattern, max_value, gamma, target=target,
                                    with_output=False)
            ctx.save_for_backward(raw, gains, ccm, target)
            ctx.meta = (pattern, max_value, gamma)
            return loss[0]

        @staticmethod
        def backward(ctx, grad_loss):
            raw, gains, ccm, target = ctx.saved_tensors
            pattern, max_value, gamma = ctx.meta
            gr, gg, gm = chain_backward(raw, gains, ccm.view(9), pattern, max_value, gamma, target=target,
                                        want_raw=ctx.needs_input_grad[0])
            s = grad_loss
            return (None if gr is None else gr * s), gg * s, gm.view_as(ccm) * s, None, None, None, None

    return ISPChain, ISPChainMSE


_fns = None


def functions():
    """(ISPChain, ISPChainMSE) autograd Functions."""
    global _fns
    if _fns is None:
        _fns = _autograd()
    return _fns


def fit_gains_ccm(raw, target, pattern: int = 0, max_value: float = 4095.0, gamma: float = 2.2, steps: int = 200,
                  lr: float = 0.1, betas=(0.7, 0.9), g0: float = 1.75, final_lr_frac: float = 1e-4):
    """Recover white-balance gains and CCM from raw frames and a rendered
    target with Adam over the fused chain + MSE (ISPChainMSE).

    Parameterisation: log gains q and P = CCM . diag(gains).  In
    unsaturated pixels the output depends on P alone, so P is fitted by the
    bulk of the image while q only moves the white-balance saturation
    thresholds; the two no longer trade off against each other along the
    gains x CCM scale ambiguity, which plain (gains, CCM) descent crawls
    along.  Adam with betas (0.7, 0.9) keeps its step size tracking the
    shrinking gradients; the learning rate follows a cosine from lr to
    lr * final_lr_frac.  Start: gains g0 (middle of [1, 2.5]), CCM = I.

    Returns (gains [3], ccm [3, 3], final loss) as tensors on raw's device."""
    import math

    import torch
    _, ISPChainMSE = functions()
    dev = raw.device
    q = torch.full((3,), math.log(g0), device=dev, requires_grad=True)
    P = (torch.eye(3, device=dev) * g0).requires_grad_(True)
    opt = torch.optim.Adam([q, P], lr=lr, betas=betas)
    sched = torch.optim.lr_scheduler.LambdaLR(
        opt, lambda s: final_lr_frac + 0.5 * (1.0 - final_lr_frac) * (1.0 + math.cos(math.pi * s / steps)))
    for _ in range(steps):
        opt.zero_grad(set_to_none=True)
        g = q.exp()
        loss = ISPChainMSE.apply(raw, g, P / g[None, :], target, pattern, max_value, gamma)
        loss.backward()
        opt.step()
        sched.step()
    with torch.no_grad():
        g = q.exp()
        M = P / g[None, :]
        loss = ISPChainMSE.apply(raw, g, M, target, pattern, max_value, gamma)
    return g.detach(), M.detach(), loss.detach()

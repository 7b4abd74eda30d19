"""Differentiable FFT convolution for PyTorch over the library's forward and
adjoint kernels (ctypes calls on torch's current stream; no torch FFT).

Forward is one mufft_plan_conv round per signal, batched: the two inputs are
transformed (zero padded to N, fft.h:139-143), multiplied and inverse
transformed with 1/N normalisation (fft.c:602, 804-809).

    stereo (MUFFT_CONV_METHOD_FLAG_STEREO_MONO): x complex64 [B, N/2]
        (L = real, R = imaginary), h float32 [N/2] -> y complex64 [B, N]
    mono (MONO_MONO): x float32 [B, N/2], h float32 [N/2] -> y float32 [B, N]

Backward uses the adjoint kernels (gradient convention dL/dRe + i dL/dIm):
  G   = adjoint(output transform)(dy)
  dx  = adjoint(first transform)(G * conj(H) / N)[:, :N/2]
  dH  = sum_b G_b * conj(X_b) / N
  dh  = adjoint(second transform)(dH)[:N/2]
"""
from __future__ import annotations

import torch

from . import api as A

ZP_BOTH = A.CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST | A.CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND


class FFTConv:
    """Batched linear convolution of signals with one real filter of N/2 taps."""

    def __init__(self, N: int, stereo: bool = True):
        zp = A.FLAG_ZERO_PAD_UPPER_HALF
        self.N, self.stereo = N, stereo
        if stereo:
            self.p_first = A.Plan1D("c2c", N, A.FORWARD, zp)
            self.p_second = A.Plan1D("r2c", N, flags=zp | A.FLAG_FULL_R2C)
            self.p_out = A.Plan1D("c2c", N, A.INVERSE)
            self.nbins = N
        else:
            self.p_first = A.Plan1D("r2c", N, flags=zp)
            self.p_second = A.Plan1D("r2c", N, flags=zp)
            self.p_out = A.Plan1D("c2r", N)
            self.nbins = N // 2 + 1

    def __call__(self, x: torch.Tensor, h: torch.Tensor) -> torch.Tensor:
        return _ConvFn.apply(x, h, self)


class _ConvFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, h, op: FFTConv):
        N, B = op.N, x.shape[0]
        dev, s = x.device, torch.cuda.current_stream(x.device)
        x = x.contiguous()
        h = h.contiguous()
        X = torch.empty((B, op.nbins), dtype=torch.complex64, device=dev)
        H = torch.empty((1, op.nbins), dtype=torch.complex64, device=dev)
        op.p_first.execute_batched(X, x, B, stream=s)
        op.p_second.execute_batched(H, h, 1, stream=s)
        y = torch.empty((B, N), dtype=torch.complex64 if op.stereo else torch.float32, device=dev)
        op.p_out.execute_mul_batched(y, X, H, B, idist2=0, norm=1.0 / N, stream=s)
        ctx.save_for_backward(X, H)
        ctx.op = op
        return y

    @staticmethod
    def backward(ctx, dy):
        X, H = ctx.saved_tensors
        op = ctx.op
        N, B = op.N, X.shape[0]
        dev, s = X.device, torch.cuda.current_stream(X.device)
        dy = dy.contiguous()
        if not op.stereo:
            dy = dy.to(torch.float32)
        G = torch.empty((B, op.nbins), dtype=torch.complex64, device=dev)
        op.p_out.execute_adjoint_batched(G, dy, B, stream=s)
        dx = dh = None
        if ctx.needs_input_grad[0]:
            full = torch.empty((B, N), dtype=torch.complex64 if op.stereo else torch.float32, device=dev)
            op.p_first.execute_mul_batched(full, G, H, B, idist2=0, norm=1.0 / N, conj2=True,
                                           adjoint=True, stream=s)
            dx = full[:, : N // 2].contiguous()
        if ctx.needs_input_grad[1]:
            dH = torch.empty((1, op.nbins), dtype=torch.complex64, device=dev)
            A.mul_reduce(dH, G, X, op.nbins, B, norm=1.0 / N, conj_b=True, stream=s)
            hf = torch.empty((1, N), dtype=torch.float32, device=dev)
            op.p_second.execute_adjoint_batched(hf, dH, 1, stream=s)
            dh = hf[0, : N // 2].contiguous()
        return dx, dh, None

"""Differentiable convolution (torch.autograd over the adjoint kernels).

* Gradients at N = 256 against FP64 central finite differences of the exact
  (FP64) convolution, every coordinate of x and h.
* The full-size configuration (64 stereo signals, N = 65536, learnable real
  filter of 32768 taps initialised N(0, (1e-2)^2), MSE to the outputs of a
  hidden filter) against the FP64 closed-form gradient, and a few SGD steps
  that must lower the loss.
"""
import numpy as np
import pytest

from oracle import fp64, inputs

pytestmark = pytest.mark.gpu


def conv64(x, h, N):
    return fp64.conv_fft(x, h, N)


@pytest.mark.parametrize("stereo", [True, False])
def test_grad_vs_fp64_finite_differences(gpu, stereo):
    import torch
    from mufft_amd.autograd import FFTConv
    N, B = 256, 2
    rng = np.random.default_rng(5)
    if stereo:
        x = (rng.uniform(-1, 1, (B, N // 2)) + 1j * rng.uniform(-1, 1, (B, N // 2))).astype(np.complex64)
        w = (rng.normal(size=(B, N)) + 1j * rng.normal(size=(B, N)))
    else:
        x = rng.uniform(-1, 1, (B, N // 2)).astype(np.float32)
        w = rng.normal(size=(B, N))
    h = rng.normal(0, 0.3, N // 2).astype(np.float32)

    def loss64(xv, hv):  # L = Re sum conj(w) * y
        y = conv64(xv, hv, N)
        return float(np.real(np.sum(np.conj(w) * y)))

    op = FFTConv(N, stereo=stereo)
    tx = torch.from_numpy(x).to(gpu).requires_grad_(True)
    th = torch.from_numpy(h).to(gpu).requires_grad_(True)
    y = op(tx, th)
    tw = torch.from_numpy(w.astype(np.complex64 if stereo else np.float32)).to(gpu)
    L = torch.real(torch.sum(torch.conj(tw) * y)) if stereo else torch.sum(tw * y)
    L.backward()
    gx, gh = tx.grad.cpu().numpy(), th.grad.cpu().numpy()
    eps = 1e-4
    x64, h64 = x.astype(np.complex128 if stereo else np.float64), h.astype(np.float64)
    num_x = np.zeros_like(x64)
    for b in range(B):
        for i in range(N // 2):
            for unit in ((1.0, 1j) if stereo else (1.0,)):
                xp = x64.copy(); xp[b, i] += eps * unit
                xm = x64.copy(); xm[b, i] -= eps * unit
                d = (loss64(xp, h64) - loss64(xm, h64)) / (2 * eps)
                num_x[b, i] += d * unit  # dL/dRe + i dL/dIm
    num_h = np.zeros_like(h64)
    for i in range(N // 2):
        hp = h64.copy(); hp[i] += eps
        hm = h64.copy(); hm[i] -= eps
        num_h[i] = (loss64(x64, hp) - loss64(x64, hm)) / (2 * eps)
    assert np.max(np.abs(gx - num_x)) <= 1e-4 * np.max(np.abs(num_x)), "input gradient"
    assert np.max(np.abs(gh - num_h)) <= 1e-4 * np.max(np.abs(num_h)), "filter gradient"


def test_full_size_stereo_gradient_and_training(gpu):
    import torch
    from mufft_amd.autograd import FFTConv
    N, B, taps = 65536, 64, 32768
    x = inputs.uniform(0x5001, B * taps * 2).view(np.complex64).reshape(B, taps)
    h0 = inputs.gaussian(0x5002, taps, sigma=1e-2)
    hidden = inputs.gaussian(0x5003, taps, sigma=1e-2)
    op = FFTConv(N, stereo=True)
    tx = torch.from_numpy(x).to(gpu)
    target = op(tx, torch.from_numpy(hidden).to(gpu)).detach()
    th = torch.from_numpy(h0).to(gpu).requires_grad_(True)
    y = op(tx, th)
    loss = torch.mean(torch.abs(y - target) ** 2)
    loss.backward()
    # FP64 closed form: dL/dh = Re(sum_b IFFT_N(conj(X_b) * FFT_N(g_b)))[:taps], g = 2 (y - t) / numel
    x64 = x.astype(np.complex128)
    y64 = fp64.conv_fft(x64, h0.astype(np.float64), N)
    t64 = fp64.conv_fft(x64, hidden.astype(np.float64), N)
    g = 2.0 * (y64 - t64) / (B * N)
    G = np.fft.fft(g, axis=-1)
    X = np.fft.fft(x64, n=N, axis=-1)
    dh64 = np.real(np.fft.ifft(np.sum(np.conj(X) * G, axis=0)))[:taps]
    gh = th.grad.cpu().numpy()
    rel = np.sqrt(np.mean((gh - dh64) ** 2) / np.mean(dh64 ** 2))
    assert rel < 1e-4, rel
    # a few SGD steps on the filter reduce the loss
    opt = torch.optim.SGD([th], lr=0.5 / float(torch.max(torch.abs(th.grad))) * 1e-3)
    losses = []
    for _ in range(5):
        opt.zero_grad()
        l = torch.mean(torch.abs(op(tx, th) - target) ** 2)
        l.backward()
        opt.step()
        losses.append(float(l.detach()))
    assert losses[-1] < losses[0]

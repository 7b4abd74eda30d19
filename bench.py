"""Benchmark of the muFFT hot path on one MI355X.

One step = batched 1D FP32 transforms, inputs resident in HBM:
  c2c forward + c2c inverse, N = 4096, 16384 transforms, and
  r2c then c2r, N = 65536 real, 512 signals.
Flop convention of the reference bench (bench.c:468, 500): 5*N*log2(N) per
c2c transform and 5*N*log2(N) per r2c+c2r pair.

The JSON line carries the roofline of the dominant kernel (the fused
N = 4096 c2c kernel: algorithmic bytes 16*N per transform over its
HIP-event-timed launch duration) and a CPU baseline: the reference library
(oracle/_ref, AVX path, one plan per thread) on a bounded sample of the same
workload.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "batched c2c FFT GFLOP/s (5N·log2N) + % HBM peak; uneven conv batch speedup"

C2C_N, C2C_B = 4096, 16384
REAL_N, REAL_B = 65536, 512


def flops_per_step() -> float:
    c2c = 2 * 5.0 * C2C_N * math.log2(C2C_N) * C2C_B
    real = 5.0 * REAL_N * math.log2(REAL_N) * REAL_B
    return c2c + real


def traffic_from_profiles():
    """HBM bytes per launch of the c2c kernel from committed rocprofv3 --pmc summaries."""
    f = ROOT / "profiles" / "traffic_c2c_4096_x16384.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get("traffic_bytes_per_launch")
    except Exception:
        return None


def host_cpu_info():
    """Host CPU model, logical CPU count and the CPUs this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def baseline_threads() -> int:
    """Threads for the multi-core CPU baseline: the CPUs this process may use,
    capped by OMP_NUM_THREADS when set.  A shared GPU host may grant one
    GPU's job only a share of its cores (and set OMP_NUM_THREADS to it), so
    the baseline uses that share rather than every core of the machine."""
    n = host_cpu_info()["usable_cpus"]
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(x_c2c: np.ndarray, x_real: np.ndarray, conv_jobs, budget_s: float = 12.0,
                 one_core_s: float = 4.0):
    """Reference muFFT (AVX path) on host cores, one plan per thread, on a
    sample of the bench step; also on one core as shipped (a smaller sample);
    plus the reference's own figures for the other configurations: one c2c
    N=1024 per call on one core, the 2D images and the 640-job convolution
    round (each timed in C by bench.c's method, plans outside the loop)."""
    from oracle import ref

    if not ref.available():
        return None
    L = ref.lib()

    xc = np.ascontiguousarray(x_c2c, dtype=np.complex64)
    xr = np.ascontiguousarray(x_real, dtype=np.float32)

    def measure(threads, n_c2c, n_real, budget):
        # The timed loop runs on native threads inside the reference library's
        # process (oracle/ref_bench.c), one set of plans per thread.
        def run(nrep):
            s = L.mufft_ref_bench_run(threads, nrep, C2C_N, xc.ctypes.data, n_c2c, REAL_N, xr.ctypes.data, n_real)
            if s <= 0:
                raise RuntimeError("reference baseline run failed")
            return s

        # calibrate one repetition, then run a repetition count filling the budget
        nrep = max(1, int(budget / max(run(1), 1e-3)))
        wall = run(nrep)
        flops = nrep * (n_c2c * 2 * 5.0 * C2C_N * math.log2(C2C_N)
                        + n_real * 5.0 * REAL_N * math.log2(REAL_N))
        return flops / wall / 1e9, nrep

    def budgeted(fn, budget):
        """fn(nrep) -> seconds; one calibration call, then a run filling the budget."""
        t1 = fn(1)
        if t1 <= 0:
            raise RuntimeError("reference timing failed")
        nrep = max(1, int(budget / max(t1, 1e-6)))
        t = fn(nrep)
        if t <= 0:
            raise RuntimeError("reference timing failed")
        return t / nrep, nrep

    threads = baseline_threads()
    n_c2c = min(len(x_c2c), 64 * threads)
    n_real = min(len(x_real), 2 * threads)
    gf, nrep = measure(threads, n_c2c, n_real, budget_s)
    gf1, nrep1 = measure(1, min(len(x_c2c), 64), min(len(x_real), 2), one_core_s)

    configs = {}
    # one c2c N=1024 per call, one core (the reference's calling pattern)
    x1 = np.ascontiguousarray(x_c2c[0, :1024], dtype=np.complex64)
    per, n = budgeted(lambda r: L.mufft_ref_time_1d(0, 1024, -1, x1.ctypes.data, r * 1000), 1.0)
    configs["c2c_1024_single_call"] = {"us_per_call": round(per / 1000 * 1e6, 3), "threads": 1,
                                       "calls": n * 1000,
                                       "gflops": round(5.0 * 1024 * 10 / (per / 1000) / 1e9, 3)}
    # 2D images: pixels uniform[0,1) + N(0, 0.1^2)
    rng = np.random.default_rng(0x3001)
    img = (rng.random((16, 1024, 1024), np.float32) + 0.1 * rng.standard_normal((16, 1024, 1024), np.float32)
           + 1j * (rng.random((16, 1024, 1024), np.float32)
                   + 0.1 * rng.standard_normal((16, 1024, 1024), np.float32))).astype(np.complex64)
    per, n = budgeted(lambda r: L.mufft_ref_time_2d(0, 1024, 1024, -1, img.ctypes.data, 16, threads, r), 2.0)
    configs["2d_c2c_1024x1024x16"] = {"ms": round(per * 1e3, 3), "threads": threads, "repetitions": n,
                                      "gflops": round(16 * 5.0 * 1024 * 1024 * 20 / per / 1e9, 3)}
    del img
    rng = np.random.default_rng(0x3002)
    imr = (rng.random((4, 2048, 2048), np.float32) + 0.1 * rng.standard_normal((4, 2048, 2048), np.float32))
    imr = np.ascontiguousarray(imr, dtype=np.float32)
    t2 = min(threads, 4)
    per, n = budgeted(lambda r: L.mufft_ref_time_2d(1, 2048, 2048, -1, imr.ctypes.data, 4, t2, r), 2.0)
    fl2 = 4 * 2.5 * 2048 * 2048 * 22
    configs["2d_r2c_2048x2048x4"] = {"ms": round(per * 1e3, 3), "threads": t2, "repetitions": n,
                                     "gflops": round(fl2 / per / 1e9, 3)}
    spec = np.zeros((4, 2048, 2048), np.complex64)
    spec[..., :1025] = np.fft.rfft2(imr)[..., :1025] / 2048.0
    per, n = budgeted(lambda r: L.mufft_ref_time_2d(2, 2048, 2048, 1, spec.ctypes.data, 4, t2, r), 2.0)
    configs["2d_c2r_2048x2048x4"] = {"ms": round(per * 1e3, 3), "threads": t2, "repetitions": n,
                                     "gflops": round(fl2 / per / 1e9, 3)}
    del imr, spec
    # the 640-job convolution round
    nj = len(conv_jobs)
    NA = (ctypes.c_uint * nj)(*[j[0] for j in conv_jobs])
    MA = (ctypes.c_uint * nj)(*[j[1] for j in conv_jobs])
    FA = (ctypes.c_void_p * nj)(*[j[2].ctypes.data for j in conv_jobs])
    SA = (ctypes.c_void_p * nj)(*[j[3].ctypes.data for j in conv_jobs])
    per, n = budgeted(lambda r: L.mufft_ref_time_conv(threads, r, nj, NA, MA, FA, SA), 2.0)
    per1, n1 = budgeted(lambda r: L.mufft_ref_time_conv(1, r, nj, NA, MA, FA, SA), 2.0)
    configs["conv_round_640_jobs"] = {"ms": round(per * 1e3, 3), "threads": threads, "repetitions": n,
                                      "one_core_ms": round(per1 * 1e3, 3)}
    return {
        "value": round(gf, 3),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "reference",
        "host": host_cpu_info(),
        "sample": (f"{n_c2c} c2c N={C2C_N} forward+inverse and {n_real} r2c+c2r N={REAL_N} "
                   f"per repetition, {nrep} repetitions, {threads} native threads x one plan set each, "
                   f"muFFT built -O3 with SSE/SSE3/AVX, flags=0 (AVX selected); the sample (about "
                   f"{(n_c2c * C2C_N * 8 + n_real * REAL_N * 4) >> 20} MiB of inputs) stays in the host caches"),
        "one_core": {"value": round(gf1, 3), "unit": "GFLOP/s",
                     "sample": f"64 c2c N={C2C_N} forward+inverse and 2 r2c+c2r N={REAL_N} per "
                               f"repetition, {nrep1} repetitions, one thread (the reference as shipped)"},
        "configs": configs,
    }


def conv_jobs_host():
    """The uneven convolution workload on the host: 512 mono jobs, N =
    2^(8 + u mod 9) real, signal N/2 uniform, filter N/2 taps
    z_n*exp(-n/(N/8)), z ~ N(0,1); plus 128 stereo jobs N = 16384 (complex
    signal N/2).  Returns (N, method, signal, filter) per job."""
    from oracle import inputs

    u = inputs.splitmix64(0x4001, 512)
    sizes = [int(2 ** (8 + int(v % np.uint64(9)))) for v in u]
    jobs = []
    for i, N in enumerate(sizes):
        a = inputs.uniform(0x4002, N // 2, offset=i * 65536)
        h = (inputs.gaussian(0x4003 + i, N // 2) * np.exp(-np.arange(N // 2) / (N / 8))).astype(np.float32)
        jobs.append((N, 0 | 4 | 8, a, h))
    for i in range(128):
        N = 16384
        a = inputs.uniform(0x4004, N, offset=i * N).view(np.complex64)  # N/2 complex (L, R)
        h = (inputs.gaussian(0x4103 + i, N // 2) * np.exp(-np.arange(N // 2) / (N / 8))).astype(np.float32)
        jobs.append((N, 1 | 4 | 8, np.ascontiguousarray(a), h))
    return jobs


def conv_batch_bench(dev, stream, reps: int = 10):
    """Uneven convolution batch: 512 mono jobs, N = 2^(8 + u mod 9) real, signal
    N/2 uniform, filter N/2 taps z_n*exp(-n/(N/8)), z ~ N(0,1); plus 128 stereo
    jobs N = 16384 (complex signal N/2).  One graph launch (one persistent
    kernel: forward then inverse tasks per job group, per-XCD queues) versus
    the per-plan drop-in path (plans cached per N, 3 synchronous calls per
    job) and the batched conv API (one batched call per (N, method) group and
    step); the same kernel launched without the graph is reported too."""
    import torch

    import mufft_amd as M
    from oracle import fp64, inputs

    jobs, host = [], []
    for N, method, a, h in conv_jobs_host():
        out = torch.zeros(N, dtype=torch.complex64 if method & 1 else torch.float32, device=dev)
        jobs.append((N, method, torch.from_numpy(a).to(dev), torch.from_numpy(h).to(dev), out))
        host.append((a, h))
    torch.cuda.synchronize()
    batch = M.ConvBatch(jobs)
    ntasks, grid = batch.info()
    for _ in range(2):
        batch.execute(stream=stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        batch.execute(stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    t_batch = e0.elapsed_time(e1) / reps * 1e-3
    assert batch.status() == 0
    # the same kernel launched directly (no graph)
    e2 = torch.cuda.Event(enable_timing=True)
    e3 = torch.cuda.Event(enable_timing=True)
    e2.record(stream)
    for _ in range(reps):
        batch.execute(stream=stream, graph=False)
    e3.record(stream)
    torch.cuda.synchronize()
    t_direct = e2.elapsed_time(e3) / reps * 1e-3
    assert batch.status() == 0
    errs = []
    for k in (0, 1, 2, 511, 512, 639):
        N, method, _, _, out = jobs[k]
        a, h = host[k]
        ref = fp64.conv_fft(a.astype(np.complex128 if method & 1 else np.float64), h.astype(np.float64), N)
        errs.append(fp64.rel_rms(out.cpu().numpy(), ref))
    # per-plan drop-in: one plan per (N, method), 3 synchronous calls per job
    plans = {}
    blocks = {}
    for N, method, *_ in jobs:
        if (N, method) not in plans:
            plans[(N, method)] = M.PlanConv(N, 0, method)
            bd = plans[(N, method)].block_size // 8
            blocks[(N, method)] = (torch.zeros(bd, dtype=torch.complex64, device=dev),
                                   torch.zeros(bd, dtype=torch.complex64, device=dev))
    def dropin_round():
        for N, method, a, h, out in jobs:
            p = plans[(N, method)]
            b0, b1 = blocks[(N, method)]
            p.execute_input(0, b0, a)
            p.execute_input(1, b1, h)
            p.execute_output(out, b0, b1)
    dropin_round()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dropin_round()
    torch.cuda.synchronize()
    t_dropin = time.perf_counter() - t0
    # The batched conv API on the same jobs: inputs packed per (N, method)
    # group outside the timed region, then per group one batched call for
    # each input block and one for the output (3 launches per group).
    groups = {}
    for N, method, a, h, out in jobs:
        groups.setdefault((N, method), []).append((a, h, out))
    packed = []
    for (N, method), g in groups.items():
        p = plans[(N, method)]
        bd = p.block_size // 8
        A = torch.stack([a for a, _, _ in g])
        H = torch.stack([h for _, h, _ in g])
        O = torch.empty((len(g),) + tuple(g[0][2].shape), dtype=g[0][2].dtype, device=dev)
        B0 = torch.empty((len(g), bd), dtype=torch.complex64, device=dev)
        B1 = torch.empty_like(B0)
        packed.append((p, A, H, O, B0, B1, len(g)))

    def batched_round():
        for p, A, H, O, B0, B1, n in packed:
            p.execute_input_batched(0, B0, A, n, stream=stream)
            p.execute_input_batched(1, B1, H, n, stream=stream)
            p.execute_output_batched(O, B0, B1, n, stream=stream)

    for _ in range(2):
        batched_round()
    torch.cuda.synchronize()
    e4 = torch.cuda.Event(enable_timing=True)
    e5 = torch.cuda.Event(enable_timing=True)
    e4.record(stream)
    for _ in range(reps):
        batched_round()
    e5.record(stream)
    torch.cuda.synchronize()
    t_batched = e4.elapsed_time(e5) / reps * 1e-3
    total_bytes = sum((14 if m & 1 else 8) * N for N, m, *_ in jobs)
    return {
        "jobs": {"mono": 512, "stereo": 128, "mean_N": float(np.mean([j[0] for j in jobs]))},
        "tasks": ntasks,
        "grid": grid,
        "persistent_graph_ms": round(t_batch * 1e3, 4),
        "persistent_direct_ms": round(t_direct * 1e3, 4),
        "per_plan_dropin_ms": round(t_dropin * 1e3, 3),
        "speedup": round(t_dropin / t_batch, 2),
        "batched_api_ms": round(t_batched * 1e3, 4),
        "batched_api_launches": 3 * len(packed),
        "speedup_vs_batched_api": round(t_batched / t_batch, 2),
        "alg_bytes": total_bytes,
        "achieved_GBs": round(total_bytes / t_batch / 1e9, 1),
        "max_rel_rms_vs_fp64": max(errs),
    }


def two_d_bench(dev, stream, reps: int = 10):
    """The 2D configurations: c2c 1024 x 1024 x 16 images forward, r2c and c2r
    2048 x 2048 x 4 images; pixels uniform[0,1) + N(0, 0.1^2).  Algorithmic
    bytes count one read and one write of each image (the Nx/2+1 meaningful
    columns of the padded real layout)."""
    import torch

    import mufft_amd as M
    from oracle import fp64

    def timed(fn):
        with torch.cuda.stream(stream):
            for _ in range(2):
                fn()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn()
            e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    g = torch.Generator(device=dev)
    g.manual_seed(0x3001)
    out = {}
    n, b = 1024, 16
    x = torch.complex(torch.rand((b, n, n), generator=g, device=dev) + 0.1 * torch.randn((b, n, n), generator=g, device=dev),
                      torch.rand((b, n, n), generator=g, device=dev) + 0.1 * torch.randn((b, n, n), generator=g, device=dev))
    y = torch.empty_like(x)
    p = M.Plan2D("c2c", n, n, M.FORWARD)
    t = timed(lambda: p.execute_batched(y, x, b, stream=stream))
    alg = 16.0 * n * n * b
    err = fp64.rel_rms(y[0].cpu().numpy(), np.fft.fft2(x[0].cpu().numpy().astype(np.complex128)))
    out["c2c_1024x1024x16"] = {"us": round(t * 1e6, 1), "hbm_frac": round(alg / t / 8e12, 4),
                               "rel_rms_vs_fp64": err}
    del x, y
    n, b = 2048, 4
    g.manual_seed(0x3002)
    xr = torch.rand((b, n, n), generator=g, device=dev) + 0.1 * torch.randn((b, n, n), generator=g, device=dev)
    yc = torch.empty((b, n, n), dtype=torch.complex64, device=dev)  # padded rows of n complex
    zr = torch.empty_like(xr)
    pr = M.Plan2D("r2c", n, n, M.FORWARD)
    pc = M.Plan2D("c2r", n, n, M.INVERSE)
    alg = b * n * (4.0 * n + 8.0 * (n // 2 + 1))
    t_r = timed(lambda: pr.execute_batched(yc, xr, b, stream=stream))
    t_c = timed(lambda: pc.execute_batched(zr, yc, b, stream=stream))
    err = fp64.rel_rms(yc[0, :, : n // 2 + 1].cpu().numpy(), np.fft.rfft2(xr[0].cpu().numpy().astype(np.float64)))
    rt = fp64.rel_rms(zr[0].cpu().numpy() / (n * n), xr[0].cpu().numpy())
    out["r2c_2048x2048x4"] = {"us": round(t_r * 1e6, 1), "hbm_frac": round(alg / t_r / 8e12, 4),
                              "rel_rms_vs_fp64": err}
    out["c2r_2048x2048x4"] = {"us": round(t_c * 1e6, 1), "hbm_frac": round(alg / t_c / 8e12, 4),
                              "roundtrip_rel_rms": rt}
    return out


def host_paths_bench():
    """PCIe-inclusive rates of the host-memory paths (not part of `value`):
    one drop-in call on pinned buffers (used in place) and on pageable buffers
    (staged), and the c2c N=4096 x 16384 batch streamed from host to host."""
    import torch

    import mufft_amd as M
    out = {}
    n = 1024
    p = M.Plan1D("c2c", n, M.FORWARD)
    for mem in ("pinned", "pageable"):
        x = torch.randn(n, dtype=torch.complex64, pin_memory=(mem == "pinned"))
        y = torch.empty(n, dtype=torch.complex64, pin_memory=(mem == "pinned"))
        for _ in range(10):
            p.execute(y, x)
        t0 = time.perf_counter()
        for _ in range(200):
            p.execute(y, x)
        out[f"dropin_c2c_1024_{mem}_us"] = round((time.perf_counter() - t0) / 200 * 1e6, 1)
    q = M.Plan1D("c2c", C2C_N, M.FORWARD)
    for mem in ("pinned", "pageable"):
        x = torch.empty((C2C_B, C2C_N), dtype=torch.complex64, pin_memory=(mem == "pinned"))
        x.real.uniform_(-1, 1)
        x.imag.uniform_(-1, 1)
        y = torch.empty_like(x, pin_memory=(mem == "pinned"))
        q.execute_batched(y, x, C2C_B)
        t0 = time.perf_counter()
        for _ in range(3):
            q.execute_batched(y, x, C2C_B)
        dt = (time.perf_counter() - t0) / 3
        out[f"batched_host_c2c_4096x16384_{mem}_ms"] = round(dt * 1e3, 2)
        out[f"batched_host_c2c_4096x16384_{mem}_GFLOPs"] = round(5.0 * C2C_N * math.log2(C2C_N) * C2C_B / dt / 1e9, 1)
        del x, y
    return out


def diff_conv_bench(dev, stream, steps: int = 10, warmup: int = 3):
    """Differentiable convolution training step: 64 stereo signals, N = 65536,
    learnable real filter of 32768 taps (N(0, (1e-2)^2)), MSE to the outputs of a
    hidden filter, SGD on the filter.  One step = forward (3 transforms + fused
    multiply), loss, backward (adjoint transforms, fused multiplies, filter
    gradient reduction) and the optimiser update, on one torch stream."""
    import torch

    from mufft_amd.autograd import FFTConv
    from oracle import fp64, inputs

    N, B, taps = 65536, 64, 32768
    x = inputs.uniform(0x5001, B * taps * 2).view(np.complex64).reshape(B, taps)
    h0 = inputs.gaussian(0x5002, taps, sigma=1e-2)
    hidden = inputs.gaussian(0x5003, taps, sigma=1e-2)
    op = FFTConv(N, stereo=True)
    with torch.cuda.stream(stream):
        tx = torch.from_numpy(x).to(dev)
        target = op(tx, torch.from_numpy(hidden).to(dev)).detach()
        th = torch.from_numpy(h0).to(dev).requires_grad_(True)
        loss0 = torch.mean(torch.abs(op(tx, th) - target) ** 2)
        loss0.backward()
        gh = th.grad.detach().cpu().numpy()
        opt = torch.optim.SGD([th], lr=0.5e-3 / float(torch.max(torch.abs(th.grad))))

        def train_step():
            opt.zero_grad(set_to_none=False)
            loss = torch.mean(torch.abs(op(tx, th) - target) ** 2)
            loss.backward()
            opt.step()
            return loss

        for _ in range(warmup):
            train_step()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            last = train_step()
        e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    # FP64 closed form of the first step's filter gradient.
    x64 = x.astype(np.complex128)
    g = 2.0 * (fp64.conv_fft(x64, h0.astype(np.float64), N) -
               fp64.conv_fft(x64, hidden.astype(np.float64), N)) / (B * N)
    dh64 = np.real(np.fft.ifft(np.sum(np.conj(np.fft.fft(x64, n=N, axis=-1)) *
                                      np.fft.fft(g, axis=-1), axis=0)))[:taps]
    rel = float(np.sqrt(np.mean((gh - dh64) ** 2) / np.mean(dh64 ** 2)))
    return {
        "workload": "64 stereo signals x N=65536 (32768 complex samples), learnable 32768-tap real filter, MSE, SGD",
        "ms_per_step": round(ms, 4),
        "signals_per_s": round(B / (ms * 1e-3), 1),
        "filter_grad_rel_rms_vs_fp64": rel,
        "loss_first": float(loss0.detach()),
        "loss_last": float(last.detach()),
    }


# The per-path throughput table: (name, kind, N or image side, batch); 2D
# entries carry "2d_" kinds.  Every launch moves 256 MiB to 1 GiB.
PATHS = [
    ("c2c 4096", "c2c", 4096, 16384), ("r2c 65536", "r2c", 65536, 512), ("c2r 65536", "c2r", 65536, 512),
    ("c2c 1024", "c2c", 1024, 32768), ("c2c 8192", "c2c", 8192, 4096), ("c2c 16384", "c2c", 16384, 2048),
    ("c2c 32768", "c2c", 32768, 1024), ("c2c 65536", "c2c", 65536, 512),
    ("2D c2c 1024^2", "2d_c2c", 1024, 16), ("2D r2c 2048^2", "2d_r2c", 2048, 4), ("2D c2r 2048^2", "2d_c2r", 2048, 4),
    ("2D c2c 4096^2", "2d_c2c", 4096, 2), ("2D c2c 8192^2", "2d_c2c", 8192, 1),
    ("four-step 2^17", "c2c", 1 << 17, 512), ("four-step 2^20", "c2c", 1 << 20, 64),
    ("four-step 2^22", "c2c", 1 << 22, 16), ("four-step 2^23", "c2c", 1 << 23, 8),
    ("four-step 2^24", "c2c", 1 << 24, 2), ("four-step 2^26", "c2c", 1 << 26, 1),
    ("r2c 2^21 (four-step + resolve)", "r2c", 1 << 21, 64), ("c2r 2^21 (resolve + four-step)", "c2r", 1 << 21, 64),
]


def paths_bench(reps: int = 12):
    """Throughput of every planner path on inputs resident in HBM, with cold
    caches: a 512 MiB streaming copy (which evicts the 256 MB Infinity Cache)
    precedes each launch, and the launch alone is timed with HIP events on
    its stream; the median of `reps` launches.  Algorithmic bytes: one read
    and one write of every point (c2c 16 B, r2c / c2r 4 + 8 B per real
    sample pair as in the main line); 2D counts one read and one write of each
    image.  Prints one JSON line."""
    import torch

    import mufft_amd as M
    from mufft_amd import _lib

    L = _lib.load()
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(0x5001)
    f0 = torch.empty(64 << 20, dtype=torch.complex64, device=dev)
    f1 = torch.empty_like(f0)

    def cold_us(fn):
        ts = []
        for i in range(reps + 2):
            assert L.mufft_hip_stream_copy(f1.data_ptr(), f0.data_ptr(), f0.numel() * 8, None) == 0
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            if i >= 2:
                ts.append(e0.elapsed_time(e1) * 1e3)
        ts.sort()
        return ts[len(ts) // 2]

    rows = []
    for name, kind, n, b in PATHS:
        if kind.startswith("2d_"):
            k = kind[3:]
            if k == "r2c":
                x = torch.rand((b, n, n), generator=g, device=dev)
                y = torch.empty((b, n, n), dtype=torch.complex64, device=dev)
            elif k == "c2r":
                x = torch.randn((b, n, n), generator=g, device=dev, dtype=torch.complex64)
                y = torch.empty((b, n, n), device=dev)
            else:
                x = torch.randn((b, n, n), generator=g, device=dev, dtype=torch.complex64)
                y = torch.empty_like(x)
            p = M.Plan2D(k, n, n, M.INVERSE if k == "c2r" else M.FORWARD)
            alg = b * n * (16.0 * n if k == "c2c" else 4.0 * n + 8.0 * (n // 2 + 1))
        else:
            if kind == "r2c":
                x = torch.rand((b, n), generator=g, device=dev)
                y = torch.empty((b, n // 2 + 1), dtype=torch.complex64, device=dev)
            elif kind == "c2r":
                x = torch.randn((b, n // 2 + 1), generator=g, device=dev, dtype=torch.complex64)
                y = torch.empty((b, n), device=dev)
            else:
                x = torch.randn((b, n), generator=g, device=dev, dtype=torch.complex64)
                y = torch.empty_like(x)
            p = M.Plan1D(kind, n, M.INVERSE if kind == "c2r" else M.FORWARD)
            alg = b * (16.0 * n if kind == "c2c" else 4.0 * n + 8.0 * (n // 2 + 1))
        us = cold_us(lambda: p.execute_batched(y, x, b))
        rows.append({"path": name, "batch": b, "us": round(us, 1), "TBs": round(alg / us / 1e6, 2),
                     "hbm_frac": round(alg / us / 1e6 / (HBM_PEAK_GBS / 1e3), 3)})
        print(f"{name} x {b}: {us:.1f} us, {alg / us / 1e6:.2f} TB/s", file=sys.stderr, flush=True)
        p.close()
        del x, y
    print(json.dumps({"paths": rows, "method": "cold caches (512 MiB copy before each launch), median of "
                      f"{reps} HIP-event-timed launches"}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--paths", action="store_true",
                    help="time every planner path with cold caches (one JSON line) instead of the step")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the batch of independent "
              "transforms would shard across GPUs with no collective.", file=sys.stderr)
        sys.exit(2)
    if args.paths:
        paths_bench()
        return

    import torch

    import mufft_amd as M
    from oracle import fp64, inputs

    dev = torch.device("cuda:0")
    t_gen = time.time()
    x_c2c = inputs.complex_uniform(0x2001, C2C_N * C2C_B).reshape(C2C_B, C2C_N)
    x_real = inputs.uniform(0x2002, REAL_N * REAL_B).reshape(REAL_B, REAL_N)
    gen_s = time.time() - t_gen

    dx = torch.from_numpy(x_c2c).to(dev)
    dX = torch.empty_like(dx)
    dy = torch.empty_like(dx)
    dr = torch.from_numpy(x_real).to(dev)
    dS = torch.empty((REAL_B, REAL_N // 2 + 1), dtype=torch.complex64, device=dev)
    dz = torch.empty_like(dr)

    # One plain 1 GiB device copy outside the timed region: the reference
    # kernel of known traffic that calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE.
    dcal = torch.empty_like(dx)
    dcal.copy_(dx)
    torch.cuda.synchronize()
    # Achievable streaming bandwidth: the library's copy kernel (the FFT kernels'
    # 8-byte accesses and non-temporal policy) on the same 512 MiB.
    from mufft_amd import _lib
    L = _lib.load()
    nbytes = dx.numel() * 8
    for _ in range(3):
        L.mufft_hip_stream_copy(dcal.data_ptr(), dx.data_ptr(), nbytes, None)
    c0 = torch.cuda.Event(enable_timing=True)
    c1 = torch.cuda.Event(enable_timing=True)
    c0.record()
    for _ in range(10):
        assert L.mufft_hip_stream_copy(dcal.data_ptr(), dx.data_ptr(), nbytes, None) == 0
    c1.record()
    torch.cuda.synchronize()
    copy_gbs = 2.0 * nbytes / (c0.elapsed_time(c1) / 10 * 1e-3) / 1e9
    del dcal

    # Plan-time autotuning (outside the timed region, like the reference's
    # plan creation, bench.c:214-221): every fused-kernel variant of the four
    # transforms is timed at the workload's batch and checked against FP64;
    # the fastest valid one wins.
    from mufft_amd import api as A
    tuned = {}
    for n, kind in ((C2C_N, "c2c_forward"), (C2C_N, "c2c_inverse"), (REAL_N, "r2c"), (REAL_N, "c2r")):
        var, us = A.autotune(n, kind, C2C_B if kind.startswith("c2c") else REAL_B)
        tuned[f"{kind}_{n}"] = {"variant": hex(var), "us_tuning_batch": round(us, 1)}

    stream = torch.cuda.Stream(dev)  # default priority

    # The c2c plans: the tuner's pick replaces the static default only if it
    # is faster by more than 2% over interleaved rounds of whole launches
    # (the c2c variants near the top differ by a few percent, so a single
    # tuning pass can pick a noise winner).  Plans with explicit variants.
    def pick_c2c(direction, kind):
        tv = int(tuned[f"{kind}_{C2C_N}"]["variant"], 16)
        dv = L.mufft_hip_default_variant(C2C_N, 0 if direction == M.FORWARD else 1)
        if tv == dv:
            return dv, {}
        plans = {v: M.Plan1D("c2c", C2C_N, direction, variant=v) for v in (dv, tv)}
        t = {v: [] for v in plans}
        for _ in range(3):
            for v, p in plans.items():
                p.execute_batched(dX, dx, C2C_B, stream=stream)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(4):
                    p.execute_batched(dX, dx, C2C_B, stream=stream)
                torch.cuda.synchronize()
                t[v].append((time.perf_counter() - t0) * 1e3 / 4)
        for p in plans.values():
            p.close()
        best = tv if min(t[tv]) < 0.98 * min(t[dv]) else dv
        return best, {f"{v:#x}": round(min(x), 4) for v, x in t.items()}

    fwd_var, fwd_ms = pick_c2c(M.FORWARD, "c2c_forward")
    inv_var, inv_ms = pick_c2c(M.INVERSE, "c2c_inverse")
    c2c_choice = {"forward": hex(fwd_var), "inverse": hex(inv_var), "launch_ms": {"forward": fwd_ms, "inverse": inv_ms}}
    fwd = M.Plan1D("c2c", C2C_N, M.FORWARD, variant=fwd_var)
    inv = M.Plan1D("c2c", C2C_N, M.INVERSE, variant=inv_var)

    # The real legs run right after two 1 GiB c2c launches, which per-transform
    # tuning does not see (on MI355X the tuned two-workgroup variants lost ~9%
    # of the step to the one-workgroup defaults there).  So the (r2c, c2r) pair
    # is chosen by timing the whole step for a few candidates of each
    # transform, still at plan time.  Each candidate pair gets plans of its
    # own (per-plan variants, nothing process-wide changes); the pairs are
    # timed in five interleaved rounds and each keeps its fastest round; the
    # static default pair stays unless another beats it by more than the
    # default's own noise, the gap between its median and fastest rounds (at
    # least 1%; the median ignores an outlier round).
    def make_pair(rv, cv):
        return M.Plan1D("r2c", REAL_N, variant=rv), M.Plan1D("c2r", REAL_N, variant=cv)

    def pair_round_ms(pr, pc, reps=6):
        def one():
            fwd.execute_batched(dX, dx, C2C_B, stream=stream)
            inv.execute_batched(dy, dX, C2C_B, stream=stream)
            pr.execute_batched(dS, dr, REAL_B, stream=stream)
            pc.execute_batched(dz, dS, REAL_B, stream=stream)

        one()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            one()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / reps

    # candidates: the tuned variant, the static default and one workgroup of
    # 32 points per thread (V32ST; for c2r with chunked loads), which won the
    # step on some MI355X machines
    one_wg_r, one_wg_c = 0x1120, 0x10001120
    def_r, def_c = L.mufft_hip_default_variant(REAL_N, 2), L.mufft_hip_default_variant(REAL_N, 3)
    cand_r = sorted({int(tuned[f"r2c_{REAL_N}"]["variant"], 16), def_r, one_wg_r})
    cand_c = sorted({int(tuned[f"c2r_{REAL_N}"]["variant"], 16), def_c, one_wg_c})
    pairs = {(rv, cv): make_pair(rv, cv) for rv in cand_r for cv in cand_c}
    rounds = {k: [] for k in pairs}
    for _ in range(5):
        for k, (pr, pc) in pairs.items():
            rounds[k].append(pair_round_ms(pr, pc))
    pair_ms = {k: min(v) for k, v in rounds.items()}
    dflt = rounds[(def_r, def_c)]
    noise = max(0.01, (float(np.median(dflt)) - min(dflt)) / min(dflt))
    best = min(pair_ms, key=pair_ms.get)
    if pair_ms[best] > pair_ms[(def_r, def_c)] * (1.0 - noise):
        best = (def_r, def_c)
    best_r, best_c = best
    real_pair = {"r2c": hex(best_r), "c2r": hex(best_c), "noise": round(noise, 4),
                 "step_ms": {f"{r:#x}/{c:#x}": round(v, 4) for (r, c), v in pair_ms.items()},
                 "default_rounds_ms": [round(v, 4) for v in dflt]}
    r2c, c2r = pairs.pop(best)
    for pr, pc in pairs.values():
        pr.close()
        pc.close()
    # Confirmation of the chosen pair over 50 whole steps (about 25 ms,
    # reported).  It also leaves the GPU at its sustained clocks: after the
    # short selection rounds, with a synchronisation every few steps, the
    # first ~40 steps of a continuous run were up to 5% slower on MI355X
    # (0.54 -> 0.51 ms per step, per-step events).
    real_pair["confirm_ms"] = round(pair_round_ms(r2c, c2r, reps=50), 4)

    ev = []

    def step(record=False):
        if record:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
        fwd.execute_batched(dX, dx, C2C_B, stream=stream)
        if record:
            e1.record(stream)
            ev.append((e0, e1))
        inv.execute_batched(dy, dX, C2C_B, stream=stream)
        r2c.execute_batched(dS, dr, REAL_B, stream=stream)
        c2r.execute_batched(dz, dS, REAL_B, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # The forward c2c launch is bracketed by HIP events on about ten of the
    # timed steps (every `ev_every`-th): on MI355X each event pair costs
    # about 13 us of step time (0.499 -> 0.513 ms with a pair on every step).
    ev_every = max(1, args.steps // 10)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(record=(i % ev_every == 0))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    value = flops_per_step() / (wall / args.steps) / 1e9

    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    alg_bytes = 16.0 * C2C_N * C2C_B
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9

    # accuracy on a sample of the measured outputs
    idx = [0, 1, C2C_B // 2, C2C_B - 1]
    Xs = dX[idx].cpu().numpy()
    err_fwd = fp64.rel_rms(Xs, fp64.c2c(x_c2c[idx], -1))
    err_rt = fp64.rel_rms(dy[idx].cpu().numpy() / C2C_N, x_c2c[idx])
    ridx = [0, REAL_B - 1]
    err_r2c = fp64.rel_rms(dS[ridx].cpu().numpy(), fp64.r2c(x_real[ridx]))
    err_rrt = fp64.rel_rms(dz[ridx].cpu().numpy() / REAL_N, x_real[ridx])

    cpu = None if args.no_cpu else cpu_baseline(x_c2c, x_real, conv_jobs_host())
    del dx, dX, dy, dr, dS, dz
    torch.cuda.empty_cache()
    conv = conv_batch_bench(dev, stream)
    dconv = diff_conv_bench(dev, stream)
    host = host_paths_bench()
    twod = two_d_bench(dev, stream)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (SplitMix64 uniform[-1,1), seeds 0x2001/0x2002)",
        "config": {
            "workload": "batched 1D FP32: c2c forward+inverse N=4096 x 16384, r2c then c2r N=65536 x 512",
            "c2c": {"N": C2C_N, "batch": C2C_B, "directions": ["forward", "inverse"]},
            "real": {"N": REAL_N, "batch": REAL_B, "sequence": ["r2c", "c2r"]},
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "fft_rows_kernel<4096, forward, c2c>",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(),
            "kernel_ms": round(kern_ms, 5),
            "event_samples": len(ev),
            "alg_bytes_per_launch": int(alg_bytes),
            "stream_copy_GBs": round(copy_gbs, 1),
            "frac_of_stream_copy": round(achieved / copy_gbs, 4),
        },
        "cpu_baseline": cpu,
        "accuracy": {
            "c2c_fwd_rel_rms_vs_fp64": err_fwd,
            "c2c_roundtrip_rel_rms": err_rt,
            "r2c_rel_rms_vs_fp64": err_r2c,
            "real_roundtrip_rel_rms": err_rrt,
        },
        "tuned_variants": tuned,
        "c2c_variants": c2c_choice,
        "real_pair_by_step": real_pair,
        "conv_batch_speedup": conv["speedup"],
        "conv_batch": conv,
        "diff_conv": dconv,
        "host_paths": host,
        "two_d": twod,
        "input_gen_s": round(gen_s, 2),
    }
    if cpu:
        # the reference's own figures beside the GPU's for every configuration
        cc = cpu["configs"]
        out["gpu_vs_cpu"] = {
            "c2c_1024_single_call_us": {
                "gpu_dropin_pinned": host["dropin_c2c_1024_pinned_us"],
                "gpu_dropin_pageable": host["dropin_c2c_1024_pageable_us"],
                "cpu_one_core": cc["c2c_1024_single_call"]["us_per_call"]},
            "2d_c2c_1024x1024x16_ms": {"gpu": round(twod["c2c_1024x1024x16"]["us"] / 1e3, 4),
                                       "cpu": cc["2d_c2c_1024x1024x16"]["ms"]},
            "2d_r2c_2048x2048x4_ms": {"gpu": round(twod["r2c_2048x2048x4"]["us"] / 1e3, 4),
                                      "cpu": cc["2d_r2c_2048x2048x4"]["ms"]},
            "2d_c2r_2048x2048x4_ms": {"gpu": round(twod["c2r_2048x2048x4"]["us"] / 1e3, 4),
                                      "cpu": cc["2d_c2r_2048x2048x4"]["ms"]},
            "conv_round_640_jobs_ms": {"gpu_batch_graph": conv["persistent_graph_ms"],
                                       "gpu_batched_api": conv["batched_api_ms"],
                                       "cpu": cc["conv_round_640_jobs"]["ms"],
                                       "cpu_one_core": cc["conv_round_640_jobs"]["one_core_ms"]},
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Test infrastructure only: the CPU oracle for the muFFT hot path.

Nothing in the shipped library imports this package. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker or the timed CPU baseline, never as the product path.

Contents
--------
restate   float32 restatement of muFFT's C path: planner (fft.c:217-435),
          twiddles (fft.c:101-134, 441-455), executor (fft.c:799-989) and
          every scalar kernel (kernel.c:22-738).  Pinned bit-exactly against
          the compiled reference (``oracle/_ref``, flags = CPU_NO_SIMD) and
          against the FP64 oracle within tolerance.
fp64      FP64 DFT/FFT and direct O(N^2) convolution (test.c:375-403 in
          double).  np.fft is pinned by brute-force DFT and closed forms.
inputs    SplitMix64 seeded input generators (uniform, Gaussian).
ref       ctypes binding to the compiled reference library in oracle/_ref.
"""

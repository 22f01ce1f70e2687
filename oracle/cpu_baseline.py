"""CPU baseline: the reference's own op sequence for cu_get_reconst_xyprojection, on host cores.

Test / benchmark infrastructure only (the CPU-baseline leg of bench.py). The reference has no
CPU path (every function is typed on CuArray), so this restates holofunc.jl:254-263 on
Array inputs with the same arithmetic: complex64 FFTs (torch.fft on CPU, multithreaded),
fftshift / ifftshift, the ComplexF32 running product ``F .*= Hdz`` and
``min(proj, abs2(ifft(ifftshift(F))))``; transfer functions in the reference's arithmetic
(Float32 d and sqrt, FP64 exp -> ComplexF32, holofunc.jl:8-15, 38-45).
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import asm


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo, as lscpu prints it)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:  # pragma: no cover
        pass
    import platform

    return platform.processor() or "unknown"


def cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def reconst_xyprojection_cpu(holo: np.ndarray, h0: np.ndarray, hdz: np.ndarray, slices: int):
    """holofunc.jl:80-82 + 254-263 with torch CPU complex64."""
    import torch

    psi = torch.from_numpy(np.sqrt(holo.astype(np.float32))).to(torch.complex64)
    H0 = torch.from_numpy(h0.astype(np.complex64))
    Hdz = torch.from_numpy(hdz.astype(np.complex64))
    F = torch.fft.fftshift(torch.fft.fft2(psi)) * H0
    x = torch.fft.ifft2(torch.fft.ifftshift(F))
    proj = (x * x.conj()).abs().to(torch.float32)
    for _ in range(1, slices):
        F *= Hdz
        x = torch.fft.ifft2(torch.fft.ifftshift(F))
        proj = torch.fmin(proj, (x * x.conj()).abs().to(torch.float32))
    return proj.numpy()


def measure(holo: np.ndarray, z0: float, dz: float, wavlen: float, dx: float, slices: int,
            budget_s: float = 15.0, threads: int | None = None) -> dict:
    """Times the CPU restatement on ``threads`` host threads (default: every core in the
    affinity mask) over a bounded sample of about ``budget_s`` seconds of work: one run covers
    the whole workload, or as many slices as fit a third of the budget, and runs repeat (3 to
    25) until the budget is spent; returns slices/s over the median run."""
    import torch

    n_threads = threads or cores()
    torch.set_num_threads(n_threads)
    ny, nx = holo.shape
    d32 = asm.transfer_sqrt_arr(nx, wavlen, dx, ny)
    h0 = asm.transfer_reference(z0, wavlen, d32)
    hdz = asm.transfer_reference(dz, wavlen, d32)
    reconst_xyprojection_cpu(holo, h0, hdz, 2)  # warm-up (first-call costs)
    calib = min(slices, 16)
    t = time.perf_counter()
    reconst_xyprojection_cpu(holo, h0, hdz, calib)
    per_slice = max((time.perf_counter() - t) / calib, 1e-5)
    sample = int(max(2, min(slices, (budget_s / 3.0) / per_slice)))
    times = []
    runs = 3
    while len(times) < runs:
        t = time.perf_counter()
        reconst_xyprojection_cpu(holo, h0, hdz, sample)
        times.append(time.perf_counter() - t)
        if len(times) == 1:  # the first full run sizes the repetition count
            runs = int(min(25, max(3, round(budget_s / times[0]))))
    med = float(np.median(times))
    return {"value": sample / med, "unit": "z-slices/s", "cores": n_threads, "kind": "port", "model": cpu_model(),
            "sample": f"{nx}x{ny} hologram, {sample} of {slices} slices, median of {runs} runs "
                      f"({sum(times):.1f} s of CPU work), torch {torch.__version__} CPU complex64 FFT "
                      f"on {n_threads} threads"}

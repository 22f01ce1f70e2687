"""How far the reference's own arithmetic lands from FP64 at the configured depths (test
infrastructure, CPU only; prints a markdown table).

The reference builds d in FP64 and stores it as Float32, takes the Float32 sqrt before its FP64
exp (holofunc.jl:12, 42), rounds H to ComplexF32 and advances the spectrum by the ComplexF32
running product F .*= Hdz (holofunc.jl:260). ``asm`` restates that as its 'reference' mode; this
script compares it with the FP64 'exact' evaluation on the seeded inputs:

  512^2, seed-101 200-disk hologram, 100 slices from z0 = 1e4 um at dz = 1e3 um
  1024^2, seed-102 1000-disk hologram, 1000 slices from 1e4 um at 200 um
  1024^2, seed-103 500-disk two-plane pair 2000 um apart, GS 100 iterations (transfer arrays
      in each arithmetic), then 512 slices from 1e4 um at 200 um

Run: python3 -m oracle.emulation_deviation [--quick]   (--quick: the 512^2 case only)
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from . import asm

LAM, DX = 0.6328, 10.0


def _workers():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:  # pragma: no cover
        return 1


def reference_projection(wavefront, slices, z_front, z_dz, workers=None):
    """asm's 'reference' mode (holofunc.jl:254-263 in the reference's arithmetic) with threaded
    FP64 inverse FFTs of the ComplexF32 running spectrum."""
    import scipy.fft as sfft

    w = workers or _workers()
    ny, nx = wavefront.shape
    d32 = asm.transfer_sqrt_arr(nx, LAM, DX, ny)
    h0 = asm.transfer_reference(z_front, LAM, d32).astype(np.complex64)
    hdz = asm.transfer_reference(z_dz, LAM, d32).astype(np.complex64)
    spec = np.fft.fftshift(sfft.fft2(wavefront.astype(np.complex128), workers=w)).astype(np.complex64)
    spec = (spec * h0).astype(np.complex64)
    proj = None
    for k in range(slices):
        if k:
            spec = (spec * hdz).astype(np.complex64)
        psi = sfft.ifft2(np.fft.ifftshift(spec).astype(np.complex128), workers=w)
        inten = psi.real ** 2 + psi.imag ** 2
        proj = inten if proj is None else np.fmin(proj, inten, out=proj)
    return proj.astype(np.float32)


def transfer_phase_error(n, z):
    """Max and RMS phase error (rad) of the reference's H(z) against FP64, propagating components."""
    d32 = asm.transfer_sqrt_arr(n, LAM, DX)
    d64 = asm.transfer_sqrt_arr(n, LAM, DX, dtype=np.float64)
    ph32 = (2 * np.pi * z / LAM) * np.sqrt(d32.astype(np.float32)).astype(np.float64)
    ph64 = (2 * np.pi * z / LAM) * np.sqrt(d64)
    e = np.angle(np.exp(1j * (ph32 - ph64)))
    return float(np.abs(e).max()), float(np.sqrt((e ** 2).mean()))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "particleholography.jl_amd"))
    from particleholography import synth

    rows = []
    cases = [("512^2 x 100, dz 1e3", 101, 200, 512, 100, -1e4, -1e3, 1e5)]
    if not args.quick:
        cases.append(("1024^2 x 1000, dz 200", 102, 1000, 1024, 1000, -1e4, -200.0, 2e5))
    for name, seed, nd, n, slices, z0, dz, zmax in cases:
        holo = synth.hologram(seed, nd, n, zmin=1e4, zmax=zmax, wavlen=LAM, dx=DX)
        wf = asm.gabor_wavefront(holo)
        exact = asm.reconst_xyprojection_running(wf, slices, z_front=z0, z_dz=dz, wavlen=LAM, dx=DX)
        ref = reference_projection(wf, slices, z0, dz)
        last = z0 + (slices - 1) * dz
        rows.append((name, f"{abs(z0):.0f}-{abs(last):.0f}", asm.rel_l2(ref, exact), transfer_phase_error(n, abs(last))))
    if not args.quick:
        n, zg, iters, slices = 1024, 2000.0, 100, 512
        i1, i2 = synth.gs_pair(103, 500, n, zg)
        d32 = asm.transfer_sqrt_arr(n, LAM, DX)
        hf32, hb32 = asm.transfer_reference(zg, LAM, d32), asm.transfer_reference(-zg, LAM, d32)
        psi_ref = asm.phase_retrieval(i1, i2, iters, hf32.astype(np.complex64), hb32.astype(np.complex64))
        psi_exact = asm.phase_retrieval(i1, i2, iters, z=zg, wavlen=LAM, dx=DX)
        gs_rms = asm.weighted_phase_rms(psi_ref, psi_exact)
        exact = asm.reconst_xyprojection_running(psi_exact.astype(np.complex64), slices, z_front=-1e4, z_dz=-200.0,
                                                 wavlen=LAM, dx=DX)
        ref = reference_projection(psi_ref.astype(np.complex64), slices, -1e4, -200.0)
        rows.append((f"1024^2 GS {iters} it + {slices} slices (GS phase RMS {gs_rms:.2e} rad)",
                     f"1e4-{1e4 + (slices - 1) * 200:.0f}", asm.rel_l2(ref, exact), transfer_phase_error(n, 1e4 + (slices - 1) * 200)))
    print("| config | depths (um) | projection rel L2, reference arithmetic vs FP64 | H phase error at the deepest slice, max / RMS (rad) |")
    print("|---|---|---|---|")
    for name, depths, err, (pmax, prms) in rows:
        print(f"| {name} | {depths} | {err:.2e} | {pmax:.2e} / {prms:.2e} |")


if __name__ == "__main__":
    main()

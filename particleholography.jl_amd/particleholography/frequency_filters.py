"""Low-pass filters of ParticleHolography.jl's frequency_filters.jl over libholo.

``cu_rectangle_filter`` and ``cu_super_gaussian_filter`` build the centred Float32 filter on
the device (FP64 arithmetic in the reference's order, frequency_filters.jl:5-74);
``cu_apply_low_pass_filter_`` (the reference's ``cu_apply_low_pass_filter!``) and
``cu_apply_low_pass_filter`` compute IFFT2(ifftshift(lpf .* fftshift(FFT2(psi))))
(frequency_filters.jl:88-107) in two fused passes: forward row FFT, then forward column FFT
times the filter and inverse column FFT, then inverse row FFT scaled by 1/(nx*ny).
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import check
from .holofunc import CuLowPassFilter, CuWavefront, _contig, _dev, _h, _p, _stream

__all__ = ["cu_rectangle_filter", "cu_super_gaussian_filter", "cu_apply_low_pass_filter_",
           "cu_apply_low_pass_filter"]


def _filter(kind: int, prop_dist: float, wavlen: float, imglen: int, pixel_pitch: float,
            device: int) -> CuLowPassFilter:
    out = torch.empty((imglen, imglen), dtype=torch.float32, device=f"cuda:{device}")
    check(_lib.load().holo_lowpass_filter(_h(device), _stream(device), kind, float(prop_dist),
                                          float(wavlen), int(imglen), float(pixel_pitch), _p(out)))
    return CuLowPassFilter(out)


def cu_rectangle_filter(prop_dist: float, wavlen: float, imglen: int, pixel_picth: float,
                        device: int = 0) -> CuLowPassFilter:
    """frequency_filters.jl:31-38: 1 inside the cutoff circle (Fugal 2009 Eq. 13-14), 0 outside."""
    return _filter(_lib.HOLO_LPF_RECTANGLE, prop_dist, wavlen, imglen, pixel_picth, device)


def cu_super_gaussian_filter(prop_dist: float, wavlen: float, imglen: int, pixel_picth: float,
                             device: int = 0) -> CuLowPassFilter:
    """frequency_filters.jl:66-74: order-3 super-Gaussian, one half at the cutoff (Eq. 15)."""
    return _filter(_lib.HOLO_LPF_SUPER_GAUSSIAN, prop_dist, wavlen, imglen, pixel_picth, device)


def _apply(src: torch.Tensor, lpf: CuLowPassFilter, dst: torch.Tensor) -> None:
    """The filter must be a Float32 tensor of the wavefront's shape on its device (the reference's
    CuLowPassFilter{Float32}, types.jl:50-52); a non-contiguous one is copied."""
    device = _dev(src)
    if not isinstance(lpf.data, torch.Tensor) or _dev(lpf.data) != device:
        raise AssertionError(f"low-pass filter must be a tensor on cuda:{device}")
    filt = _contig(lpf.data, torch.float32)
    assert tuple(filt.shape) == tuple(src.shape), "low-pass filter and wavefront must have the same size"
    ny, nx = src.shape
    check(_lib.load().holo_apply_low_pass_filter(_h(device), _stream(device), nx, ny, _p(dst),
                                                 _p(src), _p(filt)))


def cu_apply_low_pass_filter_(holo: CuWavefront, lpf: CuLowPassFilter) -> None:
    """frequency_filters.jl:88-91 (cu_apply_low_pass_filter!): filters holo.data in place."""
    data = holo.data
    if data.is_contiguous() and data.dtype == torch.complex64:
        _apply(data, lpf, data)
    else:
        src = _contig(data, torch.complex64)
        _apply(src, lpf, src)
        data.copy_(src)


def cu_apply_low_pass_filter(holo: CuWavefront, lpf: CuLowPassFilter) -> CuWavefront:
    """frequency_filters.jl:105-107: a new filtered wavefront."""
    src = _contig(holo.data, torch.complex64)
    out = torch.empty_like(src)
    _apply(src, lpf, out)
    return CuWavefront(out)

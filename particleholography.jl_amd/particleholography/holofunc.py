"""Host-side mirror of ParticleHolography.jl's holofunc.jl / types.jl over libholo.

Same function names, argument order and meaning as the reference (src/holofunc.jl:5-6 exports),
with torch device tensors in place of CuArrays. A Julia array of size (nx, ny) is a torch
tensor of shape (ny, nx) here (same memory layout: x contiguous); a Julia volume of size
(nx, ny, slices) is a tensor of shape (slices, ny, nx).

Transfer functions carry their provenance (z, wavelength, pixel pitch) when built by
cu_transfer. With provenance on both transfers, reconstruction and phase retrieval generate
H on the device in FP64-wrapped form (``h_mode="auto"``/``"generated"``); transfers built from
raw arrays, or ``h_mode="arrays"``, use the arrays with the reference's ComplexF32 running
product. Errors map to the reference's: size mismatches raise AssertionError (holofunc.jl:101,
185, 217, 249, 301, 341), a bad return_type raises ArgumentError (holofunc.jl:164).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import torch

from . import _lib
from ._lib import ArgumentError, check, holo_tf

__all__ = [
    "CuTransferSqrtPart", "CuTransfer", "CuWavefront", "CuLowPassFilter", "N0f8", "Float16", "Float32", "Float64",
    "cu_transfer_sqrt_arr", "cu_transfer", "cu_gabor_wavefront", "cu_phase_retrieval_holo",
    "cu_get_reconst_vol", "cu_get_reconst_xyprojection", "cu_get_reconst_vol_and_xyprojection",
    "cu_get_reconst_complex_vol", "cu_asm_prop_", "cu_asm_prop", "cu_2d_pad",
    "cu_get_reconst_vol_and_xyprojection_padded", "cu_fft2", "cu_ifft2", "ArgumentError",
]


class N0f8:  # marker for FixedPointNumbers.N0f8 (stored as uint8 bit patterns)
    pass


Float16, Float32, Float64 = torch.float16, torch.float32, torch.float64
ComplexF32 = torch.complex64

_DTYPE_CODE = {N0f8: _lib.HOLO_N0F8, torch.float16: _lib.HOLO_F16, torch.float32: _lib.HOLO_F32,
               torch.float64: _lib.HOLO_F64}
_TORCH_OUT = {N0f8: torch.uint8, torch.float16: torch.float16, torch.float32: torch.float32,
              torch.float64: torch.float64}


# ---------------------------------------------------------------------------------------
# types.jl:13-61
# ---------------------------------------------------------------------------------------
@dataclass
class _Wrapped:
    data: torch.Tensor

    def size(self):
        """Julia size: (nx, ny) for a (ny, nx) tensor."""
        return tuple(reversed(self.data.shape))

    def eltype(self):
        return self.data.dtype

    def ndims(self):
        return self.data.dim()


@dataclass
class CuTransferSqrtPart(_Wrapped):
    """Square-root part d of the transfer function (types.jl:13-15); optional provenance."""
    wavlen: Optional[float] = None
    dx: Optional[float] = None


@dataclass
class CuTransfer(_Wrapped):
    """Centred ComplexF32 transfer function (types.jl:25-27); optional provenance."""
    z: Optional[float] = None
    wavlen: Optional[float] = None
    dx: Optional[float] = None

    @property
    def generated(self) -> bool:
        return self.z is not None and self.wavlen is not None and self.dx is not None


@dataclass
class CuWavefront(_Wrapped):
    """Complex wavefront (types.jl:37-39)."""


@dataclass
class CuLowPassFilter(_Wrapped):
    """Centred real low-pass filter (types.jl:50-52), multiplied with the centred spectrum."""


# ---------------------------------------------------------------------------------------
# plumbing
# ---------------------------------------------------------------------------------------
def _dev(t: torch.Tensor) -> int:
    if not t.is_cuda:
        raise AssertionError("arrays must be device (GPU) tensors")
    return t.device.index or 0


def _stream(device: int):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _h(device: int):
    return _lib.handle(device).ptr


def _p(t: Optional[torch.Tensor]):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p()


def _contig(t: torch.Tensor, dtype) -> torch.Tensor:
    if t.dtype != dtype:
        raise AssertionError(f"expected {dtype}, got {t.dtype}")
    return t.contiguous()


def _tf_struct(tf: CuTransfer, use_generated: bool, shape, device: int) -> holo_tf:
    """The C ABI's holo_tf for a transfer. An array-backed transfer (read by the kernels) must be a
    ComplexF32 (ny, nx) tensor on the call's device, as the reference's CuTransfer{ComplexF32}
    (types.jl:25-27); a non-contiguous one is copied, and the copy stays referenced by the
    returned struct until the call has been enqueued (the caching allocator keeps it valid for
    the stream's later work)."""
    s = holo_tf()
    if use_generated:
        s.data, s.z, s.wavlen, s.dx, s.generated = None, float(tf.z), float(tf.wavlen), float(tf.dx), 1
        return s
    a = tf.data
    if not isinstance(a, torch.Tensor) or a.dtype != torch.complex64:
        raise AssertionError(f"transfer arrays must be ComplexF32 (torch.complex64) tensors, got "
                             f"{getattr(a, 'dtype', type(a))}")
    if tuple(a.shape) != tuple(shape):
        raise AssertionError(f"transfer array must be {tuple(shape)}, got {tuple(a.shape)}")
    if not a.is_cuda:
        raise AssertionError("transfer arrays must be device (GPU) tensors")
    if (a.device.index or 0) != device:
        raise AssertionError(f"transfer array on cuda:{a.device.index or 0}, call on cuda:{device}")
    a = a.contiguous()
    s.data, s.z, s.wavlen, s.dx, s.generated = a.data_ptr(), 0.0, 0.0, 0.0, 0
    s._keep = a
    return s


def _choose_generated(h_mode: str, *tfs: CuTransfer) -> bool:
    if h_mode == "arrays":
        return False
    gen = all(t.generated for t in tfs)
    if h_mode == "generated":
        if not gen:
            raise AssertionError("h_mode='generated' needs transfers built by cu_transfer")
        return True
    if h_mode != "auto":
        raise ArgumentError(f"h_mode must be 'auto', 'generated' or 'arrays', got {h_mode!r}")
    return gen


# ---------------------------------------------------------------------------------------
# holofunc.jl
# ---------------------------------------------------------------------------------------
def cu_transfer_sqrt_arr(datlen: int, wavlen: float, dx: float, device: int = 0,
                         dtype=torch.float32) -> CuTransferSqrtPart:
    """holofunc.jl:30-36: d = 1 - (lambda fx)^2 - (lambda fy)^2 on the centred grid."""
    out = torch.empty((datlen, datlen), dtype=dtype, device=f"cuda:{device}")
    code = _lib.HOLO_F32 if dtype == torch.float32 else _lib.HOLO_F64
    check(_lib.load().holo_transfer_sqrt(_h(device), _stream(device), datlen, datlen, float(wavlen),
                                         float(dx), code, _p(out)))
    return CuTransferSqrtPart(out, wavlen=float(wavlen), dx=float(dx))


def cu_transfer(z0: float, datLen: int, wavLen: float, d_sqr: CuTransferSqrtPart,
                precision: str = "fp64") -> CuTransfer:
    """holofunc.jl:61-67: H = exp(2 pi i z0 / wavLen * sqrt(d)), centred ComplexF32.

    precision "fp64" (default): phase k*z*sqrt(d) in FP64, wrapped mod 2 pi before the cast;
    d recomputed in FP64 when ``d_sqr`` carries (wavlen, dx), else FP64 sqrt of its values.
    precision "reference": the reference's arithmetic (Float32 sqrt of d, FP64 exp)."""
    d = d_sqr.data
    assert d.shape == (datLen, datLen), f"d_sqr must be ({datLen}, {datLen}), got {tuple(d.shape)}"
    device = _dev(d)
    out = torch.empty((datLen, datLen), dtype=torch.complex64, device=d.device)
    d = d.contiguous()
    dcode = _lib.HOLO_F32 if d.dtype == torch.float32 else _lib.HOLO_F64
    if precision == "reference":
        mode = _lib.HOLO_TF_REFERENCE
    elif precision == "fp64":
        mode = _lib.HOLO_TF_FP64 if (d_sqr.dx is not None and d_sqr.wavlen == float(wavLen)) else _lib.HOLO_TF_FP64_D
    else:
        raise ArgumentError(f"precision must be 'fp64' or 'reference', got {precision!r}")
    dx = d_sqr.dx if d_sqr.dx is not None else 0.0
    check(_lib.load().holo_transfer(_h(device), _stream(device), datLen, datLen, float(z0), float(wavLen),
                                    _p(d), dcode, mode, float(dx), _p(out)))
    prov = d_sqr.dx is not None and precision == "fp64"
    return CuTransfer(out, z=float(z0) if prov else None, wavlen=float(wavLen) if prov else None,
                      dx=float(dx) if prov else None)


def cu_gabor_wavefront(holo: torch.Tensor) -> CuWavefront:
    """holofunc.jl:80-82: ComplexF32(sqrt(holo) + 0im)."""
    holo = _contig(holo, torch.float32)
    device = _dev(holo)
    ny, nx = holo.shape
    out = torch.empty((ny, nx), dtype=torch.complex64, device=holo.device)
    check(_lib.load().holo_gabor_wavefront(_h(device), _stream(device), nx, ny, _p(holo), _p(out)))
    return CuWavefront(out)


def cu_phase_retrieval_holo(holo1: torch.Tensor, holo2: torch.Tensor, transfer: CuTransfer,
                            invtransfer: CuTransfer, priter: int, datlen: int,
                            use_graph: bool = True, h_mode: str = "auto") -> CuWavefront:
    """holofunc.jl:100-129: Gerchberg-Saxton phase retrieval, returns psi at holo1's plane."""
    shapes = (tuple(holo1.shape), tuple(holo2.shape), tuple(transfer.data.shape), tuple(invtransfer.data.shape))
    assert all(s == (datlen, datlen) for s in shapes), \
        f"All arrays must have the same size as ({datlen}, {datlen}). Got {shapes}."
    holo1 = _contig(holo1, torch.float32)
    holo2 = _contig(holo2, torch.float32)
    device = _dev(holo1)
    gen = _choose_generated(h_mode, transfer, invtransfer)
    f, b = _tf_struct(transfer, gen, (datlen, datlen), device), _tf_struct(invtransfer, gen, (datlen, datlen), device)
    out = torch.empty((datlen, datlen), dtype=torch.complex64, device=holo1.device)
    check(_lib.load().holo_phase_retrieval(_h(device), _stream(device), datlen, datlen, _p(holo1), _p(holo2),
                                           ctypes.byref(f), ctypes.byref(b), int(priter), _p(out),
                                           1 if use_graph else 0))
    return CuWavefront(out)


def cu_2d_pad(inarr: torch.Tensor) -> torch.Tensor:
    """holofunc.jl:139-149: 2N x 2N filled with mean(inarr), inarr at offset N/2."""
    inarr = inarr.to(torch.complex64).contiguous()
    device = _dev(inarr)
    ny, nx = inarr.shape
    out = torch.empty((2 * ny, 2 * nx), dtype=torch.complex64, device=inarr.device)
    check(_lib.load().holo_2d_pad(_h(device), _stream(device), nx, ny, _p(inarr), _p(out)))
    return out


def _check_sizes(wavefront, tf_front, tf_dz, factor=1):
    wf = tuple(wavefront.data.shape)
    want = tuple(factor * s for s in wf)
    if factor == 1:
        assert wf == tuple(tf_front.data.shape) == tuple(tf_dz.data.shape), \
            f"All arrays must have the same size. Got {wf}, {tuple(tf_front.data.shape)}, {tuple(tf_dz.data.shape)}."
    else:
        assert want == tuple(tf_front.data.shape) == tuple(tf_dz.data.shape), \
            ("size(transfer_front.data) and size(transfer_dz.data) must be equal to 2*size(wavefront.data). "
             f"Got {wf}, {tuple(tf_front.data.shape)}, {tuple(tf_dz.data.shape)}.")


def cu_get_reconst_xyprojection(wavefront: CuWavefront, transfer_front: CuTransfer,
                                transfer_dz: CuTransfer, slices: int, h_mode: str = "auto") -> torch.Tensor:
    """holofunc.jl:248-266: min over slices of |psi_k|^2 (Float32 (ny, nx))."""
    _check_sizes(wavefront, transfer_front, transfer_dz)
    wf = _contig(wavefront.data, torch.complex64)
    device = _dev(wf)
    ny, nx = wf.shape
    gen = _choose_generated(h_mode, transfer_front, transfer_dz)
    f, d = _tf_struct(transfer_front, gen, (ny, nx), device), _tf_struct(transfer_dz, gen, (ny, nx), device)
    proj = torch.empty((ny, nx), dtype=torch.float32, device=wf.device)
    check(_lib.load().holo_reconst_xyprojection(_h(device), _stream(device), nx, ny, _p(wf), ctypes.byref(f),
                                                ctypes.byref(d), int(slices), _p(proj)))
    return proj


def reconst_xyprojection_holo(holo: torch.Tensor, transfer_front: CuTransfer, transfer_dz: CuTransfer,
                              slices: int, h_mode: str = "auto", out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """cu_gabor_wavefront + cu_get_reconst_xyprojection with sqrt(holo) fused into the first FFT pass."""
    holo = _contig(holo, torch.float32)
    device = _dev(holo)
    ny, nx = holo.shape
    assert (ny, nx) == tuple(transfer_front.data.shape) == tuple(transfer_dz.data.shape), "size mismatch"
    gen = _choose_generated(h_mode, transfer_front, transfer_dz)
    f, d = _tf_struct(transfer_front, gen, (ny, nx), device), _tf_struct(transfer_dz, gen, (ny, nx), device)
    proj = out if out is not None else torch.empty((ny, nx), dtype=torch.float32, device=holo.device)
    check(_lib.load().holo_reconst_xyprojection_holo(_h(device), _stream(device), nx, ny, _p(holo),
                                                     ctypes.byref(f), ctypes.byref(d), int(slices), _p(proj)))
    return proj


def _reconst_vol(wavefront, transfer_front, transfer_dz, slices, return_type, with_proj, h_mode):
    _check_sizes(wavefront, transfer_front, transfer_dz)
    if return_type is ComplexF32:
        code, tdtype = _lib.HOLO_C64, torch.complex64
    elif return_type in _DTYPE_CODE:
        code, tdtype = _DTYPE_CODE[return_type], _TORCH_OUT[return_type]
    else:
        raise ArgumentError(f"return_type must be Subtype of AbstractFloat or N0f8. Got {return_type}")
    wf = _contig(wavefront.data, torch.complex64)
    device = _dev(wf)
    ny, nx = wf.shape
    gen = _choose_generated(h_mode, transfer_front, transfer_dz)
    f, d = _tf_struct(transfer_front, gen, (ny, nx), device), _tf_struct(transfer_dz, gen, (ny, nx), device)
    vol = torch.empty((int(slices), ny, nx), dtype=tdtype, device=wf.device)
    proj = torch.empty((ny, nx), dtype=torch.float32, device=wf.device) if with_proj else None
    check(_lib.load().holo_reconst_vol(_h(device), _stream(device), nx, ny, _p(wf), ctypes.byref(f),
                                       ctypes.byref(d), int(slices), code, _p(vol), _p(proj)))
    return vol, proj


def cu_get_reconst_vol(wavefront: CuWavefront, transfer_front: CuTransfer, transfer_dz: CuTransfer,
                       slices: int, return_type=N0f8, h_mode: str = "auto") -> torch.Tensor:
    """holofunc.jl:184-200: clamp(|psi_k|^2, 0, 1) as return_type, shape (slices, ny, nx)."""
    return _reconst_vol(wavefront, transfer_front, transfer_dz, slices, return_type, False, h_mode)[0]


def cu_get_reconst_complex_vol(wavefront: CuWavefront, transfer_front: CuTransfer,
                               transfer_dz: CuTransfer, slices: int, h_mode: str = "auto") -> torch.Tensor:
    """holofunc.jl:216-232: psi_k (ComplexF32), shape (slices, ny, nx)."""
    return _reconst_vol(wavefront, transfer_front, transfer_dz, slices, ComplexF32, False, h_mode)[0]


def cu_get_reconst_vol_and_xyprojection(wavefront: CuWavefront, transfer_front: CuTransfer,
                                        transfer_dz: CuTransfer, slices: int, return_type=N0f8,
                                        h_mode: str = "auto"):
    """holofunc.jl:300-321: (vol, min over slices of the stored values)."""
    if return_type is ComplexF32:
        raise ArgumentError("return_type must be Subtype of AbstractFloat or N0f8. Got ComplexF32")
    return _reconst_vol(wavefront, transfer_front, transfer_dz, slices, return_type, True, h_mode)


def cu_get_reconst_vol_and_xyprojection_padded(wavefront: CuWavefront, transfer_front: CuTransfer,
                                               transfer_dz: CuTransfer, slices: int, return_type=N0f8,
                                               h_mode: str = "auto"):
    """holofunc.jl:339-364: reconstruct the mean-padded 2N wavefront, keep the centre N x N of
    every slice as N0f8 (whatever return_type is, holofunc.jl:344) and its xy projection."""
    if return_type is not N0f8 and return_type not in _DTYPE_CODE:
        raise ArgumentError(f"return_type must be Subtype of AbstractFloat or N0f8. Got {return_type}")
    _check_sizes(wavefront, transfer_front, transfer_dz, factor=2)
    wf = _contig(wavefront.data, torch.complex64)
    device = _dev(wf)
    ny, nx = wf.shape
    gen = _choose_generated(h_mode, transfer_front, transfer_dz)
    f, d = (_tf_struct(transfer_front, gen, (2 * ny, 2 * nx), device),
            _tf_struct(transfer_dz, gen, (2 * ny, 2 * nx), device))
    vol = torch.empty((int(slices), ny, nx), dtype=torch.uint8, device=wf.device)
    proj = torch.empty((ny, nx), dtype=torch.float32, device=wf.device)
    check(_lib.load().holo_reconst_vol_and_xyprojection_padded(_h(device), _stream(device), nx, ny, _p(wf),
                                                               ctypes.byref(f), ctypes.byref(d), int(slices),
                                                               _p(vol), _p(proj)))
    return vol, proj


def cu_asm_prop_(outholo: CuWavefront, inholo: CuWavefront, d_sqr: CuTransferSqrtPart, zprop: float,
                 datlen: int, wavlen: float) -> None:
    """holofunc.jl:383-388 (cu_asm_prop!): outholo.data = IFFT2(FFT2(in) .* H(zprop))."""
    src = _contig(inholo.data, torch.complex64)
    assert tuple(src.shape) == (datlen, datlen) == tuple(outholo.data.shape), "size mismatch"
    device = _dev(src)
    if d_sqr.dx is not None and d_sqr.wavlen == float(wavlen):
        tf = CuTransfer(d_sqr.data, z=float(zprop), wavlen=float(wavlen), dx=d_sqr.dx)
        s = _tf_struct(tf, True, (datlen, datlen), device)
    else:
        tf = cu_transfer(zprop, datlen, wavlen, d_sqr)
        s = _tf_struct(tf, False, (datlen, datlen), device)
    dst = outholo.data
    tmp = dst if dst.is_contiguous() and dst.dtype == torch.complex64 else torch.empty_like(src)
    check(_lib.load().holo_asm_prop(_h(device), _stream(device), datlen, datlen, _p(tmp), _p(src), ctypes.byref(s)))
    if tmp is not dst:
        dst.copy_(tmp)


cu_asm_prop = cu_asm_prop_


def _fft2(x: torch.Tensor, inverse: bool) -> torch.Tensor:
    x = _contig(x, torch.complex64)
    device = _dev(x)
    ny, nx = x.shape
    out = torch.empty_like(x)
    check(_lib.load().holo_fft2(_h(device), _stream(device), nx, ny, _p(x), _p(out), 1 if inverse else 0))
    return out


def cu_fft2(x: torch.Tensor) -> torch.Tensor:
    """Unnormalized 2-D forward FFT (the CUFFT.fft calls of holofunc.jl)."""
    return _fft2(x, False)


def cu_ifft2(x: torch.Tensor) -> torch.Tensor:
    """2-D inverse FFT scaled by 1/(nx*ny) (CUFFT.ifft, AbstractFFTs convention)."""
    return _fft2(x, True)

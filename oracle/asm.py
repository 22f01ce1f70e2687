"""CPU oracle for the angular-spectrum / Gerchberg-Saxton path (test infrastructure only).

Only tests, the smoke check and the CPU-baseline leg of the benchmark may import it; the
product path never calls it.

Array convention: numpy arrays are C-ordered with shape ``(ny, nx)``. Element ``[j, i]``
is the reference's Julia element ``[i+1, j+1]`` (Julia dim 1 = ``x`` = the contiguous axis),
so a numpy buffer and the reference's column-major ``CuArray`` have the same memory layout.

Two arithmetic modes are restated:

* ``exact`` -- FP64 everywhere: FP64 ``d``, FP64 ``sqrt``, FP64 ``exp``, FP64 FFTs. This is
  the mathematics of the method (docs/src/whats_inline_holography.md:11-15, 38-42, 60-89).
* ``reference`` -- the reference's own arithmetic for the transfer functions: Float32 ``d``
  (holofunc.jl:8-15), Float32 ``sqrt`` then FP64 ``exp`` rounded to ComplexF32
  (holofunc.jl:38-45), and a ComplexF32 running product ``F .*= Hdz`` (holofunc.jl:259-263).
  FFTs are done in FP64 on the rounded ComplexF32 values (cuFFT's own rounding is not
  restated; it is a closed library).

Parity status of every function is listed in ``PARITY`` at the bottom of this file.
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2.0 * np.pi  # Julia ``2im * pi`` -> ComplexF64(0, 2.0*pi)


# --------------------------------------------------------------------------------------
# A2: square-root part  (holofunc.jl:8-15, 30-36)
# --------------------------------------------------------------------------------------
def axis_term(n: int, wavlen: float, dx: float) -> np.ndarray:
    """((x - n/2 - 1.0) * wavlen / n / dx)^2 for 1-based x = 1..n, in FP64 with the
    reference's operation order (holofunc.jl:12). Index is the *centred* (fftshifted) one."""
    x = np.arange(1, n + 1, dtype=np.float64)
    t = (x - n / 2 - 1.0) * wavlen / n / dx
    return t * t  # Julia literal ``^2`` lowers to ``t*t``


def transfer_sqrt_arr(nx: int, wavlen: float, dx: float, ny: int | None = None,
                      dtype=np.float32) -> np.ndarray:
    """d[x,y] = 1 - ax - ay, evaluated left to right in FP64 and stored as ``dtype``
    (holofunc.jl:12: ``1.0 - (...)^2 - (...)^2``). Centred layout, shape (ny, nx)."""
    ny = nx if ny is None else ny
    ax = axis_term(nx, wavlen, dx)
    ay = axis_term(ny, wavlen, dx)
    d = (1.0 - ax[None, :]) - ay[:, None]
    return d.astype(dtype)


# --------------------------------------------------------------------------------------
# A3: transfer function  (holofunc.jl:38-45, 61-67)
# --------------------------------------------------------------------------------------
def transfer_reference(z: float, wavlen: float, d32: np.ndarray) -> np.ndarray:
    """Reference arithmetic: theta = ((2pi*z)/wavlen) * Float64(sqrtf(d32)); exp in FP64;
    rounded to ComplexF32 (holofunc.jl:42). Negative d gives NaN, as sqrtf does."""
    with np.errstate(invalid="ignore"):
        s32 = np.sqrt(d32.astype(np.float32))  # Float32 sqrt, correctly rounded
    theta = ((TWO_PI * z) / wavlen) * s32.astype(np.float64)
    return (np.cos(theta) + 1j * np.sin(theta)).astype(np.complex64)


def transfer_exact(z: float, wavlen: float, dx: float, nx: int, ny: int | None = None) -> np.ndarray:
    """Exact FP64 transfer function exp(i 2pi z/wavlen sqrt(d)) on the centred grid, complex128.
    Evanescent components (d < 0) are set to 0 (the build's documented choice; the
    reference yields NaN there)."""
    d = transfer_sqrt_arr(nx, wavlen, dx, ny, dtype=np.float64)
    s = np.sqrt(np.maximum(d, 0.0))
    theta = (TWO_PI * z / wavlen) * s
    h = np.cos(theta) + 1j * np.sin(theta)
    h[d < 0] = 0.0
    return h


def natural(h_centred: np.ndarray) -> np.ndarray:
    """ifftshift: centred (fftshifted) layout -> natural FFT order. AbstractFFTs and numpy
    agree (circshift by -div(n,2) per axis)."""
    return np.fft.ifftshift(h_centred)


# --------------------------------------------------------------------------------------
# A4, A12: wavefront and FFT conventions
# --------------------------------------------------------------------------------------
def gabor_wavefront(holo: np.ndarray) -> np.ndarray:
    """ComplexF32(sqrt(holo) + 0im)  (holofunc.jl:80-82); sqrt in Float32."""
    with np.errstate(invalid="ignore"):
        return np.sqrt(holo.astype(np.float32)).astype(np.complex64)


def fft2(x):
    return np.fft.fft2(x.astype(np.complex128))


def ifft2(x):
    """Inverse FFT scaled by 1/(nx*ny) (AbstractFFTs convention)."""
    return np.fft.ifft2(x.astype(np.complex128))


def propagate(psi: np.ndarray, h_centred: np.ndarray) -> np.ndarray:
    """IFFT2(ifftshift(fftshift(FFT2(psi)) .* H)) = IFFT2(FFT2(psi) .* ifftshift(H))
    (holofunc.jl:114, 121, 386). FP64."""
    return ifft2(fft2(psi) * natural(h_centred))


def asm_prop(psi: np.ndarray, z: float, wavlen: float, dx: float, mode: str = "exact") -> np.ndarray:
    """cu_asm_prop! (holofunc.jl:383-388): builds H(z) and propagates once."""
    ny, nx = psi.shape
    if mode == "exact":
        h = transfer_exact(z, wavlen, dx, nx, ny)
    else:
        h = transfer_reference(z, wavlen, transfer_sqrt_arr(nx, wavlen, dx, ny))
    return propagate(psi, h)


# --------------------------------------------------------------------------------------
# A6-A9: slice reconstruction  (holofunc.jl:184-200, 216-232, 248-266, 268-282, 300-321)
# --------------------------------------------------------------------------------------
def _slice_fields(wavefront, slices, mode, z_front=None, z_dz=None, wavlen=None, dx=None,
                  h_front=None, h_dz=None):
    """Yields psi_k for k = 0..slices-1, psi_k at z_front + k*z_dz.

    mode 'exact': H generated in FP64 for each slice (complex128 spectrum, FP64 FFTs).
    mode 'reference': ComplexF32 spectrum ``fftshift(fft(psi)) .* H0`` then ``.*= Hdz``
    per slice in ComplexF32 (holofunc.jl:254-260), FP64 inverse FFT of the rounded values.
    """
    ny, nx = wavefront.shape
    if mode == "exact":
        F = fft2(wavefront)
        d = transfer_sqrt_arr(nx, wavlen, dx, ny, dtype=np.float64)
        s = natural(np.sqrt(np.maximum(d, 0.0)))
        evan = natural(d < 0)
        for k in range(slices):
            z = z_front + k * z_dz
            theta = (TWO_PI * z / wavlen) * s
            h = np.cos(theta) + 1j * np.sin(theta)
            h[evan] = 0.0
            yield ifft2(F * h)
    elif mode == "reference":
        F = np.fft.fftshift(fft2(wavefront)).astype(np.complex64)
        F = (F * h_front.astype(np.complex64)).astype(np.complex64)
        hdz = h_dz.astype(np.complex64)
        for k in range(slices):
            if k > 0:
                F = (F * hdz).astype(np.complex64)
            yield ifft2(np.fft.ifftshift(F))
    else:
        raise ValueError(mode)


def reconst_xyprojection(wavefront, slices, mode="exact", **kw) -> np.ndarray:
    """min over slices of abs2(psi_k), unclamped, returned as Float32 (holofunc.jl:248-266).
    ``CUDA.min.(proj, projtmp)`` runs CUDA.jl's device method of ``min`` for Float32, which
    calls libdevice ``fminf``: a NaN operand is ignored (np.fmin has the same rule)."""
    proj = None
    for psi in _slice_fields(wavefront, slices, mode, **kw):
        inten = (psi.real ** 2 + psi.imag ** 2)
        proj = inten if proj is None else np.fmin(proj, inten)
    return proj.astype(np.float32)


def reconst_xyprojection_running(wavefront, slices, *, z_front, z_dz, wavlen, dx, workers=None) -> np.ndarray:
    """The FP64 projection of ``reconst_xyprojection(mode='exact')`` for long slice stacks
    (holofunc.jl:248-266): the spectrum is advanced from slice to slice by the complex128 product
    with H(z_dz) (the composition identity H(a) H(b) = H(a + b) of propagating components; FP64
    drift over 1000 products ~1e-13), and the FFTs run on ``workers`` threads (scipy.fft, FP64).
    Pinned against the per-slice evaluation in tests/test_oracle.py."""
    import scipy.fft as sfft

    if workers is None:
        try:
            import os

            workers = max(1, min(16, len(os.sched_getaffinity(0))))
        except AttributeError:  # pragma: no cover
            workers = 1
    ny, nx = wavefront.shape
    spec = sfft.fft2(wavefront.astype(np.complex128), workers=workers)
    d = transfer_sqrt_arr(nx, wavlen, dx, ny, dtype=np.float64)
    s = natural(np.sqrt(np.maximum(d, 0.0)))
    evan = natural(d < 0)

    def cis(z):
        theta = (TWO_PI * z / wavlen) * s
        h = np.cos(theta) + 1j * np.sin(theta)
        h[evan] = 0.0
        return h

    g = spec * cis(z_front)
    hdz = cis(z_dz)
    proj = None
    for k in range(slices):
        if k:
            g *= hdz
        psi = sfft.ifft2(g, workers=workers)
        inten = psi.real ** 2 + psi.imag ** 2
        proj = inten if proj is None else np.fmin(proj, inten, out=proj)
    return proj.astype(np.float32)


def n0f8(v: np.ndarray) -> np.ndarray:
    """reinterpret(N0f8, round(UInt8, clamp(v,0,1)*255)) with ties-to-even
    (holofunc.jl:154-156). ``v`` is promoted to FP64 first (clamp with 0.0/1.0 literals)."""
    c = np.clip(v.astype(np.float64), 0.0, 1.0)
    return np.rint(c * 255.0).astype(np.uint8)


def store_value(inten_f32: np.ndarray, out_dtype) -> np.ndarray:
    """_ifft_and_abs epilogue (holofunc.jl:158-166): abs2 (Float32) -> clamp(.,0.0,1.0) in FP64
    -> return_type. ``out_dtype`` in {np.uint8 (N0f8), np.float16, np.float32, np.float64}."""
    if out_dtype == np.uint8:
        return n0f8(inten_f32)
    c = np.clip(inten_f32.astype(np.float64), 0.0, 1.0)
    return c.astype(out_dtype)


def reconst_vol(wavefront, slices, out_dtype=np.uint8, mode="exact", **kw) -> np.ndarray:
    """cu_get_reconst_vol (holofunc.jl:184-200). Shape (slices, ny, nx) = Julia (nx, ny, slices)."""
    out = []
    for psi in _slice_fields(wavefront, slices, mode, **kw):
        inten = (psi.real ** 2 + psi.imag ** 2).astype(np.float32)
        out.append(store_value(inten, out_dtype))
    return np.stack(out)


def reconst_complex_vol(wavefront, slices, mode="exact", **kw) -> np.ndarray:
    """cu_get_reconst_complex_vol (holofunc.jl:216-232): psi_k, unclamped."""
    return np.stack([psi for psi in _slice_fields(wavefront, slices, mode, **kw)])


def stored_to_f32(vol: np.ndarray) -> np.ndarray:
    """Float32 view of stored values: N0f8 -> u/255 (correctly rounded), floats -> Float32."""
    if vol.dtype == np.uint8:
        return (vol.astype(np.float32) / np.float32(255.0)).astype(np.float32)
    return vol.astype(np.float32)


def xy_projection_from_vol(vol: np.ndarray) -> np.ndarray:
    """_cu_get_xy_projection_from_vol! (holofunc.jl:268-282): ``val < min ? val : min`` over z,
    starting from slice 1, so a NaN is ignored after the first slice; stored into Float32."""
    vals = stored_to_f32(vol)
    m = vals[0].copy()
    for k in range(1, vals.shape[0]):
        v = vals[k]
        m = np.where(v < m, v, m)
    return m


def reconst_vol_and_xyprojection(wavefront, slices, out_dtype=np.uint8, mode="exact", **kw):
    """cu_get_reconst_vol_and_xyprojection (holofunc.jl:300-321)."""
    vol = reconst_vol(wavefront, slices, out_dtype, mode, **kw)
    return vol, xy_projection_from_vol(vol)


# --------------------------------------------------------------------------------------
# f2: padding and the padded reconstruction  (holofunc.jl:139-149, 339-364)
# --------------------------------------------------------------------------------------
def pad2(a: np.ndarray) -> np.ndarray:
    """cu_2d_pad: 2N x 2N filled with ComplexF32(mean(a)), a copied at offset N/2 (0-based
    ``outlen÷4``) in both axes (holofunc.jl:139-149)."""
    ny, nx = a.shape
    m = np.complex64(a.astype(np.complex128).mean())
    out = np.full((2 * ny, 2 * nx), m, dtype=np.complex64)
    oy, ox = (2 * ny) // 4, (2 * nx) // 4
    out[oy:oy + ny, ox:ox + nx] = a.astype(np.complex64)
    return out


def reconst_vol_and_xyprojection_padded(wavefront, slices, mode="exact", **kw):
    """Padded variant: reconstruct at 2N, keep the centre window [N/2, 3N/2) of each slice as
    N0f8 whatever return_type is (holofunc.jl:344-356), projection from the stored values."""
    ny, nx = wavefront.shape
    padded = pad2(wavefront)
    vols = []
    oy, ox = ny // 2, nx // 2
    for psi in _slice_fields(padded, slices, mode, **kw):
        inten = (psi.real ** 2 + psi.imag ** 2).astype(np.float32)
        vols.append(n0f8(inten[oy:oy + ny, ox:ox + nx]))
    vol = np.stack(vols)
    return vol, xy_projection_from_vol(vol)


# --------------------------------------------------------------------------------------
# A10: Gerchberg-Saxton  (holofunc.jl:100-129)
# --------------------------------------------------------------------------------------
def phase_retrieval(holo1, holo2, iters, h_fwd=None, h_bwd=None, *, z=None, wavlen=None,
                    dx=None) -> np.ndarray:
    """GS in FP64: psi1 = sqrt(I1); repeat: psi2 = P+(psi1); psi2 = sqrt(I2) e^{i angle psi2};
    psi1 = P-(psi2); psi1 = sqrt(I1) e^{i angle psi1}. angle(0) = 0. Transfer functions are the
    given centred arrays (h_fwd / h_bwd, used at FP64) or, with ``z``, exact H(+z) / H(-z).
    The angle is taken of the spatial field (holofunc.jl:115, 122)."""
    ny, nx = holo1.shape
    if z is not None:
        hf = natural(transfer_exact(z, wavlen, dx, nx, ny))
        hb = natural(transfer_exact(-z, wavlen, dx, nx, ny))
    else:
        hf = natural(h_fwd.astype(np.complex128))
        hb = natural(h_bwd.astype(np.complex128))
    with np.errstate(invalid="ignore"):
        a1 = np.sqrt(holo1.astype(np.float32)).astype(np.float64)
        a2 = np.sqrt(holo2.astype(np.float32)).astype(np.float64)
    psi1 = a1.astype(np.complex128)
    for _ in range(iters):
        psi2 = ifft2(fft2(psi1) * hf)
        psi2 = a2 * np.exp(1j * np.angle(psi2))
        psi1 = ifft2(fft2(psi2) * hb)
        psi1 = a1 * np.exp(1j * np.angle(psi1))
    return psi1


def gs_error(psi1, holo2, h_fwd_nat) -> float:
    """Fienup error-reduction error at plane 2: || |P+(psi1)| - sqrt(I2) ||_2 (used for the
    non-increase invariant, Fienup, Appl. Opt. 21, 2758 (1982))."""
    a2 = np.sqrt(holo2.astype(np.float64))
    psi2 = ifft2(fft2(psi1) * h_fwd_nat)
    return float(np.linalg.norm(np.abs(psi2) - a2))


# --------------------------------------------------------------------------------------
# metrics
# --------------------------------------------------------------------------------------
# --------------------------------------------------------------------------------------
# Low-pass filters  (frequency_filters.jl:5-107)
# --------------------------------------------------------------------------------------
def lpf_cutoff(prop_dist: float, wavlen: float, imglen: int, dx: float) -> float:
    """Cutoff radius in frequency samples, FP64 in the reference's order
    (frequency_filters.jl:33, 68): 1/wavlen * imglen^2 * dx^2 / sqrt(4 z^2 + imglen^2 dx^2)."""
    n2 = float(imglen * imglen)
    p2 = dx * dx
    return 1.0 / wavlen * n2 * p2 / np.sqrt(4.0 * (prop_dist * prop_dist) + n2 * p2)


def super_gaussian_sigma(prop_dist: float, wavlen: float, imglen: int, dx: float) -> float:
    """sigma_x = maxi / (imglen * dx) / (2 log 2)^(1/6)  (frequency_filters.jl:69)."""
    return lpf_cutoff(prop_dist, wavlen, imglen, dx) / (imglen * dx) / (2.0 * np.log(2.0)) ** (1 / 6)


def rectangle_filter(prop_dist: float, wavlen: float, imglen: int, dx: float) -> np.ndarray:
    """frequency_filters.jl:5-15: arr[y, x] = 0 where (x - n/2 + 1)^2 + (y - n/2)^2 >= maxi^2,
    1 elsewhere (1-based x = Julia dim 2, y = Julia dim 1). Shape (n, n), Float32."""
    maxi = lpf_cutoff(prop_dist, wavlen, imglen, dx)
    idx = np.arange(1, imglen + 1, dtype=np.float64)
    hx = idx - imglen / 2 + 1.0   # x: numpy row index + 1
    hy = idx - imglen / 2         # y: numpy column index + 1
    r2 = hx[:, None] * hx[:, None] + hy[None, :] * hy[None, :]
    return np.where(r2 >= maxi * maxi, 0.0, 1.0).astype(np.float32)


def super_gaussian_at(x, y, imglen: int, dx: float, sigma: float):
    """exp(-1/2 * (((x - n/2 + 1)/n/dx/sigma)^2 + ((y - n/2 + 1)/n/dx/sigma)^2)^3) at 1-based
    (possibly fractional) Julia coordinates x (dim 2), y (dim 1), FP64
    (frequency_filters.jl:46)."""
    ax = (np.asarray(x, np.float64) - imglen / 2 + 1.0) / imglen / dx / sigma
    ay = (np.asarray(y, np.float64) - imglen / 2 + 1.0) / imglen / dx / sigma
    s = ax * ax + ay * ay
    return np.exp(-0.5 * (s * s * s))


def super_gaussian_filter(prop_dist: float, wavlen: float, imglen: int, dx: float) -> np.ndarray:
    """frequency_filters.jl:40-50, 66-74 on the integer grid, rounded to Float32. Shape (n, n)."""
    sigma = super_gaussian_sigma(prop_dist, wavlen, imglen, dx)
    idx = np.arange(1, imglen + 1, dtype=np.float64)
    return super_gaussian_at(idx[:, None], idx[None, :], imglen, dx, sigma).astype(np.float32)


def apply_low_pass_filter(psi: np.ndarray, lpf: np.ndarray) -> np.ndarray:
    """ifft(ifftshift(lpf .* fftshift(fft(psi))))  (frequency_filters.jl:89, 106), FP64 FFTs
    with the reference's explicit shifts."""
    spec = np.fft.fftshift(np.fft.fft2(np.asarray(psi, np.complex128)))
    return np.fft.ifft2(np.fft.ifftshift(np.asarray(lpf, np.float64) * spec))


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def weighted_phase_rms(psi, psi_ref, weight=None) -> float:
    """sqrt(sum A^2 wrap(phi - phi_ref)^2 / sum A^2), A = |psi_ref| unless given."""
    w = np.abs(psi_ref) ** 2 if weight is None else np.asarray(weight, np.float64) ** 2
    dphi = np.angle(psi * np.conj(psi_ref))  # wrapped difference in (-pi, pi]
    return float(np.sqrt(np.sum(w * dphi ** 2) / np.sum(w)))


# --------------------------------------------------------------------------------------
# Which check pins which function (see tests/test_oracle.py)
# --------------------------------------------------------------------------------------
PARITY = {
    "transfer_sqrt_arr": "closed form: brute-force scalar re-evaluation + centre value d=1 at x=n/2+1",
    "transfer_reference": "H(0)=1, |H|=1, H(a)H(b)=H(a+b) to F32 rounding, phase vs exact <= table bound",
    "transfer_exact": "plane-wave eigenfunction, composition, O(N^4) brute-force DFT propagation",
    "propagate/asm_prop": "constant field, single-frequency grating closed form, delta<->constant, "
                          "reversibility, Parseval, brute-force DFT at 8^2/16^2",
    "reconst_xyprojection": "constant field -> 1, grating closed-form min over z, running-product "
                            "identity (reference mode vs exact), brute-force DFT",
    "reconst_vol/complex_vol/vol_and_xyprojection": "per-slice equal to propagate(); N0f8 rounding "
                                                    "ties; NaN rule of the projection kernel",
    "pad2/padded": "cu_2d_pad known-answer test (runtests.jl:362-368)",
    "phase_retrieval": "plane-wave fixed point, |psi1| == sqrt(I1), error-reduction non-increase, "
                       "brute-force DFT GS iteration at 8^2",
    "lpf_cutoff/rectangle_filter": "cutoff = n dx sin(theta)/wavlen with tan(theta) = (n dx/2)/z "
                                   "(Fugal 2009 geometry), pass-band area ~ pi maxi^2, disk centred "
                                   "one sample apart on the two axes, hand-checked 8x8 mask",
    "super_gaussian_filter": "1 at the centre, exactly 1/2 at radius maxi on both axes, radial "
                             "monotonicity, x<->y symmetry",
    "apply_low_pass_filter": "all-ones filter = identity, DC-only filter = mean, single-frequency "
                             "eigenfunction scaled by its filter value, O(N^4) brute-force DFT",
}


# --------------------------------------------------------------------------------------
# Ingest: load_gray2float, make_background, cu_make_background_mode (utils.jl:23-25,
# 217-317) and the tutorial's background removal (preprocessings.md:21-23)
# --------------------------------------------------------------------------------------
def n0f8_to_f32(u8) -> np.ndarray:
    """Float32(reinterpret(N0f8, v)) = v / 255 correctly rounded to Float32."""
    return np.asarray(u8, np.uint8).astype(np.float32) / np.float32(255)


def ingest_u8(u8, bkg=None, offset: float = 0.0) -> np.ndarray:
    """Float32((Float64(v/255) - bkg) + offset) with bkg 0 when absent; with a background and
    offset 64/255 this is ``img .= img .- bkg .+ 64/255`` on a Float32 img."""
    f = n0f8_to_f32(u8).astype(np.float64)
    b = 0.0 if bkg is None else np.asarray(bkg, np.float64)
    return ((f - b) + offset).astype(np.float32)


def background_mode(stack) -> np.ndarray:
    """Per-pixel mode of a uint8 stack (count, ...) / 255.0: votes per value, then the first
    (smallest) value with the largest count (utils.jl:229-237, 254-271, 313-315)."""
    s = np.asarray(stack, np.uint8)
    flat = s.reshape(s.shape[0], -1)
    npx = flat.shape[1]
    mode = np.empty(npx, np.int64)
    chunk = 1 << 16
    for c0 in range(0, npx, chunk):
        part = flat[:, c0:c0 + chunk].astype(np.int64)
        w = part.shape[1]
        idx = part * w + np.arange(w)[None, :]  # vote (value, pixel) -> value * w + pixel
        votes = np.bincount(idx.ravel(), minlength=256 * w).reshape(256, w)
        mode[c0:c0 + w] = np.argmax(votes, axis=0)  # first maximum = smallest value
    return (mode.astype(np.float64) / 255.0).reshape(s.shape[1:])


def background_mean(stack) -> np.ndarray:
    """Per-pixel FP64 sum of Float32(v/255) in image order from 0.0, then / count
    (utils.jl:219-224)."""
    s = np.asarray(stack, np.uint8)
    acc = np.zeros(s.shape[1:], np.float64)
    for img in s:
        acc = acc + n0f8_to_f32(img).astype(np.float64)
    return acc / s.shape[0]


PARITY.update({
    "n0f8_to_f32": "exhaustive over 0..255 against the exact rational v/255 rounded to Float32",
    "ingest_u8": "scalar re-evaluation with Python floats (IEEE FP64) on every byte value",
    "background_mode": "cu_make_background_mode known-answer test (runtests.jl:125-138), "
                       "collections.Counter brute force with smallest-value tie break",
    "background_mean": "sequential Python-float sum, and math.fsum mean within 1e-15 relative",
})

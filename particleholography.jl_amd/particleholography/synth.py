"""Seeded synthetic inline holograms of opaque disks (input generator for tests and benchmarks).

Scene: opaque disks with diameters ~ U(dmin, dmax) um, centres x, y ~ U over the field
[0, n*dx), depths z ~ U(zmin, zmax) um from the sensor. Formation follows the method text of
the reference (docs/src/whats_inline_holography.md:20-33): a unit plane wave is multiplied by
(1 - A_i) at each disk plane, farthest first, and carried between planes and finally to the
sensor by the FP64 angular spectrum method with the natural-order transfer function
exp(i 2 pi dz / lambda sqrt(1 - (lambda fx)^2 - (lambda fy)^2)). The recorded intensity is
|psi|^2 + N(0, noise^2), clamped at 0 (a generator choice: sqrt of a negative intensity is NaN
in the reference) and cast to Float32.

To bound the cost of large scenes, disks may be binned into ``max_planes`` equally wide depth
bins; each bin is applied at its mean depth. The bin width is (zmax - zmin) / max_planes.
"""
from __future__ import annotations

import os

import numpy as np

try:  # scipy's pocketfft runs multithreaded; numpy's is single threaded
    import scipy.fft as _sfft

    def _fft2(a):
        return _sfft.fft2(a, workers=_workers())

    def _ifft2(a):
        return _sfft.ifft2(a, workers=_workers())
except ImportError:  # pragma: no cover
    _fft2, _ifft2 = np.fft.fft2, np.fft.ifft2


def _workers() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _kz_cycles(nx, ny, wavlen, dx):
    """sqrt(1 - (lambda fx)^2 - (lambda fy)^2) / lambda on the natural-order grid (FP64)."""
    fx = np.fft.fftfreq(nx, d=dx)
    fy = np.fft.fftfreq(ny, d=dx)
    d = 1.0 - (wavlen * fx[None, :]) ** 2 - (wavlen * fy[:, None]) ** 2
    return np.sqrt(np.maximum(d, 0.0)) / wavlen, d < 0


def disk_scene(seed: int, ndisks: int, nx: int, ny: int | None = None, zmin: float = 1e4,
               zmax: float = 1e5, dmin: float = 10.0, dmax: float = 60.0, dx: float = 10.0):
    """Returns (x, y, z, diameter) arrays in um, drawn from PCG64(seed) in that order."""
    ny = nx if ny is None else ny
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.uniform(0.0, nx * dx, ndisks)
    y = rng.uniform(0.0, ny * dx, ndisks)
    z = rng.uniform(zmin, zmax, ndisks)
    diam = rng.uniform(dmin, dmax, ndisks)
    return x, y, z, diam


def _apply_disk(psi, x0, y0, r, dx):
    ny, nx = psi.shape
    i0, i1 = int(np.floor((x0 - r) / dx)) - 1, int(np.ceil((x0 + r) / dx)) + 1
    j0, j1 = int(np.floor((y0 - r) / dx)) - 1, int(np.ceil((y0 + r) / dx)) + 1
    i0, j0 = max(i0, 0), max(j0, 0)
    i1, j1 = min(i1, nx - 1), min(j1, ny - 1)
    if i1 < i0 or j1 < j0:
        return
    ii = (np.arange(i0, i1 + 1) + 0.5) * dx  # pixel centres
    jj = (np.arange(j0, j1 + 1) + 0.5) * dx
    mask = (ii[None, :] - x0) ** 2 + (jj[:, None] - y0) ** 2 <= r * r
    psi[j0:j1 + 1, i0:i1 + 1][mask] = 0.0


def _planes(z, zmin, zmax, max_planes):
    """Depth planes farthest first, as (depth, member indices)."""
    if max_planes is not None and z.size > max_planes:
        width = (zmax - zmin) / max_planes
        b = np.minimum(((z - zmin) / width).astype(np.int64), max_planes - 1)
        return [(float(z[b == bi].mean()), np.nonzero(b == bi)[0])
                for bi in range(max_planes - 1, -1, -1) if np.any(b == bi)]
    return [(float(z[i]), np.array([i])) for i in np.argsort(-z, kind="stable")]


def sensor_field(scene, nx, ny, zmin, zmax, wavlen, dx, max_planes):
    """Complex field at the sensor plane (z = 0) behind the disks of ``scene`` (FP64)."""
    x, y, z, diam = scene
    kzc, evan = _kz_cycles(nx, ny, wavlen, dx)
    psi = np.ones((ny, nx), dtype=np.complex128)
    zcur = None
    for zp, members in _planes(z, zmin, zmax, max_planes):
        if zcur is not None:
            psi = _propagate(psi, zcur - zp, kzc, evan)
        for i in members:
            _apply_disk(psi, x[i], y[i], diam[i] / 2.0, dx)
        zcur = zp
    if zcur is not None:
        psi = _propagate(psi, zcur, kzc, evan)
    return psi


def _record(psi, rng, noise):
    inten = np.abs(psi) ** 2 + rng.normal(0.0, noise, size=psi.shape)
    return np.maximum(inten, 0.0).astype(np.float32)


def hologram(seed: int, ndisks: int, nx: int, ny: int | None = None, zmin: float = 1e4,
             zmax: float = 1e5, wavlen: float = 0.6328, dx: float = 10.0, noise: float = 0.01,
             max_planes: int | None = 64, return_scene: bool = False):
    """Float32 hologram of shape (ny, nx) (the reference's column-major (nx, ny) layout)."""
    ny = nx if ny is None else ny
    scene = disk_scene(seed, ndisks, nx, ny, zmin, zmax, dx=dx)
    psi = sensor_field(scene, nx, ny, zmin, zmax, wavlen, dx, max_planes)
    holo = _record(psi, np.random.Generator(np.random.PCG64(seed + 0x9E3779B9)), noise)
    return (holo, scene) if return_scene else holo


def _propagate(psi, dist, kzc, evan):
    h = np.exp(2j * np.pi * dist * kzc)
    h[evan] = 0.0
    return _ifft2(_fft2(psi) * h)


def gs_pair(seed: int, ndisks: int, n: int, separation: float, zmin: float = 1e4,
            zmax: float = 1e5, wavlen: float = 0.6328, dx: float = 10.0, noise: float = 0.01,
            max_planes: int | None = 64):
    """Two holograms of one scene: I1 at the sensor and I2 at +separation um further from the
    object (I1 is the plane nearer the object, holofunc.jl:90-91)."""
    scene = disk_scene(seed, ndisks, n, n, zmin, zmax, dx=dx)
    psi1 = sensor_field(scene, n, n, zmin, zmax, wavlen, dx, max_planes)
    kzc, evan = _kz_cycles(n, n, wavlen, dx)
    psi2 = _propagate(psi1, separation, kzc, evan)
    rng = np.random.Generator(np.random.PCG64(seed + 0x9E3779B9))
    return _record(psi1, rng, noise), _record(psi2, rng, noise)


def _sensor_field_device(seed, ndisks, nx, ny, zmin, zmax, wavlen, dx, max_planes, device):
    """The sensor-plane field of ``disk_scene(seed, ...)`` formed on the GPU (complex64 tensor):
    disk masks applied on the device, the field carried between depth planes by libholo's
    angular-spectrum propagation (complex FP32, FP64-wrapped transfer phase). Returns
    (field, propagate) where propagate(field, dist) propagates in place."""
    import ctypes

    import torch

    from . import _lib

    x, y, z, diam = disk_scene(seed, ndisks, nx, ny, zmin, zmax, dx=dx)
    dev = torch.device(f"cuda:{device}")
    psi = torch.ones((ny, nx), dtype=torch.complex64, device=dev)
    lib, h = _lib.load(), _lib.handle(device)
    st = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)

    def prop(field, dist):
        t = _lib.holo_tf(None, float(dist), float(wavlen), float(dx), 1)
        _lib.check(lib.holo_asm_prop(h.ptr, st, nx, ny, ctypes.c_void_p(field.data_ptr()),
                                     ctypes.c_void_p(field.data_ptr()), ctypes.byref(t)))

    zcur = None
    for zp, members in _planes(z, zmin, zmax, max_planes):
        if zcur is not None:
            prop(psi, zcur - zp)
        mask = np.ones((ny, nx), dtype=np.float32)
        for i in members:
            _apply_disk(mask, x[i], y[i], diam[i] / 2.0, dx)
        psi *= torch.from_numpy(mask).to(dev)
        zcur = zp
    if zcur is not None:
        prop(psi, zcur)
    return psi, prop


def _record_device(psi, gen, noise):
    import torch

    inten = psi.real ** 2 + psi.imag ** 2
    inten = inten + noise * torch.randn(tuple(psi.shape), generator=gen, device=psi.device, dtype=torch.float32)
    return inten.clamp_min(0.0).contiguous()


def hologram_device(seed: int, ndisks: int, nx: int, ny: int | None = None, zmin: float = 1e4,
                    zmax: float = 1e5, wavlen: float = 0.6328, dx: float = 10.0, noise: float = 0.01,
                    max_planes: int = 16, device: int = 0):
    """The same scene model as ``hologram`` formed on the GPU: disk masks are applied on the
    device and the field is carried between depth planes by libholo's angular-spectrum
    propagation in complex FP32 (FP64-wrapped transfer phase). For benchmark inputs in bulk;
    the tests' parity cases use the FP64 ``hologram``. Returns a Float32 (ny, nx) device tensor."""
    import torch

    ny = nx if ny is None else ny
    psi, _ = _sensor_field_device(seed, ndisks, nx, ny, zmin, zmax, wavlen, dx, max_planes, device)
    gen = torch.Generator(device=psi.device)
    gen.manual_seed(int(seed) + 0x9E3779B9)
    return _record_device(psi, gen, noise)


def gs_pair_device(seed: int, ndisks: int, nx: int, ny: int | None = None, separation: float = 2000.0,
                   zmin: float = 1e4, zmax: float = 1e5, wavlen: float = 0.6328, dx: float = 10.0,
                   noise: float = 0.01, max_planes: int = 16, device: int = 0):
    """``gs_pair`` formed on the GPU: two recordings of one scene, I1 at the sensor and I2
    ``separation`` um further from the object (I1 is the plane nearer the object,
    holofunc.jl:90-91). Returns two Float32 (ny, nx) device tensors."""
    import torch

    ny = nx if ny is None else ny
    psi, prop = _sensor_field_device(seed, ndisks, nx, ny, zmin, zmax, wavlen, dx, max_planes, device)
    gen = torch.Generator(device=psi.device)
    gen.manual_seed(int(seed) + 0x9E3779B9)
    i1 = _record_device(psi, gen, noise)
    prop(psi, separation)
    return i1, _record_device(psi, gen, noise)

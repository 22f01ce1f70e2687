"""Particle neighbourhoods from a binary or thresholded volume: ParticleHolography.jl's
particle_bounding_boxes / particle_bounding_boxes_3d (particle_detection.jl:58-96) and
cu_dilate (particle_detection.jl:12-42) over libholo.

The reference labels every slice, copies each label image to the host and scans it for the
rectangles of its valid labels (ccl.jl:256-281). Here one libholo call labels all slices on
the device and returns only the rectangles (``holo_ccl_boxes``); the z-linking of rectangles
into neighbourhoods is the host logic of ccl.jl. A uint8 (N0f8) volume with a ``threshold``
is binarised inside the labelling kernel (``d_vol .<= threshold`` without a Bool volume).

Volumes are tensors of shape (slices, ny, nx): the reference's (nx, ny, slices) array.
"""
from __future__ import annotations

import warnings
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import check
from .ccl import link_arrays

__all__ = ["cu_dilate", "slice_bounding_rectangles", "particle_bounding_boxes", "particle_bounding_boxes_3d",
           "tamura", "depth_profile", "getcenterfromslice", "gaussian_kernel", "particle_coordinates"]


def _dev(t: torch.Tensor) -> int:
    if not t.is_cuda:
        raise AssertionError("arrays must be device (GPU) tensors")
    return t.device.index or 0


def cu_dilate(vol: torch.Tensor, blocksize: int = 32) -> torch.Tensor:
    """particle_detection.jl:34-42: 3 x 3 in-plane dilation of every slice of a Bool volume;
    border pixels are false. ``blocksize`` is accepted for signature parity."""
    if vol.dtype not in (torch.bool, torch.uint8) or vol.dim() != 3:
        raise AssertionError("cu_dilate takes a Bool volume of shape (slices, ny, nx)")
    v = vol.contiguous()
    device = _dev(v)
    out = torch.empty_like(v)
    count, ny, nx = v.shape
    check(_lib.load().holo_dilate(_lib.handle(device).ptr, torch.cuda.current_stream(device).cuda_stream,
                                  nx, ny, count, v.view(torch.uint8).data_ptr(), out.view(torch.uint8).data_ptr()))
    return out


def _slice_boxes(vol: torch.Tensor, threshold: Optional[float], max_boxes: int) -> Tuple[np.ndarray, np.ndarray]:
    """Rectangles (x_min, y_min, x_max, y_max) of every slice's 8-connected components in
    ascending label order, as int32 (R, 4) rows with int64 (slices + 1,) slice offsets. Only the
    used part of the device box table is copied to the host."""
    if vol.dim() != 3:
        raise AssertionError("the volume must have shape (slices, ny, nx)")
    v = vol.contiguous()
    device = _dev(v)
    count, ny, nx = v.shape
    if threshold is not None:
        if v.dtype != torch.uint8:
            raise AssertionError("a threshold applies to an N0f8 (uint8) volume")
        kind = _lib.HOLO_CCL_N0F8_LE
    elif v.dtype == torch.float32:
        kind = _lib.HOLO_CCL_F32
    elif v.dtype in (torch.bool, torch.uint8):
        kind = _lib.HOLO_CCL_BINARY
    else:
        raise AssertionError(f"unsupported volume type {v.dtype}")
    src = v.view(torch.uint8) if v.dtype == torch.bool else v
    lib, h = _lib.load(), _lib.handle(device)
    stream = torch.cuda.current_stream(device).cuda_stream
    while True:
        boxes = torch.empty((count, max_boxes, 5), dtype=torch.int32, device=v.device)
        nbox = torch.empty(count, dtype=torch.int32, device=v.device)
        check(lib.holo_ccl_boxes(h.ptr, stream, nx, ny, count, src.data_ptr(), kind,
                                 float(threshold) if threshold is not None else 0.0, max_boxes,
                                 boxes.data_ptr(), nbox.data_ptr()))
        n = nbox.cpu().numpy().astype(np.int64)
        if count == 0 or int(n.max()) <= max_boxes:
            break
        max_boxes = int(n.max())
    offsets = np.zeros(count + 1, np.int64)
    np.cumsum(n, out=offsets[1:])
    maxn = int(n.max()) if count else 0
    if maxn == 0:
        return np.zeros((0, 4), np.int32), offsets
    b = boxes[:, :maxn].cpu().numpy()
    valid = np.arange(maxn)[None, :] < n[:, None]
    key = b[..., 0].view(np.uint32).astype(np.int64)
    key[~valid] = 1 << 40  # unused slots sort last
    b = np.take_along_axis(b, np.argsort(key, axis=1, kind="stable")[..., None], axis=1)
    # box (i0, j0, i1, j1): tensor (row j, column i) is the reference's [y = i + 1, x = j + 1]
    rects = (b[valid][:, [2, 1, 4, 3]] + 1).astype(np.int32)
    return rects, offsets


def slice_bounding_rectangles(vol: torch.Tensor, threshold: Optional[float] = None,
                              max_boxes: int = 4096) -> List[List[Tuple[int, int, int, int]]]:
    """Per slice, the rectangles (x_min, y_min, x_max, y_max) of its 8-connected components in
    ascending label order (cu_find_valid_labels + get_bounding_rectangles of each slice).
    ``vol``: Bool/uint8 (nonzero is foreground), float32 (equal positive values connect), or
    uint8 N0f8 with ``threshold`` (foreground where v/255 <= threshold)."""
    rects, offsets = _slice_boxes(vol, threshold, max_boxes)
    return [[tuple(int(c) for c in r) for r in rects[offsets[z]:offsets[z + 1]]] for z in range(offsets.size - 1)]


def _bounding_boxes(vol, threshold, three_d):
    rects, offsets = _slice_boxes(vol, threshold, 4096)
    return link_arrays(rects, offsets, three_d)


def particle_bounding_boxes(d_bin_vol: torch.Tensor, threshold: Optional[float] = None):
    """particle_detection.jl:58-77: neighbourhoods {UUID: [x_min, y_min, z_min, x_max, y_max,
    z_max]} of a Bool volume (or of ``d_vol .<= threshold`` for an N0f8 volume); rectangles
    overlapping in x-y are linked along z."""
    return _bounding_boxes(d_bin_vol, threshold, False)


def particle_bounding_boxes_3d(d_bin_vol: torch.Tensor, threshold: Optional[float] = None):
    """particle_detection.jl:79-96: as particle_bounding_boxes, linking a rectangle only to
    neighbourhoods that reach the previous slice."""
    return _bounding_boxes(d_bin_vol, threshold, True)


def tamura(arr: np.ndarray) -> float:
    """particle_detection.jl:110-112: the Tamura coefficient std / mean of a slice (standard
    deviation with the n - 1 divisor). Evaluated in FP64."""
    a = np.asarray(arr, dtype=np.float64)
    return float(a.std(ddof=1) / a.mean())


def depth_profile(f: Callable, subvol: np.ndarray) -> np.ndarray:
    """particle_detection.jl:114-116: ``f`` of every slice of a (slices, rows, columns) block.
    The Tamura metric is evaluated for all slices at once (the same FP64 std / mean)."""
    if f is tamura:
        a = np.asarray(subvol, dtype=np.float64).reshape(subvol.shape[0], -1)
        with np.errstate(invalid="ignore", divide="ignore"), warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)  # one-pixel slices: std is NaN, as in Julia
            return a.std(axis=1, ddof=1) / a.mean(axis=1)
    return np.array([f(subvol[k]) for k in range(subvol.shape[0])], dtype=np.float64)


def getcenterfromslice(arr: np.ndarray) -> Tuple[float, float]:
    """particle_detection.jl:118-131: the intensity-weighted centre (x, y) of a slice, 1-based.
    ``arr`` is in the tensor layout (row r = the reference's x = r + 1, column c = its
    y = c + 1)."""
    a = np.asarray(arr, dtype=np.float64)
    total = a.sum()
    x = float((a.sum(axis=1) * np.arange(1, a.shape[0] + 1)).sum() / total)
    y = float((a.sum(axis=0) * np.arange(1, a.shape[1] + 1)).sum() / total)
    return x, y


def gaussian_kernel(sigma: float) -> np.ndarray:
    """The profile smoothing kernel of particle_coordinates, ImageFiltering.jl's
    ``Kernel.gaussian((sigma,))``: exp(-x^2 / (2 sigma^2)) for x = -w..w, w = 2 ceil(sigma),
    normalised to sum 1."""
    w = 2 * int(np.ceil(sigma))
    x = np.arange(-w, w + 1, dtype=np.float64)
    g = np.exp(-x * x / (2.0 * sigma * sigma))
    return g / g.sum()


def _smooth_replicate(v: np.ndarray, kernel: np.ndarray) -> np.ndarray:
    """Correlation with a centred odd kernel, borders replicated (ImageFiltering's default)."""
    w = kernel.size // 2
    padded = np.concatenate([np.full(w, v[0]), v, np.full(w, v[-1])])
    return np.correlate(padded, kernel, mode="valid")


def particle_coordinates(particle_bbs: Dict, d_vol: torch.Tensor, depth_metrics: Callable = tamura,
                         profile_smoothing_kernel: Optional[np.ndarray] = None) -> Dict:
    """particle_detection.jl:148-160: per neighbourhood [x_min, y_min, z_min, x_max, y_max, z_max]
    of ``particle_bbs``, the depth is the first maximum of the smoothed ``depth_metrics`` profile
    over the block's slices and (x, y) the intensity-weighted centre of that slice, all 1-based
    volume coordinates as float32. ``d_vol`` is an N0f8 volume of shape (slices, ny, nx); each
    block is copied to the host (it is a few hundred kB at most)."""
    if d_vol.dtype != torch.uint8 or d_vol.dim() != 3:
        raise AssertionError("particle_coordinates takes an N0f8 (uint8) volume of shape (slices, ny, nx)")
    kernel = gaussian_kernel(5.0) if profile_smoothing_kernel is None else np.asarray(profile_smoothing_kernel, np.float64)
    out = {}
    for key, (x0, y0, z0, x1, y1, z1) in particle_bbs.items():
        # the reference's d_vol[y0:y1, x0:x1, z0:z1]: tensor [z, x - 1, y - 1]
        block = d_vol[z0 - 1:z1, x0 - 1:x1, y0 - 1:y1].cpu().numpy().astype(np.float32) / np.float32(255)
        profile = _smooth_replicate(depth_profile(depth_metrics, block), kernel).astype(np.float32)
        z = int(np.argmax(profile)) + 1
        x, y = getcenterfromslice(block[z - 1])
        out[key] = [np.float32(x + x0 - 1), np.float32(y + y0 - 1), np.float32(z + z0 - 1)]
    return out

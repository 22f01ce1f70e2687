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

from typing import List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import check
from .ccl import link_slices

__all__ = ["cu_dilate", "slice_bounding_rectangles", "particle_bounding_boxes", "particle_bounding_boxes_3d"]


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


def slice_bounding_rectangles(vol: torch.Tensor, threshold: Optional[float] = None,
                              max_boxes: int = 4096) -> List[List[Tuple[int, int, int, int]]]:
    """Per slice, the rectangles (x_min, y_min, x_max, y_max) of its 8-connected components in
    ascending label order (cu_find_valid_labels + get_bounding_rectangles of each slice).
    ``vol``: Bool/uint8 (nonzero is foreground), float32 (equal positive values connect), or
    uint8 N0f8 with ``threshold`` (foreground where v/255 <= threshold)."""
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
        n = nbox.cpu().numpy()
        if count == 0 or int(n.max()) <= max_boxes:
            break
        max_boxes = int(n.max())
    b = boxes.cpu().numpy()
    out = []
    for z in range(count):
        rows = b[z, :n[z]]
        rows = rows[np.argsort(rows[:, 0].view(np.uint32), kind="stable")]
        # tensor (row j, column i) is the reference's [y = i + 1, x = j + 1]
        out.append([(int(r[2]) + 1, int(r[1]) + 1, int(r[4]) + 1, int(r[3]) + 1) for r in rows])
    return out


def _bounding_boxes(vol, threshold, three_d):
    return link_slices(slice_bounding_rectangles(vol, threshold), three_d)


def particle_bounding_boxes(d_bin_vol: torch.Tensor, threshold: Optional[float] = None):
    """particle_detection.jl:58-77: neighbourhoods {UUID: [x_min, y_min, z_min, x_max, y_max,
    z_max]} of a Bool volume (or of ``d_vol .<= threshold`` for an N0f8 volume); rectangles
    overlapping in x-y are linked along z."""
    return _bounding_boxes(d_bin_vol, threshold, False)


def particle_bounding_boxes_3d(d_bin_vol: torch.Tensor, threshold: Optional[float] = None):
    """particle_detection.jl:79-96: as particle_bounding_boxes, linking a rectangle only to
    neighbourhoods that reach the previous slice."""
    return _bounding_boxes(d_bin_vol, threshold, True)

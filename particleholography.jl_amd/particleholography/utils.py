"""Ingest of 8-bit holograms: ParticleHolography.jl's load_gray2float / load_grayimg
(utils.jl:23-41), make_background (utils.jl:217-240), cu_make_background_mode
(utils.jl:243-317) and the tutorial's background removal ``img .- bkg .+ 64/255``
(preprocessings.md:21-23).

Layout. The reference's ``load`` returns an (H, W) Julia matrix whose element [r, c] is
raster row r, column c; in memory (column-major) each raster column is contiguous. The
mirror keeps that memory layout, like every other array here (a Julia array of size
(nx, ny) is a tensor of shape (ny, nx)): ``load_gray2float`` returns the raster transposed,
shape (W, H), element [c, r] = raster row r, column c. Physical axes therefore meet the
filters and transfers exactly as in the reference.

Decoding is host IO (FileIO / ImageMagick in the reference, Pillow here); colour images are
reduced to gray with the ITU-R 601 luma weights and rounded to 8 bits, which is exact for
gray images (R = G = B, as the reference's test images are). The per-pixel work runs on the
device through libholo: ``cu_ingest_u8`` uploads 1 byte per pixel and converts (with the
background removal fused), ``make_background`` / ``cu_make_background_mode`` reduce a
resident uint8 stack with one read of every byte.
"""
from __future__ import annotations

import json
import math
import uuid
from typing import Dict, List, Optional, Sequence, Union

import numpy as np
import torch

from . import _lib
from ._lib import check

__all__ = ["load_grayimg", "load_gray2float", "make_background", "cu_make_background_mode",
           "cu_make_background", "cu_ingest_u8", "dictsave", "dictload"]

BACKGROUND_OFFSET = 64 / 255  # preprocessings.md:23


def _decode_raster_u8(path: str) -> np.ndarray:
    """8-bit gray raster (H, W) of an image file."""
    try:
        from PIL import Image
    except ImportError as e:  # pragma: no cover - Pillow ships with the image
        raise ImportError("load_gray2float needs Pillow to decode image files") from e
    with Image.open(path) as im:
        if im.mode in ("L", "P", "1", "LA", "PA"):
            return np.asarray(im.convert("L"), dtype=np.uint8).copy()
        if im.mode in ("I;16", "I;16B", "I;16L", "I"):
            raise ValueError(f"{path}: 16-bit images are not supported")
        rgb = np.asarray(im.convert("RGB"), dtype=np.uint8)
    if np.array_equal(rgb[..., 0], rgb[..., 1]) and np.array_equal(rgb[..., 1], rgb[..., 2]):
        return rgb[..., 0].copy()
    c = rgb.astype(np.float32) / np.float32(255)
    y = np.float32(0.299) * c[..., 0] + np.float32(0.587) * c[..., 1] + np.float32(0.114) * c[..., 2]
    return np.clip(np.rint(y * np.float32(255)), 0, 255).astype(np.uint8)


def load_grayimg(path: str) -> np.ndarray:
    """utils.jl:38-41: the N0f8 gray image as raw uint8, reference layout (W, H)."""
    return np.ascontiguousarray(_decode_raster_u8(path).T)


def load_gray2float(path: str) -> np.ndarray:
    """utils.jl:23-25: Float32 gray values v/255, reference layout (W, H)."""
    return load_grayimg(path).astype(np.float32) / np.float32(255)


def _n0f8_bytes(img) -> np.ndarray:
    """An image of the list as raw uint8: uint8 as is, anything else as
    round(UInt8, clamp(Float32(img) * 255, 0, 255)) (utils.jl:298-302, ties to even)."""
    if isinstance(img, torch.Tensor):
        img = img.detach().cpu().numpy()
    a = np.asarray(img)
    if a.dtype == np.uint8:
        return a
    return np.clip(np.rint(a.astype(np.float32) * np.float32(255)), 0, 255).astype(np.uint8)


def _upload_stack(images: Sequence, device: int) -> torch.Tensor:
    """Stacks host images into pinned memory and copies them to the device as (count, ny, nx)."""
    first = _n0f8_bytes(images[0])
    shape = first.shape
    pinned = torch.empty((len(images),) + tuple(shape), dtype=torch.uint8, pin_memory=True)
    view = pinned.numpy()
    for k, img in enumerate(images):
        b = first if k == 0 else _n0f8_bytes(img)
        assert b.shape == shape, f"image {k} has size {b.shape}, expected {shape}"
        view[k] = b
    return pinned.to(f"cuda:{device}", non_blocking=True)


def cu_make_background(stack: torch.Tensor, mode: str = "mode") -> torch.Tensor:
    """Per-pixel background (Float64, (ny, nx), on the stack's device) of a device uint8 stack
    (count, ny, nx): the most frequent value (smallest on ties) / 255 for ``mode="mode"``,
    the FP64 mean of v/255 for ``mode="mean"``."""
    kind = {"mode": _lib.HOLO_BACKGROUND_MODE, ":mode": _lib.HOLO_BACKGROUND_MODE,
            "mean": _lib.HOLO_BACKGROUND_MEAN, ":mean": _lib.HOLO_BACKGROUND_MEAN}.get(str(mode))
    if kind is None:
        raise ValueError(f"mode must be :mode or :mean, got {mode!r}")
    if stack.dtype != torch.uint8 or stack.dim() != 3:
        raise AssertionError("the stack must be a uint8 tensor of shape (count, ny, nx)")
    if not stack.is_cuda:
        raise AssertionError("arrays must be device (GPU) tensors")
    stack = stack.contiguous()
    count, ny, nx = stack.shape
    device = stack.device.index or 0
    out = torch.empty((ny, nx), dtype=torch.float64, device=stack.device)
    stream = torch.cuda.current_stream(device).cuda_stream
    check(_lib.load().holo_make_background(_lib.handle(device).ptr, stream, nx, ny, count,
                                           stack.data_ptr(), kind, out.data_ptr()))
    return out


def cu_make_background_mode(grayimglist: Sequence, device: int = 0) -> np.ndarray:
    """utils.jl:286-317: per-pixel mode of the images (uint8 or N0f8-valued floats) / 255,
    returned as a host Float64 array like the reference."""
    if len(grayimglist) == 0:
        raise AssertionError("the image list is empty")
    stack = _upload_stack(grayimglist, device)
    return cu_make_background(stack, "mode").cpu().numpy()


def make_background(pathlist: List[str], mode: str = "mode", device: int = 0) -> np.ndarray:
    """utils.jl:217-240: background of the hologram files by per-pixel mode (default) or mean,
    as a host Float64 array in the reference layout."""
    if len(pathlist) == 0:
        raise AssertionError("the path list is empty")
    stack = _upload_stack([load_grayimg(p) for p in pathlist], device)
    return cu_make_background(stack, mode).cpu().numpy()


def cu_ingest_u8(images: Union[torch.Tensor, np.ndarray], bkg=None, offset: Optional[float] = None,
                 device: int = 0) -> torch.Tensor:
    """Device Float32 holograms from raw 8-bit images ((ny, nx) or (count, ny, nx), host or
    device): v/255 as load_gray2float, or with a background ``bkg`` (Float64 (ny, nx)) the
    tutorial's Float32((v/255 - bkg) + offset), offset 64/255 unless given."""
    if isinstance(images, np.ndarray):
        images = torch.from_numpy(np.ascontiguousarray(images))
    if images.dtype != torch.uint8:
        raise AssertionError("images must be uint8 (raw N0f8 values)")
    squeeze = images.dim() == 2
    stack = images[None] if squeeze else images
    if not stack.is_cuda:
        pinned = torch.empty(stack.shape, dtype=torch.uint8, pin_memory=True)
        pinned.copy_(stack)
        stack = pinned.to(f"cuda:{device}", non_blocking=True)
    stack = stack.contiguous()
    count, ny, nx = stack.shape
    dev = stack.device.index or 0
    b = None
    if bkg is not None:
        b = torch.as_tensor(bkg, dtype=torch.float64)
        b = b.to(stack.device).contiguous()
        assert tuple(b.shape) == (ny, nx), "background and images must have the same size"
        if offset is None:
            offset = BACKGROUND_OFFSET
    off = 0.0 if offset is None else float(offset)
    out = torch.empty((count, ny, nx), dtype=torch.float32, device=stack.device)
    stream = torch.cuda.current_stream(dev).cuda_stream
    check(_lib.load().holo_ingest_u8(_lib.handle(dev).ptr, stream, nx, ny, count, stack.data_ptr(),
                                     b.data_ptr() if b is not None else None, off, out.data_ptr()))
    return out[0] if squeeze else out


def _json_f32(v) -> str:
    f = np.float32(v)
    return str(f) if math.isfinite(f) else "null"


def dictsave(filename: str, d: Dict) -> None:
    """utils.jl:340-345: a particle dictionary {UUID: [Float32, ...]} as one JSON object, keys the
    UUID strings, values printed as the shortest Float32 decimals (as Julia prints Float32)."""
    items = (f"{json.dumps(str(k))}:[{','.join(_json_f32(v) for v in vals)}]" for k, vals in d.items())
    with open(filename, "w") as io:
        io.write("{" + ",".join(items) + "}")


def dictload(filename: str) -> Dict[uuid.UUID, List[np.float32]]:
    """utils.jl:358-365: the dictionary written by dictsave, UUID keys, Float32 values."""
    with open(filename) as io:
        data = json.load(io)
    return {uuid.UUID(k): [np.float32(v) for v in vals] for k, vals in data.items()}

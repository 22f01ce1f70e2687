"""Batched reconstruction of many holograms of uneven sizes, and the tile autotuner.

``reconst_batch`` runs many ``cu_get_reconst_xyprojection`` calls (holofunc.jl:248-266)
together through ``holo_reconst_batch``: holograms are grouped by size, and every round of
slices of a group is one column-pass launch and one row-pass launch over all of its holograms.
Each projection equals the one the single-hologram call returns.

``autotune`` times candidate tile configurations of both reconstruction passes on the device
for one size and keeps the fastest in the library handle; results are cached per device name
in a JSON file (``$PARTICLEHOLOGRAPHY_TUNING`` or ``~/.cache/particleholography/tuning.json``)
and re-applied by ``load_tuning``.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass
from pathlib import Path
from typing import List, Optional

import torch

from . import _lib
from ._lib import check, holo_job, holo_tf


@dataclass
class ReconJob:
    """One hologram to reconstruct: Float32 intensity (ny, nx) or ComplexF32 wavefront."""
    holo: torch.Tensor
    z_front: float
    z_dz: float
    slices: int
    wavlen: float = 0.6328
    dx: float = 10.0
    proj: Optional[torch.Tensor] = None


def reconst_batch(jobs: List[ReconJob], device: int = 0, handle=None) -> List[torch.Tensor]:
    """xy min-projections of every job (Float32 (ny, nx) tensors), computed together on the
    current stream (``handle``: a libholo handle other than the thread's default)."""
    arr = (holo_job * max(1, len(jobs)))()
    keep = []
    for i, j in enumerate(jobs):
        src = j.holo.contiguous()
        if src.dtype not in (torch.float32, torch.complex64):
            raise AssertionError(f"job {i}: hologram must be float32 or complex64, got {src.dtype}")
        if not src.is_cuda:
            raise AssertionError(f"job {i}: arrays must be device (GPU) tensors")
        ny, nx = src.shape
        if j.proj is None:
            j.proj = torch.empty((ny, nx), dtype=torch.float32, device=src.device)
        assert tuple(j.proj.shape) == (ny, nx) and j.proj.dtype == torch.float32, f"job {i}: bad projection"
        keep.append(src)
        arr[i].nx, arr[i].ny = nx, ny
        arr[i].src = src.data_ptr()
        arr[i].src_real = 1 if src.dtype == torch.float32 else 0
        arr[i].front = holo_tf(None, float(j.z_front), float(j.wavlen), float(j.dx), 1)
        arr[i].dz = holo_tf(None, float(j.z_dz), float(j.wavlen), float(j.dx), 1)
        arr[i].slices = int(j.slices)
        arr[i].proj = j.proj.data_ptr()
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    hd = handle if handle is not None else _lib.handle(device)
    check(_lib.load().holo_reconst_batch(hd.ptr, stream, arr, len(jobs)))
    return [j.proj for j in jobs]


def _cache_path() -> Path:
    p = os.environ.get("PARTICLEHOLOGRAPHY_TUNING")
    return Path(p) if p else Path.home() / ".cache" / "particleholography" / "tuning.json"


def _device_key(device: int) -> str:
    return torch.cuda.get_device_name(device)


FIELDS = ("batch", "spg_gen", "spg_out", "cols_per_wg", "rows_per_wg")
VARIANTS = ("col_db", "row_pf")


def _variants(nx: int, ny: int, device: int, set_to: Optional[dict] = None) -> dict:
    cd, rp = ctypes.c_int(0), ctypes.c_int(0)
    if set_to is not None:
        cd.value, rp.value = int(set_to.get("col_db", 0)), int(set_to.get("row_pf", 0))
    check(_lib.load().holo_size_variants(_lib.handle(device).ptr, nx, ny, ctypes.byref(cd), ctypes.byref(rp),
                                         1 if set_to is not None else 0))
    return {"col_db": cd.value, "row_pf": rp.value}


def autotune(nx: int, ny: int, wavlen: float = 0.6328, dx: float = 10.0, device: int = 0,
             save: bool = True) -> dict:
    """Times candidate tiles for an (nx, ny) reconstruction and keeps the fastest."""
    out = (ctypes.c_int * 5)()
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    check(_lib.load().holo_autotune(_lib.handle(device).ptr, stream, nx, ny, float(wavlen), float(dx), out))
    choice = dict(zip(FIELDS, list(out)))
    choice.update(_variants(nx, ny, device))
    if save:
        path = _cache_path()
        data = json.loads(path.read_text()) if path.exists() else {}
        data.setdefault(_device_key(device), {})[f"{nx}x{ny}"] = choice
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(data, indent=1, sort_keys=True))
    return choice


def set_tuning_for_size(nx: int, ny: int, choice: dict, device: int = 0) -> None:
    """Sets a size's tiles; its kernel variants too when ``choice`` names them."""
    vals = [int(choice.get(f, 0)) for f in FIELDS]
    check(_lib.load().holo_set_tuning_for_size(_lib.handle(device).ptr, nx, ny, *vals))
    if any(v in choice for v in VARIANTS):
        cur = _variants(nx, ny, device)
        _variants(nx, ny, device, {v: choice.get(v, cur[v]) for v in VARIANTS})


def get_tuning_for_size(nx: int, ny: int, slices: int = 0, device: int = 0) -> dict:
    out = (ctypes.c_int * 5)()
    check(_lib.load().holo_get_tuning_for_size(_lib.handle(device).ptr, nx, ny, int(slices), out))
    choice = dict(zip(FIELDS, list(out)))
    choice.update(_variants(nx, ny, device))
    return choice


def load_tuning(device: int = 0) -> int:
    """Applies cached autotuner choices for this device; returns how many sizes were set."""
    path = _cache_path()
    if not path.exists():
        return 0
    entries = json.loads(path.read_text()).get(_device_key(device), {})
    for size, choice in entries.items():
        nx, ny = (int(v) for v in size.split("x"))
        set_tuning_for_size(nx, ny, choice, device)
    return len(entries)


@dataclass
class GSJob:
    """One Gerchberg-Saxton pair: intensities at two planes separated by +z (holo1 nearer the
    object), ``iters`` iterations. Transfers are generated from (z, wavlen, dx) unless
    ``fwd`` / ``bwd`` (centred ComplexF32 arrays, as from cu_transfer) are given."""
    holo1: torch.Tensor
    holo2: torch.Tensor
    z: float
    iters: int
    wavlen: float = 0.6328
    dx: float = 10.0
    fwd: Optional[torch.Tensor] = None
    bwd: Optional[torch.Tensor] = None
    out: Optional[torch.Tensor] = None


def phase_retrieval_batch(jobs: List[GSJob], device: int = 0, handle=None) -> List[torch.Tensor]:
    """Retrieved wavefronts at holo1's plane (ComplexF32 (ny, nx)), computed together on the
    current stream (``handle``: a libholo handle other than the thread's default)."""
    from ._lib import holo_gs_job

    arr = (holo_gs_job * max(1, len(jobs)))()
    keep = []
    for i, j in enumerate(jobs):
        h1, h2 = j.holo1.contiguous(), j.holo2.contiguous()
        if h1.dtype != torch.float32 or h2.dtype != torch.float32 or h1.shape != h2.shape:
            raise AssertionError(f"job {i}: holograms must be float32 tensors of one shape")
        if not h1.is_cuda:
            raise AssertionError(f"job {i}: arrays must be device (GPU) tensors")
        ny, nx = h1.shape
        if j.out is None:
            j.out = torch.empty((ny, nx), dtype=torch.complex64, device=h1.device)
        keep += [h1, h2]
        arr[i].nx, arr[i].ny = nx, ny
        arr[i].holo1, arr[i].holo2 = h1.data_ptr(), h2.data_ptr()
        if j.fwd is not None:
            for name, a in (("fwd", j.fwd), ("bwd", j.bwd)):
                if not isinstance(a, torch.Tensor) or not a.is_cuda or a.device != h1.device:
                    raise AssertionError(f"job {i}: transfer {name} must be a tensor on {h1.device}")
                if a.dtype != torch.complex64 or tuple(a.shape) != (ny, nx):
                    raise AssertionError(f"job {i}: transfer {name} must be complex64 of shape {(ny, nx)}")
            f, b = j.fwd.contiguous(), j.bwd.contiguous()
            keep += [f, b]
            arr[i].fwd = holo_tf(f.data_ptr(), 0.0, 0.0, 0.0, 0)
            arr[i].bwd = holo_tf(b.data_ptr(), 0.0, 0.0, 0.0, 0)
        else:
            arr[i].fwd = holo_tf(None, float(j.z), float(j.wavlen), float(j.dx), 1)
            arr[i].bwd = holo_tf(None, -float(j.z), float(j.wavlen), float(j.dx), 1)
        arr[i].iters = int(j.iters)
        arr[i].out = j.out.data_ptr()
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    hd = handle if handle is not None else _lib.handle(device)
    check(_lib.load().holo_phase_retrieval_batch(hd.ptr, stream, arr, len(jobs)))
    return [j.out for j in jobs]

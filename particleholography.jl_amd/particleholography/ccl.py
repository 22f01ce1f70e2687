"""Connected components and particle neighbourhoods: ParticleHolography.jl's ccl.jl over libholo.

Device work (libholo): ``cu_connected_component_labeling`` (ccl.jl:186-218),
``cu_find_valid_labels`` and ``get_bounding_rectangles`` (ccl.jl:240-264) through the box
kernels. The label of a pixel is 1 + the smallest memory offset of its 8-connected component
of equal positive values, 0 for background -- the labels the reference's union-find leaves.

Host work, as in the reference: the z-linking of per-slice rectangles into particle
neighbourhoods (ccl.jl:266-466). Neighbourhoods are a dict from UUID to
``[x_min, y_min, z_min, x_max, y_max, z_max]`` (1-based, x = Julia dim 2, y = Julia dim 1).
Julia's Dict iterates in hash order of the UUID keys; this dict iterates in insertion order,
and entries deleted while a loop runs are skipped as in Julia. The results agree wherever
the reference's merging does not depend on iteration order.

Layout: a Julia (height, width) label image is a tensor of shape (width, height) here (same
memory), so a rectangle's x comes from the tensor's row index and y from its column index.
"""
from __future__ import annotations

import ctypes
import math
import random
import uuid
from typing import Dict, List, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import check

__all__ = ["cu_connected_component_labeling", "count_labels", "cu_find_valid_labels",
           "get_bounding_rectangles", "gen_particle_neighborhoods", "update_particle_neighborhoods_",
           "update_particle_neighborhoods3d_", "finalize_particle_neighborhoods_", "judge_overlap2d",
           "judge_overlap3d", "new_rect", "delete_duplicates_", "link_slices", "link_arrays"]

_TYPEMAX, _TYPEMIN = 2 ** 63 - 1, -(2 ** 63)


def _dev(t: torch.Tensor) -> int:
    if not t.is_cuda:
        raise AssertionError("arrays must be device (GPU) tensors")
    return t.device.index or 0


def _kind_of(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return _lib.HOLO_CCL_F32
    if t.dtype in (torch.bool, torch.uint8):
        return _lib.HOLO_CCL_BINARY
    raise AssertionError(f"images must be float32, bool or uint8, got {t.dtype}")


def cu_connected_component_labeling(input_img: torch.Tensor) -> torch.Tensor:
    """ccl.jl:202-227: 8-way labels (uint32, same shape) of a 2-D image (float32: equal
    positive values connect; bool/uint8: nonzero is foreground). A 3-D tensor is labelled
    slice by slice."""
    assert input_img.numel() <= 2 ** 32 - 1, "Image is too large. Maximum length is 2^32-1."
    img = input_img.contiguous()
    device = _dev(img)
    squeeze = img.dim() == 2
    stack = img[None] if squeeze else img
    count, ny, nx = stack.shape
    labels = torch.empty(stack.shape, dtype=torch.uint32, device=img.device)
    kind = _kind_of(stack)
    src = stack.view(torch.uint8) if stack.dtype == torch.bool else stack
    check(_lib.load().holo_ccl_label(_lib.handle(device).ptr, torch.cuda.current_stream(device).cuda_stream,
                                     nx, ny, count, src.data_ptr(), kind, 0.0, labels.data_ptr()))
    return labels[0] if squeeze else labels


def count_labels(labels) -> int:
    """ccl.jl:229-237: number of entries whose label is their own 1-based linear index."""
    a = _host_u32(labels).reshape(-1)
    return int(np.count_nonzero(a == np.arange(1, a.size + 1, dtype=np.uint64)))


def _host_u32(labels) -> np.ndarray:
    if isinstance(labels, torch.Tensor):
        labels = labels.detach().cpu().numpy()
    return np.asarray(labels).astype(np.uint64)


def _boxes_of_labels(labels: torch.Tensor, max_boxes: int = 1 << 16) -> List[Tuple[int, int, int, int, int]]:
    """(label, i0, j0, i1, j1) of every component of a device label image, by label."""
    lab = labels.contiguous()
    if lab.dtype != torch.uint32:
        lab = lab.to(torch.int64).to(torch.uint32)
    device = _dev(lab)
    ny, nx = lab.shape
    while True:
        boxes = torch.empty((max_boxes, 5), dtype=torch.int32, device=lab.device)
        nbox = torch.empty(1, dtype=torch.int32, device=lab.device)
        check(_lib.load().holo_label_boxes(_lib.handle(device).ptr, torch.cuda.current_stream(device).cuda_stream,
                                           nx, ny, 1, lab.data_ptr(), max_boxes, boxes.data_ptr(),
                                           nbox.data_ptr()))
        n = int(nbox.item())
        if n <= max_boxes:
            break
        max_boxes = n
    b = boxes[:n].cpu().numpy()
    b = b[np.argsort(b[:, 0].view(np.uint32), kind="stable")]
    return [(int(np.uint32(r[0])), int(r[1]), int(r[2]), int(r[3]), int(r[4])) for r in b]


def cu_find_valid_labels(labels: torch.Tensor) -> List[int]:
    """ccl.jl:250-254: the 1-based linear indices whose label equals the index (one per
    component), ascending."""
    return [b[0] for b in _boxes_of_labels(labels)]


def get_bounding_rectangles(labels, valid_labels: Sequence[int]) -> List[Tuple[int, int, int, int]]:
    """ccl.jl:256-281: (x_min, y_min, x_max, y_max) of each label of ``valid_labels``, in that
    order, 1-based with x the Julia column (tensor row) and y the Julia row (tensor column);
    a label without pixels gives (typemax, typemax, typemin, typemin)."""
    lab = labels if isinstance(labels, torch.Tensor) else torch.from_numpy(np.asarray(labels, np.uint32))
    if not lab.is_cuda:
        lab = lab.to("cuda")
    found = {b[0]: b for b in _boxes_of_labels(lab)}
    out = []
    for l in valid_labels:
        b = found.get(int(l))
        if b is None:
            out.append((_TYPEMAX, _TYPEMAX, _TYPEMIN, _TYPEMIN))
        else:
            _, i0, j0, i1, j1 = b
            out.append((j0 + 1, i0 + 1, j1 + 1, i1 + 1))
    return out


# ------------------------------------------------------------------------------------------
# particle neighbourhoods (ccl.jl:266-466), host logic
# ------------------------------------------------------------------------------------------
_KEY_DRAWS: List[Tuple[int, int]] = []  # (node, clock_seq) draws of random.Random(1234), in order
_KEY_RNG = random.Random(1234)


def _key_draw(k: int) -> Tuple[int, int]:
    while len(_KEY_DRAWS) <= k:
        _KEY_DRAWS.append((_KEY_RNG.getrandbits(48) | (1 << 40), _KEY_RNG.getrandbits(14)))
    return _KEY_DRAWS[k]


def _uuid_source():
    """A key source replaying one seeded sequence: the reference seeds a fresh
    MersenneTwister(1234) per call, so every source starts with the same draws (cached here,
    seeding costs more than the key)."""
    count = iter(range(1 << 62))

    def new_key():
        node, seq = _key_draw(next(count))
        return uuid.uuid1(node=node, clock_seq=seq)
    return new_key


def gen_particle_neighborhoods(bounding_rectangles, slicenum: int) -> Dict[uuid.UUID, List[int]]:
    """ccl.jl:266-277."""
    new_key = _uuid_source()
    out: Dict[uuid.UUID, List[int]] = {}
    for x_min, y_min, x_max, y_max in bounding_rectangles:
        out[new_key()] = [x_min, y_min, slicenum, x_max, y_max, slicenum]
    return out


def judge_overlap2d(rect1, rect2) -> bool:
    """ccl.jl:279-288 (strict inequalities)."""
    x_min1, y_min1, x_max1, y_max1 = rect1
    x_min2, y_min2, x_max2, y_max2 = rect2
    return x_min1 < x_max2 and x_max1 > x_min2 and y_min1 < y_max2 and y_max1 > y_min2


def judge_overlap3d(rect1, rect2) -> bool:
    """ccl.jl:290-299."""
    x_min1, y_min1, z_min1, x_max1, y_max1, z_max1 = rect1
    x_min2, y_min2, z_min2, x_max2, y_max2, z_max2 = rect2
    return (x_min1 < x_max2 and x_max1 > x_min2 and y_min1 < y_max2 and y_max1 > y_min2
            and z_min1 < z_max2 and z_max1 > z_min2)


def new_rect(rect1, rect2) -> List[int]:
    """ccl.jl:301-312."""
    x_min1, y_min1, x_max1, y_max1 = rect1
    x_min2, y_min2, x_max2, y_max2 = rect2
    return [min(x_min1, x_min2), min(y_min1, y_min2), max(x_max1, x_max2), max(y_max1, y_max2)]


class _Table:
    """The neighbourhoods of a dict as an (n, 6) int64 array in dict order, with appends and
    deletions, written back to the dict (values updated in place) by ``store``."""

    def __init__(self, pn):
        self.pn = pn
        self.keys = list(pn)
        n = len(self.keys)
        self.a = np.zeros((max(16, 2 * n), 6), np.int64)
        if n:
            self.a[:n] = np.array([pn[k] for k in self.keys], np.int64).reshape(n, 6)
        self.n = n
        self.dead = np.zeros(self.a.shape[0], bool)

    def append(self, key, row):
        if self.n == self.a.shape[0]:
            self.a = np.concatenate([self.a, np.zeros_like(self.a)])
            self.dead = np.concatenate([self.dead, np.zeros_like(self.dead)])
        self.a[self.n] = row
        self.keys.append(key)
        self.n += 1

    def store(self):
        pn = self.pn
        for idx, k in enumerate(self.keys):
            if self.dead[idx]:
                pn.pop(k, None)
            elif k in pn:
                pn[k][:] = [int(v) for v in self.a[idx]]
            else:
                pn[k] = [int(v) for v in self.a[idx]]


def _overlap2d_rows(br, a):
    """judge_overlap2d(br, (x_min, y_min, x_max, y_max) of each row)."""
    return (br[0] < a[:, 3]) & (br[2] > a[:, 0]) & (br[1] < a[:, 4]) & (br[3] > a[:, 1])


def _drop_unit_cubes(t: _Table) -> None:
    a = t.a[:t.n]
    unit = (np.abs(a[:, 0] - a[:, 3]) == 1) & (np.abs(a[:, 1] - a[:, 4]) == 1) & (np.abs(a[:, 2] - a[:, 5]) == 1)
    t.dead[:t.n] |= unit


def _update_table(t: _Table, bounding_rectangles, slicenum: int, three_d: bool, new_key) -> None:
    """One call of update_particle_neighborhoods! (three_d False) or
    update_particle_neighborhoods3d! (True) on a table."""
    s = int(slicenum)
    if three_d:
        sub = np.nonzero((t.a[:t.n, 5] == s - 1) & ~t.dead[:t.n])[0]
    for br in bounding_rectangles:
        if three_d:
            rows = t.a[sub]
            m = _overlap2d_rows(br, rows)
            if m.any():
                r = rows[m]
                r[:, 0] = np.minimum(r[:, 0], br[0])
                r[:, 1] = np.minimum(r[:, 1], br[1])
                r[:, 3] = np.maximum(r[:, 3], br[2])
                r[:, 4] = np.maximum(r[:, 4], br[3])
                r[:, 5] = s
                t.a[sub[m]] = r
                continue
        else:
            a = t.a[:t.n]
            m = _overlap2d_rows(br, a) & ~t.dead[:t.n]
            if m.any():
                rows = a[m]
                rows[:, 0] = np.minimum(rows[:, 0], br[0])
                rows[:, 1] = np.minimum(rows[:, 1], br[1])
                rows[:, 3] = np.maximum(rows[:, 3], br[2])
                rows[:, 4] = np.maximum(rows[:, 4], br[3])
                lt = s < rows[:, 2]
                gt = ~lt & (s > rows[:, 5])
                rows[lt, 2] = s
                rows[gt, 5] = s
                a[m] = rows
                continue
        t.append(new_key(), [br[0], br[1], s, br[2], br[3], s])
    _drop_unit_cubes(t)


def update_particle_neighborhoods_(particle_neighborhoods, bounding_rectangles, slicenum: int) -> None:
    """ccl.jl:328-363 (update_particle_neighborhoods!): a rectangle overlapping a neighbourhood
    in x-y grows it and extends its z-range; otherwise it starts a new neighbourhood. The
    rectangles are taken in order, each against every neighbourhood present at that moment
    (those started by earlier rectangles included); per rectangle the neighbourhoods do not
    interact, so that scan is one vector operation."""
    t = _Table(particle_neighborhoods)
    _update_table(t, bounding_rectangles, slicenum, False, _uuid_source())
    t.store()


def update_particle_neighborhoods3d_(particle_neighborhoods, bounding_rectangles, slicenum: int) -> None:
    """ccl.jl:378-409 (update_particle_neighborhoods3d!): only neighbourhoods that end on the
    previous slice when the call starts are extended."""
    t = _Table(particle_neighborhoods)
    _update_table(t, bounding_rectangles, slicenum, True, _uuid_source())
    t.store()


def link_arrays(rects: np.ndarray, offsets: np.ndarray, three_d: bool = False) -> Dict[uuid.UUID, List[int]]:
    """gen_particle_neighborhoods on slice 1, then one update per further slice, then
    finalize_particle_neighborhoods! (particle_detection.jl:58-96), in libholo's native
    linking (``holo_link_rectangles``). ``rects``: int32 (R, 4) rows (x_min, y_min, x_max,
    y_max) of every slice in order; ``offsets``: int64 (slices + 1,), slice s (1-based) holding
    rows offsets[s-1]:offsets[s]. Keys: one UUID source per start slice, as the reference draws a
    fresh MersenneTwister(1234) per call; the dict holds the survivors in start order."""
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    nslices = offsets.size - 1
    if nslices <= 0:
        return {}
    rects = np.ascontiguousarray(rects, dtype=np.int32).reshape(-1, 4)
    total = int(offsets[-1])
    if rects.shape[0] < total:
        raise AssertionError("fewer rectangles than offsets[-1]")
    cap = max(total, 1)
    rows = np.empty((cap, 6), np.int64)
    born = np.empty(cap, np.int32)
    alive = np.empty(cap, np.uint8)
    n = ctypes.c_int64()
    check(_lib.load().holo_link_rectangles(rects.ctypes.data, offsets.ctypes.data, nslices, int(bool(three_d)),
                                           rows.ctypes.data, born.ctypes.data, alive.ctypes.data, ctypes.byref(n)))
    pn: Dict[uuid.UUID, List[int]] = {}
    cur, new_key = None, None
    for i in np.nonzero(alive[:n.value])[0]:
        if born[i] != cur:
            cur, new_key = born[i], _uuid_source()
        pn[new_key()] = rows[i].tolist()
    return pn


def link_slices(per_slice_rectangles, three_d: bool = False):
    """link_arrays over per-slice lists of (x_min, y_min, x_max, y_max) rectangles."""
    if not per_slice_rectangles:
        return {}
    counts = np.fromiter((len(r) for r in per_slice_rectangles), np.int64, len(per_slice_rectangles))
    offsets = np.zeros(counts.size + 1, np.int64)
    np.cumsum(counts, out=offsets[1:])
    flat = np.array([c for r in per_slice_rectangles for br in r for c in br], np.int32).reshape(-1, 4)
    return link_arrays(flat, offsets, three_d)


def delete_duplicates_(particle_neighborhoods) -> None:
    """ccl.jl:446-466 (delete_duplicates!): each neighbourhood in turn absorbs the first other
    neighbourhood (in dict order) it overlaps in 3-D; passes repeat until none overlap."""
    t = _Table(particle_neighborhoods)
    changed = True
    while changed:
        changed = False
        for k in range(t.n):
            if t.dead[k]:
                continue
            alive = np.nonzero(~t.dead[:t.n])[0]
            b = t.a[alive]
            v = t.a[k]
            m = ((v[0] < b[:, 3]) & (v[3] > b[:, 0]) & (v[1] < b[:, 4]) & (v[4] > b[:, 1])
                 & (v[2] < b[:, 5]) & (v[5] > b[:, 2]) & (alive != k))
            if m.any():
                k2 = alive[np.argmax(m)]
                w = t.a[k2]
                t.a[k] = [min(v[0], w[0]), min(v[1], w[1]), min(v[2], w[2]),
                          max(v[3], w[3]), max(v[4], w[4]), max(v[5], w[5])]
                t.dead[k2] = True
                changed = True
    t.store()


def _ratio(a: int, b: int) -> float:
    if b != 0:
        return a / b
    return math.inf if a > 0 else math.nan  # Julia Int / Int: x/0 = Inf, 0/0 = NaN


def finalize_particle_neighborhoods_(particle_neighborhoods) -> None:
    """ccl.jl:411-444 (finalize_particle_neighborhoods!): merges duplicates, then drops
    neighbourhoods two slices deep, with an x/y extent ratio above 3 or below 1/3, or an x-y
    extent product below 10."""
    pn = particle_neighborhoods
    delete_duplicates_(pn)
    for k in list(pn):
        v = pn[k]
        if abs(v[2] - v[5]) == 1:
            pn.pop(k, None)
        r = _ratio(abs(v[0] - v[3]), abs(v[1] - v[4]))
        if r > 3 or r < 1 / 3:
            pn.pop(k, None)
        if abs(v[0] - v[3]) * abs(v[1] - v[4]) < 10:
            pn.pop(k, None)

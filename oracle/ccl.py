"""CPU oracle for connected-component labelling, bounding rectangles and dilation (test
infrastructure only: tests, the smoke check and the CPU-baseline leg may use it; the product
path never does).

Labels follow the reference's union-find result (ccl.jl:186-227): 8-connected components of
equal positive values, labelled 1 + the smallest flat (memory) offset of the component,
background (values <= 0) 0. Arrays are C-ordered (ny, nx), the flat offset j*nx + i.

Pinned by tests/test_oracle_ccl.py: breadth-first search on small random images, the
reference's known answer on binaryparticles.png (84 components), rectangles against direct
scans, dilation against scipy.ndimage.binary_dilation.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

EIGHT = np.ones((3, 3), dtype=bool)


def labels(img) -> np.ndarray:
    """uint32 labels of a 2-D image (float: equal positive values connect; bool/uint8:
    nonzero is foreground)."""
    a = np.asarray(img)
    if a.dtype == np.bool_ or a.dtype == np.uint8:
        keys = (a != 0).astype(np.float64)
    else:
        keys = np.where(a > 0, a, 0).astype(np.float64)
    out = np.zeros(a.shape, np.uint32)
    flat_idx = np.arange(a.size, dtype=np.int64).reshape(a.shape)
    for v in np.unique(keys[keys > 0]):
        comp, n = ndimage.label(keys == v, structure=EIGHT)
        if n == 0:
            continue
        first = np.full(n + 1, np.iinfo(np.int64).max, np.int64)
        np.minimum.at(first, comp.ravel(), flat_idx.ravel())
        m = comp > 0
        out[m] = (first[comp[m]] + 1).astype(np.uint32)
    return out


def boxes(lab) -> dict:
    """{label: (i0, j0, i1, j1)} (0-based, i = column index, j = row index, inclusive)."""
    lab = np.asarray(lab)
    out = {}
    objs = ndimage.find_objects(_dense(lab))
    keys = np.unique(lab[lab > 0])
    for k, sl in zip(keys, [o for o in objs if o is not None]):
        out[int(k)] = (sl[1].start, sl[0].start, sl[1].stop - 1, sl[0].stop - 1)
    return out


def _dense(lab):
    """Relabels to 1..n in ascending label order (find_objects indexes by label value)."""
    keys, inv = np.unique(lab, return_inverse=True)
    dense = inv.reshape(lab.shape)
    if keys[0] != 0:
        dense = dense + 1
    return dense


def dilate(vol) -> np.ndarray:
    """3 x 3 in-plane OR for interior pixels, false on each slice's border
    (particle_detection.jl:12-42). vol: (slices, ny, nx) bool."""
    v = np.asarray(vol, bool)
    out = np.zeros_like(v)
    s, ny, nx = v.shape
    if ny < 3 or nx < 3:
        return out
    acc = np.zeros((s, ny - 2, nx - 2), bool)
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            acc |= v[:, 1 + dj:ny - 1 + dj, 1 + di:nx - 1 + di]
    out[:, 1:ny - 1, 1:nx - 1] = acc
    return out

"""Pins the CCL oracle (oracle/ccl.py): breadth-first search on small images, the reference's
known answer on binaryparticles.png, rectangles by direct scan, dilation by scipy. CPU only."""
from collections import deque
from pathlib import Path

import numpy as np
import pytest
from scipy import ndimage

from oracle import ccl

GOLDEN = Path(__file__).resolve().parent / "golden"


def _bfs_labels(img):
    a = np.asarray(img)
    ny, nx = a.shape
    fg = (a != 0) if a.dtype in (np.bool_, np.uint8) else (a > 0)
    out = np.zeros((ny, nx), np.uint32)
    seen = np.zeros((ny, nx), bool)
    for j in range(ny):          # memory order: the first pixel reached is the smallest offset
        for i in range(nx):
            if not fg[j, i] or seen[j, i]:
                continue
            lab = j * nx + i + 1
            q = deque([(j, i)])
            seen[j, i] = True
            while q:
                y, x = q.popleft()
                out[y, x] = lab
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        yy, xx = y + dy, x + dx
                        if 0 <= yy < ny and 0 <= xx < nx and fg[yy, xx] and not seen[yy, xx] \
                                and a[yy, xx] == a[y, x]:
                            seen[yy, xx] = True
                            q.append((yy, xx))
    return out


@pytest.mark.parametrize("shape,density,seed", [((7, 9), 0.4, 1), ((16, 16), 0.6, 2), ((20, 33), 0.3, 3),
                                                 ((1, 12), 0.5, 4), ((12, 1), 0.5, 5)])
def test_labels_binary_bfs(shape, density, seed):
    img = np.random.default_rng(seed).random(shape) < density
    np.testing.assert_array_equal(ccl.labels(img), _bfs_labels(img))


def test_labels_equal_values_bfs():
    rng = np.random.default_rng(9)
    img = rng.choice(np.array([0.0, -1.0, 0.25, 0.5, 1.0], np.float32), size=(24, 19))
    np.testing.assert_array_equal(ccl.labels(img), _bfs_labels(img))


def test_binaryparticles_known_answer():
    """runtests.jl:198-230: cu_connected_component_labeling(1 - img) has 84 components."""
    from PIL import Image

    g = np.asarray(Image.open(GOLDEN / "binaryparticles.png").convert("L"))
    img = 1.0 - g.T.astype(np.float32) / np.float32(255)  # the reference layout (loader transposes)
    lab = ccl.labels(img)
    flat = lab.reshape(-1)
    roots = np.count_nonzero(flat == np.arange(1, flat.size + 1))
    assert roots == 84
    assert len(np.unique(flat[flat > 0])) == 84


def test_boxes_direct_scan():
    img = np.random.default_rng(4).random((40, 50)) < 0.35
    lab = ccl.labels(img)
    got = ccl.boxes(lab)
    for k in np.unique(lab[lab > 0]):
        rows, cols = np.nonzero(lab == k)
        assert got[int(k)] == (cols.min(), rows.min(), cols.max(), rows.max())


def test_dilate_scipy():
    vol = np.random.default_rng(6).random((3, 17, 23)) < 0.1
    want = np.stack([ndimage.binary_dilation(s, structure=np.ones((3, 3), bool)) for s in vol])
    want[:, 0, :] = want[:, -1, :] = False
    want[:, :, 0] = want[:, :, -1] = False
    np.testing.assert_array_equal(ccl.dilate(vol), want)

"""Host logic of particle_coordinates (particle_detection.jl:110-160) and dictsave / dictload
(utils.jl:340-365), on CPU tensors and the reference's saved particle dictionary."""
import json
import math
import uuid
from pathlib import Path

import numpy as np
import pytest
import torch

from particleholography import particle_detection as pd
from particleholography import utils

GOLDEN = Path(__file__).resolve().parent / "golden"
REF_DATA = Path("/root/reference/test/data")


def test_gaussian_kernel_is_imagefiltering_sigma5():
    k = pd.gaussian_kernel(5.0)
    assert k.size == 21  # 4 ceil(sigma) + 1 taps
    assert abs(k.sum() - 1.0) < 1e-15
    np.testing.assert_array_equal(k, k[::-1])
    assert abs(k[10 + 5] / k[10] - math.exp(-0.5)) < 1e-15
    assert pd.gaussian_kernel(1.2).size == 9


def test_smoothing_replicates_borders_and_keeps_constants():
    k = pd.gaussian_kernel(5.0)
    np.testing.assert_allclose(pd._smooth_replicate(np.full(30, 2.5), k), 2.5, rtol=1e-15)
    imp = np.zeros(41)
    imp[20] = 1.0
    np.testing.assert_allclose(pd._smooth_replicate(imp, k)[10:31], k, atol=1e-17)
    # a step: the first output sees 10 replicated copies of v[0] on its left
    v = np.r_[np.ones(5), np.zeros(30)]
    assert abs(pd._smooth_replicate(v, k)[0] - k[:15].sum()) < 1e-15
    # a short profile (shorter than the kernel) still smooths to its own length
    assert pd._smooth_replicate(np.arange(3.0), k).shape == (3,)


def test_tamura_and_centre_closed_forms():
    a = np.array([[1.0, 3.0], [1.0, 3.0]])
    assert abs(pd.tamura(a) - np.std([1, 3, 1, 3], ddof=1) / 2.0) < 1e-15
    one = np.zeros((5, 7))
    one[3, 2] = 4.0
    assert pd.getcenterfromslice(one) == (4.0, 3.0)
    two = np.zeros((4, 4))
    two[0, 0] = two[3, 3] = 1.0
    assert pd.getcenterfromslice(two) == (2.5, 2.5)


def test_vectorised_tamura_profile_equals_per_slice_loop():
    rng = np.random.default_rng(4)
    block = rng.integers(0, 256, size=(37, 6, 9)).astype(np.float32) / np.float32(255)
    loop = np.array([pd.tamura(block[k]) for k in range(block.shape[0])])
    np.testing.assert_array_equal(pd.depth_profile(pd.tamura, block), loop)
    assert np.isnan(pd.depth_profile(pd.tamura, block[:, :1, :1])).all()  # one pixel: std is NaN


def test_particle_coordinates_of_a_synthetic_block():
    """A dark spot whose contrast peaks at slice 13 (1-based), centred at tensor (row 6, col 9):
    depth 13 and centre of mass in 1-based reference coordinates."""
    slices, ny, nx = 24, 16, 20
    vol = np.full((slices, ny, nx), 200, np.uint8)
    for z in range(slices):
        depth = max(0, 150 - 25 * abs(z - 12))
        vol[z, 5:8, 8:11] = 200 - depth
    bbs = {uuid.UUID(int=1): [5, 8, 1, 12, 11, slices]}  # x = rows 5..12, y = columns 8..11
    got = pd.particle_coordinates(bbs, torch.from_numpy(vol))
    x, y, z = got[uuid.UUID(int=1)]
    assert z == np.float32(13)
    block = vol[12, 4:12, 7:11].astype(np.float64)
    want_x = (block.sum(axis=1) * np.arange(1, 9)).sum() / block.sum() + 5 - 1
    want_y = (block.sum(axis=0) * np.arange(1, 5)).sum() / block.sum() + 8 - 1
    assert (x, y) == (np.float32(want_x), np.float32(want_y))
    assert all(isinstance(v, np.float32) for v in (x, y, z))


def test_particle_coordinates_rejects_float_volumes():
    with pytest.raises(AssertionError):
        pd.particle_coordinates({}, torch.zeros((2, 3, 3)))


def test_dictsave_dictload_round_trip(tmp_path):
    d = {uuid.uuid1(): [np.float32(389.96487), np.float32(1.5), np.float32(551.0)],
         uuid.uuid1(): [np.float32(0.1), np.float32(-2.0), np.float32(3.0)]}
    p = tmp_path / "p.json"
    utils.dictsave(str(p), d)
    text = p.read_text()
    assert "389.96487" in text and "0.1," in text  # shortest Float32 decimals, as Julia prints
    assert utils.dictload(str(p)) == d


def test_dictload_reads_the_reference_particle_file():
    d = utils.dictload(str(GOLDEN / "particles.json"))
    assert len(d) == 22
    assert all(isinstance(k, uuid.UUID) and len(v) == 3 for k, v in d.items())
    raw = json.loads((GOLDEN / "particles.json").read_text())
    assert d[uuid.UUID("0bbd8354-065c-11f1-12a6-194cb3c0f1dc")] == [np.float32(v) for v in raw["0bbd8354-065c-11f1-12a6-194cb3c0f1dc"]]


@pytest.mark.skipif(not (REF_DATA / "holo1.bmp").exists(), reason="reference test data not present")
@pytest.mark.parametrize("name", ["holo1", "holo2"])
def test_png_fixtures_equal_the_reference_bitmaps(name):
    np.testing.assert_array_equal(utils.load_grayimg(str(GOLDEN / f"{name}.png")),
                                  utils.load_grayimg(str(REF_DATA / f"{name}.bmp")))
    assert json.loads((GOLDEN / "particles.json").read_text()) == json.loads((REF_DATA / "particles.json").read_text())

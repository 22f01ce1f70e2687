"""GPU parity of connected components, rectangles, dilation and particle neighbourhoods
(ccl.jl:186-466, particle_detection.jl:12-96) against the oracle, through the C ABI. Labels,
rectangles and dilations are integer results and must match exactly."""
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import ccl as occl

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def _binary(shape, density, seed):
    return np.random.default_rng(seed).random(shape) < density


SHAPES = [(1, 1), (5, 7), (16, 64), (17, 65), (33, 130), (100, 64), (256, 256), (777, 1000), (1024, 1024)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("density", [0.05, 0.45, 0.6])
def test_labels_binary_match_oracle(gpu, shape, density):
    import particleholography as ph

    img = _binary(shape, density, shape[0] * 7 + shape[1] + int(density * 100))
    got = ph.cu_connected_component_labeling(torch.from_numpy(img).to(gpu))
    assert got.dtype == torch.uint32 and tuple(got.shape) == shape
    np.testing.assert_array_equal(got.cpu().numpy(), occl.labels(img))


def test_labels_equal_values_float(gpu):
    import particleholography as ph

    rng = np.random.default_rng(3)
    img = rng.choice(np.array([0.0, -0.5, 0.25, 0.5, 1.0], np.float32), size=(300, 257))
    got = ph.cu_connected_component_labeling(torch.from_numpy(img).to(gpu)).cpu().numpy()
    np.testing.assert_array_equal(got, occl.labels(img))


def test_labels_long_chains_across_tiles(gpu):
    """A serpentine path over the whole image and an 8-connected checkerboard: components
    that cross every tile edge, many times."""
    import particleholography as ph

    n = 512
    snake = np.zeros((n, n), bool)
    snake[::4, :] = True
    snake[1::4, -1] = True
    snake[2::4, :] = True
    snake[3::4, 0] = True
    board = (np.add.outer(np.arange(n), np.arange(n)) % 2 == 0)
    for img in (snake, board, ~snake):
        got = ph.cu_connected_component_labeling(torch.from_numpy(img).to(gpu)).cpu().numpy()
        np.testing.assert_array_equal(got, occl.labels(img))


def test_binaryparticles_known_answer(gpu):
    """runtests.jl:198-230: labels of 1 - img are UInt32 (1024, 1024), 84 components."""
    import particleholography as ph

    img = ph.load_gray2float(str(GOLDEN / "binaryparticles.png"))
    out = ph.cu_connected_component_labeling(torch.from_numpy(1.0 - img).to(gpu))
    assert out.dtype == torch.uint32 and tuple(out.shape) == (1024, 1024)
    assert ph.count_labels(out) == 84
    valid = ph.cu_find_valid_labels(out)
    assert len(valid) == 84 and all(isinstance(v, int) for v in valid) and valid == sorted(valid)
    np.testing.assert_array_equal(out.cpu().numpy(), occl.labels(1.0 - img))


def test_label_stack_and_byte_input(gpu):
    import particleholography as ph

    vol = np.stack([_binary((130, 70), 0.5, s) for s in range(5)])
    got = ph.cu_connected_component_labeling(torch.from_numpy(vol.astype(np.uint8)).to(gpu)).cpu().numpy()
    for z in range(5):
        np.testing.assert_array_equal(got[z], occl.labels(vol[z]))


def test_rectangles_of_label_image(gpu):
    import particleholography as ph

    img = _binary((200, 300), 0.4, 11)
    lab = ph.cu_connected_component_labeling(torch.from_numpy(img).to(gpu))
    valid = ph.cu_find_valid_labels(lab)
    want = occl.boxes(occl.labels(img))
    assert valid == sorted(want)
    rects = ph.get_bounding_rectangles(lab.cpu().numpy(), valid)  # host labels, as the reference
    for l, r in zip(valid, rects):
        i0, j0, i1, j1 = want[l]
        assert r == (j0 + 1, i0 + 1, j1 + 1, i1 + 1)
    # a label without pixels keeps the reference's extremes
    assert ph.get_bounding_rectangles(lab, [valid[0], 10 ** 9])[1] == (2 ** 63 - 1, 2 ** 63 - 1, -2 ** 63, -2 ** 63)


def test_label_boxes_ignore_labels_that_are_not_roots(gpu):
    """Boxes from an arbitrary label image: only labels whose own pixel carries them count."""
    import particleholography as ph

    lab = np.zeros((8, 8), np.uint32)
    lab[0, 0] = 1          # valid root of label 1
    lab[2, 3] = 1          # joins label 1
    lab[5, 5] = 7          # offset 6 holds 0: not a valid label
    lab[0, 1] = 1
    lab[1, 0] = 2          # offset 1 holds 1: label 2 is not valid
    got = ph.cu_find_valid_labels(torch.from_numpy(lab).to(gpu))
    assert got == [1]
    assert ph.get_bounding_rectangles(lab, [1]) == [(1, 1, 3, 4)]


@pytest.mark.parametrize("threshold", [None, 30 / 255, 0.5, -1.0])
def test_slice_rectangles_match_oracle(gpu, threshold):
    import particleholography as ph

    rng = np.random.default_rng(21)
    vol = rng.integers(0, 256, size=(6, 96, 160)).astype(np.uint8)
    if threshold is None:
        vol = (vol < 60).astype(np.uint8)
        fg = vol != 0
    else:
        fg = (vol.astype(np.float64) / 255.0) <= threshold
    per = ph.slice_bounding_rectangles(torch.from_numpy(vol).to(gpu), threshold, max_boxes=8)  # forces a retry
    for z in range(vol.shape[0]):
        want = occl.boxes(occl.labels(fg[z]))
        assert per[z] == [(j0 + 1, i0 + 1, j1 + 1, i1 + 1) for _, (i0, j0, i1, j1) in sorted(want.items())]


def test_dilate_matches_oracle(gpu):
    import particleholography as ph

    vol = _binary((4, 67, 131), 0.05, 8)
    got = ph.cu_dilate(torch.from_numpy(vol).to(gpu))
    assert got.dtype == torch.bool
    np.testing.assert_array_equal(got.cpu().numpy(), occl.dilate(vol))


def test_particle_bounding_boxes_3d_known_answer(gpu):
    """runtests.jl:442-453 (d_bin_vol[1:4, 1:5, z] .= true in the reference layout)."""
    import particleholography as ph

    vol = torch.zeros((3, 6, 6), dtype=torch.bool, device=gpu)
    vol[:, 0:5, 0:4] = True
    pbbs = ph.particle_bounding_boxes_3d(vol)
    assert len(pbbs) == 1 and next(iter(pbbs.values())) == [1, 1, 1, 5, 4, 3]
    pbbs2 = ph.particle_bounding_boxes(vol)
    assert list(pbbs2.values()) == [[1, 1, 1, 5, 4, 3]]


def test_particle_bounding_boxes_on_a_reconstructed_volume(gpu):
    """The reference's detection flow (runtests.jl:405-437): N0f8 volume, d_vol .<= threshold,
    neighbourhoods; the fused threshold equals a Bool volume, and both equal the host linking
    run on the oracle's per-slice rectangles."""
    import particleholography as ph
    from particleholography import ccl, synth

    n, slices = 512, 64
    holo = torch.from_numpy(synth.hologram(12, 150, n)).to(gpu)
    d_sqr = ph.cu_transfer_sqrt_arr(n, 0.6328, 10.0)
    h0, hdz = ph.cu_transfer(-1e4, n, 0.6328, d_sqr), ph.cu_transfer(-1500.0, n, 0.6328, d_sqr)
    vol = ph.cu_get_reconst_vol(ph.cu_gabor_wavefront(holo), h0, hdz, slices)
    assert vol.dtype == torch.uint8
    thr = 30 / 255
    fused = ph.particle_bounding_boxes(vol, threshold=thr)
    binv = (vol.to(torch.float64) / 255.0 <= thr)
    via_bool = ph.particle_bounding_boxes(binv)
    assert list(fused.values()) == list(via_bool.values())
    fg = binv.cpu().numpy()
    rects = [[(j0 + 1, i0 + 1, j1 + 1, i1 + 1) for _, (i0, j0, i1, j1) in sorted(occl.boxes(occl.labels(s)).items())]
             for s in fg]
    want = ccl.gen_particle_neighborhoods(rects[0], 1)
    for z in range(2, slices + 1):
        ccl.update_particle_neighborhoods_(want, rects[z - 1], z)
    ccl.finalize_particle_neighborhoods_(want)
    assert list(fused.values()) == list(want.values())
    assert sum(len(r) for r in rects) > 0

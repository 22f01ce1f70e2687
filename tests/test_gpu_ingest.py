"""GPU parity of ingest (utils.jl:23-25, 217-317; preprocessings.md:21-23) against the oracle,
through the C ABI. All results are bit-exact: the conversion is one correctly rounded
division and FP64 adds, the mode is integer counting, the mean is the reference's FP64 sum in
image order."""
import numpy as np
import pytest
import torch

from oracle import asm

pytestmark = pytest.mark.gpu


def _stack(shape, seed, alphabet=None):
    rng = np.random.default_rng(seed)
    if alphabet is None:
        return rng.integers(0, 256, size=shape).astype(np.uint8)
    return rng.choice(np.asarray(alphabet, np.uint8), size=shape)


def test_background_mode_known_answer(gpu):
    """runtests.jl:125-138 (Julia [10 20; 30 40] etc. in the reference layout)."""
    import particleholography as ph

    imgs = [np.array([[10, 30], [20, 40]], np.uint8), np.array([[10, 30], [25, 45]], np.uint8),
            np.array([[10, 30], [20, 40]], np.uint8)]
    bg = ph.cu_make_background_mode(imgs)
    assert isinstance(bg, np.ndarray) and bg.dtype == np.float64 and bg.shape == (2, 2)
    for (j, i), v in {(0, 0): 10, (1, 0): 20, (0, 1): 30, (1, 1): 40}.items():
        assert abs(bg[j, i] - v / 255) < 1e-6
    # N0f8-valued float images give the same background (utils.jl:298-302)
    bgf = ph.cu_make_background_mode([im.astype(np.float32) / 255 for im in imgs])
    np.testing.assert_array_equal(bgf, bg)


@pytest.mark.parametrize("count,ny,nx,alphabet", [
    (1, 4, 4, None), (2, 3, 5, None), (7, 33, 17, [0, 1, 2]), (64, 1024, 1024, [3, 7, 9, 200]),
    (9, 768, 768, None), (300, 64, 96, [0, 128, 255]), (5, 1, 1, [1, 2]),
    # 8-bit counters up to 256 images: a counter at its 255 limit, 256 equal values (the
    # 256th image is counted in registers), ties at 256, odd sizes (byte loads); 257: 16-bit
    (255, 8, 8, [9]), (256, 8, 12, [9]), (256, 33, 20, [1, 2]), (256, 64, 64, None),
    (256, 3, 5, [0, 1]), (200, 7, 9, [250, 251, 252]), (257, 8, 8, [3, 4])])
def test_background_mode_matches_oracle(gpu, count, ny, nx, alphabet):
    import particleholography as ph

    s = _stack((count, ny, nx), count * 1000 + ny, alphabet)
    got = ph.cu_make_background(torch.from_numpy(s).to(gpu), "mode").cpu().numpy()
    np.testing.assert_array_equal(got, asm.background_mode(s))


def test_background_mode_unaligned_stack(gpu):
    import particleholography as ph

    count, ny, nx = 6, 10, 12
    s = _stack((count, ny, nx), 21, [5, 6, 7])
    base = torch.zeros(count * ny * nx + 1, dtype=torch.uint8, device=gpu)
    base[1:] = torch.from_numpy(s.reshape(-1)).to(gpu)
    view = base[1:].view(count, ny, nx)
    assert view.data_ptr() % 2 == 1
    np.testing.assert_array_equal(ph.cu_make_background(view, "mode").cpu().numpy(), asm.background_mode(s))


def test_background_mode_count_limit(gpu):
    import particleholography as ph

    s = torch.zeros((65536, 1, 1), dtype=torch.uint8, device=gpu)
    with pytest.raises(ValueError, match="65535"):
        ph.cu_make_background(s, "mode")
    assert ph.cu_make_background(s[:65535], "mode").item() == 0.0


@pytest.mark.parametrize("count,ny,nx", [(1, 4, 4), (13, 33, 17), (40, 1024, 1024), (3, 1, 7)])
def test_background_mean_matches_oracle(gpu, count, ny, nx):
    import particleholography as ph

    s = _stack((count, ny, nx), count + 7 * nx)
    got = ph.cu_make_background(torch.from_numpy(s).to(gpu), "mean").cpu().numpy()
    np.testing.assert_array_equal(got, asm.background_mean(s))


def test_make_background_from_files(gpu, tmp_path):
    """make_background(pathlist; mode) on written BMP files (utils.jl:217-240)."""
    from PIL import Image

    import particleholography as ph

    rng = np.random.default_rng(5)
    rasters = [rng.choice(np.array([40, 41, 90], np.uint8), size=(48, 64)) for _ in range(5)]
    paths = []
    for k, r in enumerate(rasters):
        p = tmp_path / f"h{k}.bmp"
        Image.fromarray(np.repeat(r[..., None], 3, axis=2), mode="RGB").save(p)
        paths.append(str(p))
    stack = np.stack([r.T for r in rasters])  # the reference layout of each image
    for mode, ref in (("mode", asm.background_mode), (":mean", asm.background_mean)):
        bg = ph.make_background(paths, mode=mode)
        assert bg.dtype == np.float64 and bg.shape == (64, 48)
        np.testing.assert_array_equal(bg, ref(stack))
    with pytest.raises(ValueError):
        ph.make_background(paths, mode="median")


@pytest.mark.parametrize("count,ny,nx", [(1, 1024, 1024), (5, 256, 256), (3, 33, 17), (2, 1, 3)])
def test_ingest_matches_oracle(gpu, count, ny, nx):
    import particleholography as ph

    s = _stack((count, ny, nx), 31 * count + nx)
    bkg = np.random.default_rng(nx).random((ny, nx))
    plain = ph.cu_ingest_u8(s).cpu().numpy()  # host input, uploaded through pinned memory
    np.testing.assert_array_equal(plain, asm.n0f8_to_f32(s))
    dev = torch.from_numpy(s).to(gpu)
    sub = ph.cu_ingest_u8(dev, bkg=bkg).cpu().numpy()  # offset 64/255 by default
    np.testing.assert_array_equal(sub, asm.ingest_u8(s, bkg, 64 / 255))
    one = ph.cu_ingest_u8(dev[0], bkg=torch.from_numpy(bkg), offset=0.125).cpu().numpy()
    np.testing.assert_array_equal(one, asm.ingest_u8(s[0], bkg, 0.125))


def test_ingest_unaligned_source(gpu):
    import particleholography as ph

    count, ny, nx = 2, 16, 16
    s = _stack((count, ny, nx), 77)
    base = torch.zeros(count * ny * nx + 3, dtype=torch.uint8, device=gpu)
    base[3:] = torch.from_numpy(s.reshape(-1)).to(gpu)
    view = base[3:].view(count, ny, nx)
    np.testing.assert_array_equal(ph.cu_ingest_u8(view).cpu().numpy(), asm.n0f8_to_f32(s))


def test_ingest_feeds_reconstruction(gpu):
    """Ingested 8-bit hologram -> projection equals the projection of load_gray2float's values."""
    import particleholography as ph
    from particleholography import synth

    holo = synth.hologram(9, 300, 512)
    u8 = np.clip(np.rint(holo * 255 / holo.max()), 0, 255).astype(np.uint8)
    img = ph.cu_ingest_u8(u8)
    d_sqr = ph.cu_transfer_sqrt_arr(512, 0.6328, 10.0)
    h0, hdz = ph.cu_transfer(-1e4, 512, 0.6328, d_sqr), ph.cu_transfer(-1000.0, 512, 0.6328, d_sqr)
    got = ph.reconst_xyprojection_holo(img, h0, hdz, 20)
    want = ph.reconst_xyprojection_holo(torch.from_numpy(asm.n0f8_to_f32(u8)).to(gpu), h0, hdz, 20)
    np.testing.assert_array_equal(got.cpu().numpy(), want.cpu().numpy())

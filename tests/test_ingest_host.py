"""Host side of ingest: image decoding and layout of load_gray2float / load_grayimg
(utils.jl:23-41). The loader returns the reference's memory layout: element [c, r] is raster
row r, column c (Julia img[r+1, c+1] of a column-major matrix)."""
import struct
from pathlib import Path

import numpy as np
import pytest

REF_DATA = Path("/root/reference/test/data")


def _write_bmp24(path, raster):
    """Minimal bottom-up 24-bit BMP writer (gray raster -> B = G = R)."""
    h, w = raster.shape
    stride = (3 * w + 3) // 4 * 4
    body = bytearray()
    for r in range(h - 1, -1, -1):
        row = np.repeat(raster[r], 3).tobytes()
        body += row + b"\0" * (stride - 3 * w)
    header = b"BM" + struct.pack("<IHHI", 54 + len(body), 0, 0, 54)
    dib = struct.pack("<IiiHHIIiiII", 40, w, h, 1, 24, 0, len(body), 2835, 2835, 0, 0)
    Path(path).write_bytes(header + dib + bytes(body))


def _read_bmp24(path):
    """Independent parse of an uncompressed 24-bit BMP: gray raster (H, W) from channel 0."""
    b = Path(path).read_bytes()
    off = struct.unpack("<I", b[10:14])[0]
    w, h = struct.unpack("<ii", b[18:26])
    bpp = struct.unpack("<H", b[28:30])[0]
    assert bpp == 24
    stride = (3 * w + 3) // 4 * 4
    rows = [np.frombuffer(b, np.uint8, 3 * w, off + k * stride)[::3] for k in range(abs(h))]
    raster = np.stack(rows)
    return raster[::-1] if h > 0 else raster


def test_loader_layout_on_written_bmp(tmp_path):
    from particleholography import utils

    rng = np.random.default_rng(4)
    raster = rng.integers(0, 256, size=(5, 7)).astype(np.uint8)  # H = 5, W = 7
    p = tmp_path / "h.bmp"
    _write_bmp24(p, raster)
    np.testing.assert_array_equal(_read_bmp24(p), raster)
    g = utils.load_grayimg(str(p))
    assert g.shape == (7, 5) and g.flags.c_contiguous
    np.testing.assert_array_equal(g, raster.T)
    f = utils.load_gray2float(str(p))
    assert f.dtype == np.float32
    np.testing.assert_array_equal(f, raster.T.astype(np.float32) / np.float32(255))


def test_loader_png_gray(tmp_path):
    from PIL import Image

    from particleholography import utils

    raster = np.arange(48, dtype=np.uint8).reshape(6, 8) * 5
    p = tmp_path / "h.png"
    Image.fromarray(raster, mode="L").save(p)
    np.testing.assert_array_equal(utils.load_grayimg(str(p)), raster.T)


@pytest.mark.skipif(not (REF_DATA / "holo1.bmp").exists(), reason="reference test data not present")
def test_loader_on_reference_test_image():
    from particleholography import utils

    raster = _read_bmp24(REF_DATA / "holo1.bmp")
    img = utils.load_gray2float(str(REF_DATA / "holo1.bmp"))
    assert img.shape == (1024, 1024) and img.dtype == np.float32
    np.testing.assert_array_equal(img, raster.T.astype(np.float32) / np.float32(255))

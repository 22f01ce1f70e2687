"""Writes the data fixtures of tests/golden from ParticleHolography.jl's test data
(test/data, MIT-licensed). Run in a tree that holds the reference at /root/reference:

    python tests/golden/make_fixtures.py

- holo1.png, holo2.png: test/data/holo1.bmp and holo2.bmp (24-bit BMP, R = G = B) re-encoded
  as 8-bit gray PNG. The re-encoding is lossless: the script checks R = G = B and that the
  PNG decodes to the same bytes.
- particles.json: test/data/particles.json as it is, the particle coordinates the reference's
  end-to-end test (runtests.jl:153-193) saved with dictsave.
- binaryparticles.png: test/data/binaryparticles.png as it is.
"""
import shutil
import sys
from pathlib import Path

import numpy as np
from PIL import Image

SRC = Path("/root/reference/test/data")
DST = Path(__file__).resolve().parent


def gray_png(name: str) -> None:
    with Image.open(SRC / f"{name}.bmp") as im:
        rgb = np.asarray(im.convert("RGB"))
    if not (np.array_equal(rgb[..., 0], rgb[..., 1]) and np.array_equal(rgb[..., 0], rgb[..., 2])):
        sys.exit(f"{name}.bmp is not gray (R, G, B differ)")
    out = DST / f"{name}.png"
    Image.fromarray(np.ascontiguousarray(rgb[..., 0]), mode="L").save(out, optimize=True)
    with Image.open(out) as im:
        if not np.array_equal(np.asarray(im), rgb[..., 0]):
            sys.exit(f"{out} does not decode to the BMP's bytes")
    print(f"wrote {out} ({out.stat().st_size} bytes)")


def main() -> None:
    for name in ("holo1", "holo2"):
        gray_png(name)
    for name in ("particles.json", "binaryparticles.png"):
        shutil.copyfile(SRC / name, DST / name)
        print(f"copied {name}")


if __name__ == "__main__":
    main()

"""Writes jpeg_vectors.npz: small u8 images and the baseline JPEG bytes
libjpeg-turbo writes for them through Pillow (quality and subsampling only
set, no Huffman optimisation) -- the same library and settings the
reference's JpegCodec::encode_with_options reaches through tj3Compress8
(jpeg/encode.rs:169-258).  Run once, in a Python with Pillow:

    python tests/golden/make_jpeg_golden.py
"""
import io
import pathlib

import numpy as np
from PIL import Image, features

OUT = pathlib.Path(__file__).resolve().parent / "jpeg_vectors.npz"

# (height, width, bands, quality, subsampling: 0 auto, 1 off, 2 4:2:0)
CASES = [(8, 8, 1, 75, 0), (23, 37, 1, 50, 0), (16, 16, 3, 75, 0), (23, 37, 3, 75, 0), (8, 8, 3, 75, 2),
         (10, 16, 3, 20, 2), (23, 37, 3, 95, 0), (17, 33, 3, 90, 1), (9, 130, 3, 5, 0), (31, 2, 3, 100, 2),
         (1, 1, 3, 75, 0), (12, 20, 4, 85, 0)]


def main() -> None:
    r = np.random.default_rng(2024)
    arrays = {}
    for i, (h, w, b, q, ss) in enumerate(CASES):
        px = (r.integers(0, 128, size=(h, w, b)) + np.linspace(0, 120, w)[None, :, None]).astype(np.uint8)
        img = Image.fromarray(px[..., 0] if b == 1 else px[..., :3])
        kw = {} if b == 1 else {"subsampling": {0: 2 if q < 90 else 0, 1: 0, 2: 2}[ss]}
        buf = io.BytesIO()
        img.save(buf, "JPEG", quality=q, optimize=False, **kw)
        arrays[f"px{i}"] = px
        arrays[f"jpeg{i}"] = np.frombuffer(buf.getvalue(), np.uint8)
        arrays[f"args{i}"] = np.array([q, ss], np.int32)
    arrays["libjpeg_turbo"] = np.frombuffer(str(features.version("libjpeg_turbo")).encode(), np.uint8)
    np.savez_compressed(OUT, **arrays)
    print(OUT, len(CASES), "cases")


if __name__ == "__main__":
    main()

"""Seeded inputs and small utilities shared by the tests."""
from __future__ import annotations

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.default_rng(seed)


def noise_u8(h: int, w: int, b: int, seed: int) -> np.ndarray:
    return rng(seed).integers(0, 256, size=(h, w, b), dtype=np.uint8)


def noise_u16(h: int, w: int, b: int, seed: int) -> np.ndarray:
    return rng(seed).integers(0, 65536, size=(h, w, b), dtype=np.uint16)


def gradient_noise_u8(h: int, w: int, b: int, seed: int) -> np.ndarray:
    """Uniform noise averaged with a smooth per-band gradient, so tap windows
    see both texture and slow ramps: p = (U{0..255} + g(x, y, band)) >> 1."""
    n = rng(seed).integers(0, 256, size=(h, w, b), dtype=np.uint16)
    y = np.arange(h, dtype=np.float64)[:, None]
    x = np.arange(w, dtype=np.float64)[None, :]
    out = np.empty((h, w, b), np.uint8)
    for band in range(b):
        if band % 2 == 0:
            g = 255.0 * (x + y) / max(w + h - 2, 1)
        else:
            g = 255.0 * (x + (h - 1 - y)) / max(w + h - 2, 1)
        g = np.floor(g).astype(np.uint16)
        out[..., band] = ((n[..., band] + g) >> 1).astype(np.uint8)
    return out


def float_image(h: int, w: int, b: int, seed: int) -> np.ndarray:
    """clip(N(0.5, 0.2), 0, 1) + 0.05 sin(2 pi (a x + b y) + phi), as float32."""
    r = rng(seed)
    base = np.clip(r.normal(0.5, 0.2, size=(h, w, b)), 0.0, 1.0)
    a, c, phi = r.uniform(0.05, 0.45), r.uniform(0.05, 0.45), r.uniform(0, 2 * np.pi)
    y = np.arange(h)[:, None, None]
    x = np.arange(w)[None, :, None]
    return (base + 0.05 * np.sin(2 * np.pi * (a * x + c * y) + phi)).astype(np.float32)


def grayscale_source(w: int, h: int) -> np.ndarray:
    """(x*17 + y*13 + 5) % 256 (tests/functional/golden/libvips/support_core.rs:37-45)."""
    y, x = np.mgrid[0:h, 0:w]
    return ((x * 17 + y * 13 + 5) % 256).astype(np.uint8)[..., None]


def rgb_source(w: int, h: int) -> np.ndarray:
    """tests/functional/golden/libvips/support_inputs.rs:26-36."""
    y, x = np.mgrid[0:h, 0:w]
    return np.stack([(x * 29 + y * 7 + 3) % 256, (x * 11 + y * 23 + 17) % 256,
                     (x * 5 + y * 13 + 29) % 256], axis=-1).astype(np.uint8)


def smooth_grayscale_source(w: int, h: int) -> np.ndarray:
    """(x*4 + y*3) % 256 (tests/functional/golden/legacy/support.rs:51-55)."""
    y, x = np.mgrid[0:h, 0:w]
    return ((x * 4 + y * 3) % 256).astype(np.uint8)[..., None]


def structural_source(w: int, h: int) -> np.ndarray:
    """(x*17 + y*13) % 256 (tests/functional/golden/basic.rs:148-156)."""
    y, x = np.mgrid[0:h, 0:w]
    return ((x * 17 + y * 13) % 256).astype(np.uint8)[..., None]



def splitmix_bytes(nbytes: int, seed: int) -> np.ndarray:
    """Host copy of the device's seeded synthetic bytes (DeviceBuffer.fill_random)."""
    from oracle.oracle import synthetic_bytes

    return synthetic_bytes(nbytes, seed)

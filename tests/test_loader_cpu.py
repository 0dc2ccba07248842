"""The file loader's conversion (image.rs:450-527) in the oracle, pinned by
known answers and an independent numpy reading of the reference (CPU)."""

import numpy as np
import pytest

import oracle as O


def _taps():
    x = np.arange(5) - 2.0
    raw = np.exp(-x * x / (2.0 * 1.1 * 1.1))
    s = 0.0
    for v in raw:  # in order, as the reference's iterator sum
        s += v
    return raw / s


def _numpy_reading(rgb):
    """image.rs:459-515 written with numpy: integer grey, horizontal pass,
    vertical pass (each as a sum in tap order), round half away from zero,
    clamp.  Padding with np.pad(mode='reflect') is reflect-101."""
    rgb = rgb.astype(np.uint64)
    if rgb.ndim == 2:
        rgb = np.repeat(rgb[..., None], 3, axis=2)
    gs = ((2126 * rgb[..., 0] + 7152 * rgb[..., 1] + 722 * rgb[..., 2] + 5000) // 10000).astype(np.float64)
    k = _taps()
    h, w = gs.shape
    p = np.pad(gs, ((0, 0), (2, 2)), mode="reflect")
    hp = np.zeros((h, w))
    for t in range(5):
        hp = hp + k[t] * p[:, t:t + w]
    p = np.pad(hp, ((2, 2), (0, 0)), mode="reflect")
    out = np.zeros((h, w))
    for t in range(5):
        out = out + k[t] * p[t:t + h, :]
    r = np.where(out >= 0, np.floor(out + 0.5), np.ceil(out - 0.5))  # half away from zero
    return np.clip(r, 0.0, 255.0)


def test_grey_formula_known_answers():
    # BT.709 integer formula with +5000 rounding (image.rs:459-466)
    px = np.array([[[255, 255, 255], [255, 0, 0], [0, 255, 0], [0, 0, 255]]], dtype=np.uint8)
    for c, want in enumerate([255, 54, 182, 18]):
        img = np.tile(px[:, c:c + 1, :], (7, 7, 1))  # a constant frame of one colour
        assert np.all(O.load_convert(img) == want)


def test_constant_frame_is_kept():
    for v in (0, 1, 77, 254, 255):
        assert np.all(O.load_convert(np.full((9, 6), v, np.uint8)) == v)


def test_single_bright_pixel_gives_rounded_outer_product():
    g = np.zeros((11, 13), np.uint8)
    g[5, 6] = 255
    out = O.load_convert(g)
    k = _taps()
    want = np.zeros((11, 13))
    for a in range(5):
        for b in range(5):
            want[5 + a - 2, 6 + b - 2] = np.round(k[a] * (k[b] * 255.0))
    np.testing.assert_array_equal(out, want)


@pytest.mark.parametrize("shape", [(3, 3), (3, 8), (7, 4), (31, 45)])
def test_matches_numpy_reading_including_borders(shape):
    rng = np.random.default_rng(sum(shape))
    rgb = rng.integers(0, 256, size=shape + (3,), dtype=np.uint8)
    np.testing.assert_array_equal(O.load_convert(rgb), _numpy_reading(rgb))
    grey = rng.integers(0, 256, size=shape, dtype=np.uint8)
    np.testing.assert_array_equal(O.load_convert(grey), _numpy_reading(grey))

"""The file loader's conversion on the device (DIC_PIX_RGB8_LOAD /
DIC_PIX_GREY8_LOAD, image.rs:450-527) against the oracle: the converted
frame is bit-exact (integer grey, FP64 blur in the reference's order,
rounding), and the coefficients equal those of the oracle-converted frame
given as 8-bit grey."""

import os

import numpy as np
import pytest

import oracle as O
import geopyv_dev as gp
from geopyv_dev import _native, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(3, 3), (3, 11), (17, 4), (64, 80), (257, 300)])
def test_device_conversion_is_bit_exact(shape):
    rng = np.random.default_rng(shape[0] * 1000 + shape[1])
    eng = _native.Engine.get(0)
    rgb = rng.integers(0, 256, size=shape + (3,), dtype=np.uint8)
    grey = rng.integers(0, 256, size=shape, dtype=np.uint8)
    for px in (rgb, grey):
        print(f"loader conversion {px.shape}", flush=True)
        img = _native.DeviceImage(eng, pixels=px, border=5, load=True)
        want = O.load_convert(px)
        np.testing.assert_array_equal(img.grey(), want)
        ref = _native.DeviceImage(eng, pixels=want.astype(np.uint8), border=5)
        np.testing.assert_array_equal(img.coef(), ref.coef())


def test_speckle_frame_and_png_file(tmp_path):
    from PIL import Image as PILImage
    sp = workloads.Speckle((512, 640), 6000, 2.0, 4)
    grey = sp.reference(2.0)
    rgb = np.stack([grey, np.roll(grey, 3, axis=1), 255 - grey], axis=2).astype(np.uint8)
    path = os.path.join(tmp_path, "frame.png")
    PILImage.fromarray(rgb).save(path)
    f = gp.Image(filepath=path, border=20)
    want = O.load_convert(rgb)
    np.testing.assert_array_equal(f.image_gs, want)
    np.testing.assert_allclose(f.coef, O.get_c(want, 20), rtol=0, atol=1e-9 * 255)
    assert repr(f) == "Image(shape=(512, 640), border=20)"


def test_update_and_stage_with_loader_pixels():
    rng = np.random.default_rng(9)
    eng = _native.Engine.get(0)
    a = rng.integers(0, 256, size=(40, 50, 3), dtype=np.uint8)
    b = rng.integers(0, 256, size=(40, 50, 3), dtype=np.uint8)
    img = _native.DeviceImage(eng, pixels=a, border=8, load=True)
    img.update(pixels=b, load=True)
    np.testing.assert_array_equal(img.grey(), O.load_convert(b))
    with pytest.raises(_native.GeodicError):
        _native.DeviceImage(eng, pixels=np.zeros((2, 9, 3), np.uint8), border=8, load=True)
    plain = _native.DeviceImage(eng, pixels=a[..., 0], border=8)
    with pytest.raises(_native.GeodicError):
        plain.grey()

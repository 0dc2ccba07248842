"""The C-ABI library loads, exports every symbol include/geodic.h declares,
and its host-side entry points (masks, argument checks) behave as the
reference does.  No device compute here."""

import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from geopyv_dev import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "geodic.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dic_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _native.load()
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert sorted(_native.EXPORTED) == names


@pytest.mark.parametrize("shape", ["circle", "square", "semicircle"])
@pytest.mark.parametrize("size", [1, 2, 5, 15, 20, 25, 40])
def test_masks_bit_exact_with_oracle(shape, size):
    c, sm = _native.mask_local(shape, size)
    co, smo = O.mask_local(shape, size)
    assert np.array_equal(c, co) and np.array_equal(sm, smo)


def test_mask_update_bit_exact_with_oracle():
    rng = np.random.default_rng(0)
    m = (rng.uniform(size=(40, 50)) > 0.3).astype(np.uint8)
    for shape in ("circle", "square", "semicircle"):
        for centre in [(0.0, 0.0), (49.0, 39.0), (3.7, 36.2), (25.0, 20.0), (47.5, 2.5), (12.0, 39.99)]:
            for size in (1, 3, 6):
                a = _native.mask_update(shape, size, centre, m)
                b = O.mask_update(shape, size, centre, m)
                assert np.array_equal(a, b), (shape, centre, size)


def test_mask_errors():
    L = _native.load()
    n = ctypes.c_int64()
    assert L.dic_mask_local(0, 0, None, None, ctypes.byref(n)) == _native.ERR_INVALID_INPUT
    assert b"radius must be > 0" in L.dic_last_error()
    assert L.dic_mask_local(1, 0, None, None, ctypes.byref(n)) == _native.ERR_INVALID_INPUT
    assert b"length must be > 0" in L.dic_last_error()
    assert L.dic_mask_local(7, 3, None, None, ctypes.byref(n)) == _native.ERR_INVALID_INPUT
    m = np.ones((10, 10), dtype=np.uint8)
    # a window entirely outside the mask makes the reference's slice panic
    with pytest.raises(RuntimeError):
        _native.mask_update("square", 2, (40.0, 40.0), m)


def test_null_arguments_are_rejected_without_a_device():
    L = _native.load()
    h = ctypes.c_void_p()
    assert L.dic_image_create(None, None, 0, 0, 4, 4, 20, ctypes.byref(h)) == _native.ERR_INVALID_INPUT
    assert L.dic_batch_results(None, None) == _native.ERR_INVALID_INPUT
    assert L.dic_engine_sync(None) == _native.ERR_INVALID_INPUT

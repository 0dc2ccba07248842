"""Host-side behaviour of the Python drop-in that needs no GPU: option
validation, solver dispatch warnings and Mask getters (reference
geopyv_dev/__init__.py:110-296, tests/python/templates/test_templates.py)."""

import warnings

import numpy as np
import pytest

import geopyv_dev as gp
from geopyv_dev import _subset_options, _checked_solver, _resolve_solver


def test_subset_options():
    assert _subset_options(None) == ("uniform", "deformable")
    assert _subset_options({"masking": "adaptive"}) == ("adaptive", "deformable")
    with pytest.raises(TypeError):
        _subset_options([])
    with pytest.raises(ValueError):
        _subset_options({"topology": "uniform"})
    with pytest.raises(ValueError):
        _subset_options({"masking": "zonal"})


def test_solver_dispatch():
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        assert _checked_solver("FAGN") == "fagn"
        assert _checked_solver("bogus") == "icgn"
        assert any(issubclass(x.category, UserWarning) for x in w)
    kw = {"algorithm": "fagn"}
    with pytest.warns(DeprecationWarning):
        assert _resolve_solver(None, kw, "algorithm") == "fagn"
    with pytest.raises(ValueError), warnings.catch_warnings():
        warnings.simplefilter("ignore")
        _resolve_solver("icgn", {"algorithm": "fagn"}, "algorithm")
    assert _resolve_solver(None, {}, "algorithm") == "icgn"


def test_mask_local_getters():
    t = gp.Mask(mask_type="local", shape="circle", size=5)
    assert (t.shape, t.dimension, t.size, t.n_px) == ("circle", "radius", 5, 81)
    assert t.coords.shape == (81, 2) and t.coords.dtype == np.float64
    np.testing.assert_array_equal(t.coords[0], [0.0, -5.0])
    np.testing.assert_array_equal(t.coords[-1], [0.0, 5.0])
    assert t.subset_mask.shape == (11, 11) and t.subset_mask[5, 5] == 1
    assert "circle" in repr(t) and "5" in repr(t)
    assert gp.Mask(mask_type="local", shape="circle").size == 25
    s = gp.Mask(mask_type="local", shape="square", size=5)
    assert s.dimension == "length" and s.n_px == 121 and np.all(s.subset_mask == 1)
    with pytest.raises(Exception):
        gp.Mask(mask_type="local", shape="triangle", size=5)
    with pytest.raises(ValueError):
        gp.Mask(mask_type="bogus")
    with pytest.raises(ValueError):
        gp.Mask(mask_type="local")


def test_mask_method():
    t = gp.Mask(mask_type="local", shape="square", size=2)
    t.mask([5.0, 5.0], np.ones((20, 20), dtype=np.uint8))
    assert t.m_n_px == 25
    t.mask([5.0, 5.0], np.zeros((20, 20), dtype=np.uint8))
    assert t.m_n_px == 0
    t = gp.Mask(mask_type="local", shape="square", size=1)
    t.mask([3.0, 3.0], np.ones((10, 10), dtype=np.uint8))
    assert t.coords.shape == (9, 2)
    np.testing.assert_array_equal(t.coords[0], [-1.0, -1.0])

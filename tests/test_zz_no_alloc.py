"""dic_batch_solve_noalloc: the plain FP64 solve whose per-pixel scratch is
the library's fixed device array (no allocation in the call).  Called
through the C ABI, it returns exactly what dic_batch_solve returns for the
same batch (same kernel and arithmetic), on two calls in a row, the second
with a smaller batch of another order and method; a batch beyond the fixed
array's 64 MiB is refused with an error before anything runs."""

import numpy as np
import pytest

from geopyv_dev import _native, workloads

pytestmark = pytest.mark.gpu

FIELDS = ("p", "c_zncc", "c_znssd", "norm", "eta_u", "eta_v", "f_m", "delta_f", "sssig", "sigma",
          "iterations", "converged", "quality_ok", "status", "n_px", "order")


def _pair():
    field = workloads.Affine((191.5, 191.5), 0.01, 0.004, -0.002, -0.003)
    sp = workloads.Speckle((384, 384), 2200, 2.3, 31)
    return sp.reference(2.0).astype(np.float64), sp.target(field, 2.0).astype(np.float64), field


def _batch(eng, n_side, lo, hi, radius, order, field):
    xs = np.linspace(lo, hi, n_side)
    coords = np.array([[x, y] for y in xs for x in xs])
    xy = np.rint(_native.mask_local("circle", radius)[0]).astype(np.int16)
    start = np.array([0, len(xy)], np.int64)
    b = _native.Batch(eng, coords, order, np.zeros(len(coords), np.int32), start, xy)
    p0 = np.zeros((len(coords), 12))
    p0[:, :2] = np.rint(field.warp(coords[:, 0], coords[:, 1])[:, :2])
    b.set_p0(p0)
    return b


def _equal(a, b):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def test_noalloc_equals_batch_solve_twice_second_smaller():
    eng = _native.Engine(0)
    ref, tar, field = _pair()
    f = _native.DeviceImage(eng, pixels=ref, border=20)
    g = _native.DeviceImage(eng, pixels=tar, border=20)
    cases = [(14, 40.0, 344.0, 18, 1, "icgn"),  # 196 subsets x 1009 px: 6.3 MB of scratch
             (5, 60.0, 324.0, 12, 2, "fagn")]   # 25 subsets, second order, six arrays per pixel
    for n_side, lo, hi, radius, order, method in cases:
        print(f"noalloc: {n_side * n_side} subsets r={radius} order {order} {method}", flush=True)
        b = _batch(eng, n_side, lo, hi, radius, order, field)
        b.solve(f, g, method=method, precision=_native.PREC_PLAIN_F64, max_norm=1e-5)
        want = b.results()
        b.solve_noalloc(f, g, method=method, max_norm=1e-5)
        got = b.results()
        _equal(want, got)
        assert np.all(got["status"] == 0) and np.mean(got["converged"]) > 0.9
        # the per-pixel arrays stay in the shared scratch: not kept on the batch
        with pytest.raises(_native.GeodicError) as ei:
            b.subset_arrays(0)
        assert ei.value.status == 8
        del b


def test_noalloc_refuses_a_batch_beyond_the_scratch():
    eng = _native.Engine(0)
    ref, tar, field = _pair()
    f = _native.DeviceImage(eng, pixels=ref, border=20)
    g = _native.DeviceImage(eng, pixels=tar, border=20)
    # 36 x 36 subsets of r = 25 (1961 px): 2.54 M pixels x 32 B = 81 MB > 64 MiB
    b = _batch(eng, 36, 40.0, 344.0, 25, 1, field)
    with pytest.raises(_native.GeodicError) as ei:
        b.solve_noalloc(f, g, max_norm=1e-5)
    assert ei.value.status == 3 and "fixed scratch" in str(ei.value)
    # the ordinary entry point still solves it
    b.solve(f, g, precision=_native.PREC_PLAIN_F64, max_norm=1e-5, max_iterations=3)
    assert np.all(b.results()["status"] == 0)

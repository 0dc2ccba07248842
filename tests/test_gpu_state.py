"""Batch state across solves on the GPU, through the C ABI:

- a list solve leaves the batch's stored records (results, history) as the
  last whole-batch solve left them, and the per-pixel arrays of an earlier
  FP64 solve are reported as gone (status NOT_SOLVED) once a list or FP32
  solve has run;
- the four-pixel list kernel with a window row wider than 72 floats (a large
  window slack) returns exactly what the whole batch returns under the same
  schedule;
- a reference image whose pixels change after a batched solve (which made
  its knot table of f and its gradient) gives the next solve exactly the
  records of an image created with the new pixels;
- calls on one engine from two threads are serialised: reading results and
  setting warps while another thread records a frame graph leaves both
  correct.
"""

import threading

import numpy as np
import pytest

from geopyv_dev import _native, workloads

pytestmark = pytest.mark.gpu

FIELDS = ("p", "c_zncc", "c_znssd", "norm", "eta_u", "eta_v", "f_m", "delta_f", "sssig", "sigma",
          "iterations", "converged", "quality_ok", "status", "n_px", "order")


@pytest.fixture(scope="module")
def case():
    field = workloads.Affine((191.5, 191.5), 0.008, 0.003, -0.002, -0.004)
    sp = workloads.Speckle((384, 384), 2200, 2.3, 17)
    ref = sp.reference(2.0).astype(np.float64)
    tar = sp.target(field, 2.0).astype(np.float64)
    eng = _native.Engine(0)
    f = _native.DeviceImage(eng, pixels=ref, border=20)
    g = _native.DeviceImage(eng, pixels=tar, border=20)
    xs = np.linspace(50.0, 334.0, 10)
    coords = np.array([[x, y] for y in xs for x in xs])
    p0 = np.zeros((len(coords), 12))
    p0[:, :2] = np.rint(field.warp(coords[:, 0], coords[:, 1])[:, :2])
    return dict(eng=eng, f=f, g=g, coords=coords, p0=p0, tar=tar, ref=ref)


def _batch(c, radius=12):
    xy = np.rint(_native.mask_local("circle", radius)[0]).astype(np.int16)
    b = _native.Batch(c["eng"], c["coords"], 1, np.zeros(len(c["coords"]), np.int32),
                      np.array([0, len(xy)], np.int64), xy)
    b.set_p0(c["p0"])
    return b


def _equal(a, b):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def test_list_solve_keeps_history_and_records(case):
    b = _batch(case)
    b.solve(case["f"], case["g"], precision=_native.PREC_PLAIN_F64, max_norm=1e-5, history=True)
    res0, hist0 = b.results(), b.history()
    assert np.any(hist0 != 0)
    idx = np.array([3, 17, 42], np.int32)
    b.solve_list(case["f"], case["g"], idx, case["p0"][idx], precision=_native.PREC_COMP_F32, max_norm=1e-5)
    b.solve_list(case["f"], case["g"], idx, case["p0"][idx], precision=_native.PREC_PLAIN_F64, max_norm=1e-5)
    _equal(res0, b.results())
    assert np.array_equal(hist0, b.history())


def test_per_pixel_arrays_gone_after_list_or_f32_solve(case):
    b = _batch(case)
    b.solve(case["f"], case["g"], precision=_native.PREC_PLAIN_F64, max_norm=1e-5)
    f0, g0 = b.subset_arrays(5)
    assert len(f0) == b.results()["n_px"][5] and np.all(np.isfinite(g0))
    idx = np.array([1, 2], np.int32)
    b.solve_list(case["f"], case["g"], idx, case["p0"][idx], precision=_native.PREC_PLAIN_F64, max_norm=1e-5)
    for s in (1, 5):
        with pytest.raises(_native.GeodicError) as ei:
            b.subset_arrays(s)
        assert ei.value.status == 8
    # an FP32 whole solve, then an FP64 list: nothing valid for an unlisted subset
    b.solve(case["f"], case["g"], precision=_native.PREC_COMP_F32, max_norm=1e-5)
    b.solve_list(case["f"], case["g"], idx, case["p0"][idx], precision=_native.PREC_PLAIN_F64, max_norm=1e-5)
    with pytest.raises(_native.GeodicError) as ei:
        b.subset_arrays(7)
    assert ei.value.status == 8
    # a new whole FP64 solve makes them available again, as before
    b.solve(case["f"], case["g"], precision=_native.PREC_PLAIN_F64, max_norm=1e-5)
    f1, g1 = b.subset_arrays(5)
    assert np.array_equal(f0, f1) and np.array_equal(g0, g1)


def test_wide_window_list_kernel_equals_whole_batch(case):
    # slack 40 px around a radius-12 template: window rows of 116 floats, so
    # the list kernel at eight wavefronts runs with the window's own stride
    b = _batch(case)
    b.set_schedule(40, 8)
    b.solve(case["f"], case["g"], precision=_native.PREC_COMP_F32, max_norm=1e-5)
    whole = b.results()
    idx = np.array([0, 9, 33, 55, 90, 99, 64], np.int32)
    lst = b.solve_list(case["f"], case["g"], idx, case["p0"][idx], precision=_native.PREC_COMP_F32, max_norm=1e-5)
    _equal(whole[idx], lst)
    assert np.all(lst["status"] == 0) and np.all(lst["converged"] == 1)


def test_two_threads_on_one_engine(case):
    # thread A records and launches frame graphs (capturing the engine's
    # stream); thread B reads results and sets warps on another batch of the
    # same engine meanwhile
    eng = case["eng"]
    b_graph = _batch(case)
    b_other = _batch(case)
    b_other.solve(case["f"], case["g"], precision=_native.PREC_COMP_F32, max_norm=1e-5)
    want = b_other.results()
    # the graph's target: the same frame, staged as pixels for the graph's prefilter
    ga = _native.DeviceImage(eng, pixels=case["tar"], border=20)
    ga.stage(pixels=case["tar"])
    errors = []

    def graphs():
        try:
            for _ in range(6):
                b_graph.reset_warps(case["p0"])
                gr = _native.FrameGraph(b_graph, case["f"], ga, warm_start=False, max_norm=1e-5)
                gr.launch()
                eng.sync()
                del gr
        except Exception as exc:  # reported below
            errors.append(exc)

    def reads():
        try:
            for _ in range(60):
                b_other.set_p0(case["p0"])
                _equal(want, b_other.results())
        except Exception as exc:
            errors.append(exc)

    ta, tb = threading.Thread(target=graphs), threading.Thread(target=reads)
    ta.start()
    tb.start()
    ta.join()
    tb.join()
    assert not errors, errors
    # the graph's batch solved the same frame as the eager batch
    _equal(want, b_graph.results())


def test_reference_update_refreshes_the_knot_table(case):
    eng = case["eng"]
    used = _native.DeviceImage(eng, pixels=case["ref"], border=20)
    # centres on whole pixels: every subset's setup reads the table
    b = _batch(dict(case, coords=np.rint(case["coords"])))
    b.solve(used, case["g"], precision=_native.PREC_COMP_F32, max_norm=1e-5)  # makes used's table
    eng.sync()
    # the pair reversed: the old target is the new reference
    used.update(case["tar"])
    fresh = _native.DeviceImage(eng, pixels=case["tar"], border=20)
    back = _native.DeviceImage(eng, pixels=case["ref"], border=20)
    b.set_p0(-case["p0"])
    b.solve(used, back, precision=_native.PREC_COMP_F32, max_norm=1e-5)
    r_used = b.results()
    b.set_p0(-case["p0"])
    b.solve(fresh, back, precision=_native.PREC_COMP_F32, max_norm=1e-5)
    r_fresh = b.results()
    assert np.mean(r_fresh["converged"]) > 0.9
    _equal(r_used, r_fresh)

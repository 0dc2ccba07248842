"""One batch of mixed warp orders (0, 1 and 2) and uneven radii 15-40 on a
reduced shear-band pair (500 x 700, 300 subsets): each order's size classes
launch side by side (the batched engine) or one launch per order (the FP64
drop-in), list solves and residual maps group their lists by order.  Bars:
the FP64 drop-in against the oracle within 1e-4 px and 1e-6 in the warp's
gradient terms with identical statuses, convergence flags and iteration
counts; the FP32 engine within 1e-4 px and 1e-6 in strain with convergence
agreeing on at least 98 % of the subsets; a subset's record in the mixed
batch equal bit for bit to its record in a batch of its own order alone."""

import numpy as np
import pytest

import oracle as O
import geopyv_dev as gp
from geopyv_dev import _native, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mixed():
    w = workloads.shear_band_pair(seed=2, shape=(500, 700))
    rng = np.random.Generator(np.random.PCG64(23))
    pick = rng.choice(len(w["coords"]), 300, replace=False)
    coords, radii, p0 = w["coords"][pick], w["radii"][pick], w["p0"][pick].copy()
    orders = rng.integers(0, 3, len(pick)).astype(np.int32)
    # the warp's unused slots stay zero (p0 is truncated to p_len as the
    # reference does, subset.rs:1135-1136)
    p0[orders == 0, 2:] = 0.0
    p0[orders == 1, 6:] = 0.0
    ref = w["ref"].astype(np.float64)
    tar = w["tar"].astype(np.float64)
    masks = {r: gp.Mask(mask_type="local", shape="circle", size=int(r)) for r in np.unique(radii)}
    local = [masks[r] for r in radii]
    index, start, xy = gp.templates_from_masks([m.coords for m in local])
    Cf, Cg = O.get_c(ref, 20), O.get_c(tar, 20)
    orc = O.batch_solve(Cf, Cg, ref.shape, 20, coords, index, xy.astype(np.float64), start, orders, p0,
                        max_norm=1e-5, max_iterations=50)
    f = gp.Image(image_gs=ref, border=20)
    g = gp.Image(image_gs=tar, border=20)
    return dict(coords=coords, radii=radii, p0=p0, orders=orders, ref=ref, tar=tar, index=index, start=start,
                xy=xy, orc=orc, f=f, g=g, Cf=Cf, Cg=Cg, local=local)


def _batch(m, sel=None):
    if sel is None:
        sel = np.arange(len(m["coords"]))
    index, start, xy = gp.templates_from_masks([m["local"][k].coords for k in sel])
    b = _native.Batch(gp._engine(), m["coords"][sel], m["orders"][sel], index, start, xy)
    b.set_p0(m["p0"][sel])
    return b


def _grad_terms(p, orders):
    # the warp's gradient terms each order solves for (slots 2..p_len)
    d = np.zeros(len(p))
    for o, plen in ((1, 6), (2, 12)):
        sel = orders == o
        if sel.any():
            d[sel] = np.abs(p[sel, 2:plen]).max(axis=1)
    return d


def test_mixed_orders_plain_f64_matches_oracle(mixed):
    m, orc = mixed, mixed["orc"]
    assert set(np.unique(m["orders"])) == {0, 1, 2}
    b = _batch(m)
    b.solve(m["f"]._dev, m["g"]._dev, precision=_native.PREC_PLAIN_F64, max_norm=1e-5)
    r = b.results()
    assert np.array_equal(r["status"], orc["status"])
    assert np.array_equal(r["order"], m["orders"])
    ok = orc["status"] == 0
    du = np.abs(r["p"][ok, :2] - orc["p"][ok, :2]).max()
    ds = _grad_terms(r["p"] - orc["p"], m["orders"])[ok].max()
    print(f"mixed orders, plain f64: max |du| {du:.2e} px, max |d gradient terms| {ds:.2e}")
    assert du < 1e-4 and ds < 1e-6
    assert np.array_equal(r["converged"], orc["converged"])
    assert np.array_equal(r["iterations"], orc["iterations"])


def test_mixed_orders_f32_engine_matches_oracle(mixed):
    m, orc = mixed, mixed["orc"]
    b = _batch(m)
    b.solve(m["f"]._dev, m["g"]._dev, precision=_native.PREC_COMP_F32, max_norm=1e-5)
    r = b.results()
    assert np.array_equal(r["order"], m["orders"])
    both = (r["converged"] == 1) & (orc["converged"] == 1)
    du = np.abs(r["p"][both, :2] - orc["p"][both, :2]).max()
    strain = np.abs(r["p"][both, 2:6] - orc["p"][both, 2:6]).max()
    agree = np.mean(r["converged"] == orc["converged"])
    print(f"mixed orders, f32 engine: max |du| {du:.2e} px, max |dstrain| {strain:.2e}, "
          f"convergence agreement {agree:.4f}")
    assert du < 1e-4 and strain < 1e-6 and agree >= 0.98  # strain measured 6e-8
    np.testing.assert_allclose(r["c_zncc"][both], orc["c_zncc"][both], atol=1e-5)
    np.testing.assert_allclose(r["eta_u"][both], orc["eta_u"][both], rtol=2e-2, atol=1e-6)


@pytest.mark.parametrize("precision", [_native.PREC_PLAIN_F64, _native.PREC_COMP_F32], ids=["f64", "f32"])
def test_mixed_batch_equals_batches_of_one_order(mixed, precision):
    m = mixed
    b = _batch(m)
    b.solve(m["f"]._dev, m["g"]._dev, precision=precision, max_norm=1e-5)
    full = b.results()
    for o in (0, 1, 2):
        sel = np.flatnonzero(m["orders"] == o)
        bo = _batch(m, sel)
        bo.solve(m["f"]._dev, m["g"]._dev, precision=precision, max_norm=1e-5)
        assert bo.results().tobytes() == full[sel].tobytes(), o


def test_mixed_list_solves(mixed):
    m = mixed
    b = _batch(m)
    f, g = m["f"]._dev, m["g"]._dev
    b.solve(f, g, precision=_native.PREC_PLAIN_F64, max_norm=1e-5)
    full = b.results()
    rng = np.random.Generator(np.random.PCG64(5))
    idx = rng.choice(len(m["coords"]), 40, replace=False).astype(np.int32)
    assert len(np.unique(m["orders"][idx])) == 3
    # FP64: the listed records equal the whole batch's, in list order
    got = b.solve_list(f, g, idx, m["p0"][idx], precision=_native.PREC_PLAIN_F64, max_norm=1e-5)
    assert got.tobytes() == full[idx].tobytes()
    # FP32 (the list runs at more wavefronts per subset): within the bar
    b.solve(f, g, precision=_native.PREC_COMP_F32, max_norm=1e-5)
    full32 = b.results()
    got32 = b.solve_list(f, g, idx, m["p0"][idx], precision=_native.PREC_COMP_F32, max_norm=1e-5)
    assert np.array_equal(got32["order"], m["orders"][idx])
    both = (got32["converged"] == 1) & (full32["converged"][idx] == 1)
    assert np.abs(got32["p"][both, :2] - full32["p"][idx][both, :2]).max() < 1e-4


def test_mixed_residual_map(mixed):
    m = mixed
    b = _batch(m)
    f, g = m["f"]._dev, m["g"]._dev
    b.solve(f, g, precision=_native.PREC_PLAIN_F64, max_norm=1e-5)
    r = b.results()
    pick = [int(np.flatnonzero(m["orders"] == o)[k]) for k in (0, 3) for o in (2, 0, 1)]
    maps = b.residual_map(f, g, pick)
    for k, s in enumerate(pick):
        eo = O.residual_map(m["Cf"], m["Cg"], m["ref"].shape, 20, m["coords"][s], m["local"][s].coords,
                            int(m["orders"][s]), r["p"][s])
        assert maps[k].shape == eo.shape
        assert np.abs(maps[k] - eo).max() < 1e-10, (k, s)

"""GPU parity of the mesh solve (dic_mesh_solve, and the list solves it is
built on) against the oracle's restatement of mesh.rs
(oracle/mesh_oracle.cpp), on the reference's synthetic fixtures and on a
speckle pair.

Bars: the FP64 drop-in gives the oracle's seed node and quality flags,
displacements within 1e-4 px, first-order strains within 1e-6 and C_ZNCC
within 1e-6 at every node, and the oracle's iteration count at 95 % of the
nodes or more (90 % for second-order subsets): a node whose last update
norm lies within round-off of max_norm may stop one iteration earlier or
later, which moves it (and its neighbours' initial warps) by ~1e-6 px; the
FP32 engine the same displacement bar.  A list solve returns exactly what the whole batch returns for the
listed subsets under the same schedule, also over mixed template sizes (dispatched largest first, returned in
list order); under the automatic schedule short lists agree with the whole batch to the FP32 engine's bar.
"""

import numpy as np
import pytest

import oracle as O
import geopyv_dev as gp
from geopyv_dev import _native, workloads
from mesh_fixtures import band_grid, circle_template, shift_band, texture

pytestmark = pytest.mark.gpu


def _case(ref, tar, border, nodes, el, radius, order=1):
    eng = _native.Engine.get(0)
    return dict(ref=ref, tar=tar, border=border, nodes=nodes, el=el, radius=radius, order=order,
                Cf=O.get_c(ref, border), Cg=O.get_c(tar, border),
                f=_native.DeviceImage(eng, pixels=ref, border=border),
                g=_native.DeviceImage(eng, pixels=tar, border=border))


@pytest.fixture(scope="module")
def band():
    tex = texture(100)
    nodes, el = band_grid()
    return _case(tex, shift_band(tex, 50, 1.5, -1.5), 15, nodes, el, 6)


@pytest.fixture(scope="module")
def speckle():
    field = workloads.Affine((256.0, 256.0), 0.01, 0.005, 0.0, -0.003)
    sp = workloads.Speckle((512, 512), 8000, 2.5, 11)
    ref = sp.reference(2.0).astype(np.float64)
    tar = sp.target(field, 2.0).astype(np.float64)
    nodes, el = gp.structured_mesh(60.0, 60.0, 452.0, 452.0, 14.0)
    return _case(ref, tar, 20, nodes, el, 20)


def _both(c, seed, precision=_native.PREC_PLAIN_F64, **kw):
    xy, start = circle_template(c["radius"])
    tpl = np.zeros(len(c["nodes"]), np.int32)
    order = kw.pop("subset_order", 1)
    common = dict(subset_order=order, tolerance=kw.pop("tolerance", 0.5), seed_tolerance=kw.pop("seed_tolerance", 0.5),
                  override=kw.pop("override", True), **kw)
    print(f"mesh solve: {len(c['nodes'])} nodes, order {order}, {common}", flush=True)
    st_o, ro = O.mesh_solve(c["Cf"], c["Cg"], c["ref"].shape, c["border"], c["nodes"], c["el"], tpl, xy, start, seed,
                            np.zeros(6 * order), **common)
    st_g, msg, rg = _native.mesh_solve(_native.Engine.get(0), c["f"], c["g"], c["nodes"], c["el"],
                                       2 if c["el"].shape[1] == 6 else 1, tpl, start, np.rint(xy).astype(np.int16),
                                       seed, np.zeros(6 * order), precision=precision, **common)
    return st_o, ro, st_g, msg, rg


def _compare(ro, rg, disp_tol=1e-4, strain_tol=1e-6, zncc_tol=1e-6, iter_frac=0.95):
    assert rg["seed_node"] == ro["seed_node"]
    np.testing.assert_array_equal(rg["quality_ok"], ro["quality_ok"])
    du = np.abs(rg["p"][:, :2] - ro["p"][:, :2]).max()
    ds = np.abs(rg["p"][:, 2:6] - ro["p"][:, 2:6]).max()
    dz = np.abs(rg["c_zncc"] - ro["c_zncc"]).max()
    same_it = float(np.mean(rg["iterations"] == ro["iterations"]))
    print(f"max |du| {du:.3e} px, |dstrain| {ds:.3e}, |dzncc| {dz:.3e}, equal iterations {same_it:.4f}, "
          f"solves {rg['n_node_solves']} vs {ro['n_solves']}, lists {rg['n_lists']}", flush=True)
    assert du <= disp_tol
    assert ds <= strain_tol
    assert dz <= zncc_tol
    assert same_it >= iter_frac
    np.testing.assert_allclose(rg["warps"][:, :6], ro["warps"][:, :6], rtol=0, atol=10 * disp_tol)


@pytest.mark.parametrize("order,wps", [(1, 2), (0, 8), (1, 8), (2, 8)])
def test_list_solve_equals_whole_batch(speckle, order, wps):
    c = speckle
    xy, start = circle_template(c["radius"])
    n = len(c["nodes"])
    b = _native.Batch(_native.Engine.get(0), c["nodes"], order, np.zeros(n, np.int32), start,
                      np.rint(xy).astype(np.int16))
    rng = np.random.default_rng(5)
    p0 = np.zeros((n, 12))
    p0[:, :2] = rng.uniform(-1, 1, size=(n, 2))
    # an explicit schedule: a short list otherwise takes more wavefronts per
    # subset than the whole batch, which rounds the FP32 engine's sums
    # differently.  With eight wavefronts the list (r = 20: at most four
    # pixels per thread) runs the unrolled list kernel, the whole batch the
    # rolled one: their sums must agree bit for bit.
    print(f"list vs whole batch: order {order}, {wps} wavefronts per subset", flush=True)
    b.set_schedule(3, wps)
    for prec in (_native.PREC_PLAIN_F64, _native.PREC_COMP_F32):
        b.set_p0(p0)
        b.solve(c["f"], c["g"], precision=prec, max_iterations=50, tolerance=0.5, max_norm=1e-5)
        whole = b.results()
        idx = rng.permutation(n)[: n // 3]
        part = b.solve_list(c["f"], c["g"], idx, p0[idx], precision=prec, max_iterations=50, tolerance=0.5,
                            max_norm=1e-5)
        differ = [k for k in ("p", "c_zncc", "c_znssd", "norm", "eta_u", "eta_v", "f_m", "delta_f", "sssig",
                              "sigma", "iterations", "converged", "quality_ok", "status")
                  if not np.array_equal(part[k], whole[idx][k])]
        assert not differ, f"precision {prec}: fields differ: {differ}"
        # a list solve writes its records only to its own output
        after = b.results()
        for k in ("p", "c_zncc", "iterations", "converged", "status"):
            np.testing.assert_array_equal(after[k], whole[k])
    with pytest.raises(_native.GeodicError):
        b.solve_list(c["f"], c["g"], [1, 1], np.zeros((2, 12)))
    with pytest.raises(_native.GeodicError):
        b.solve_list(c["f"], c["g"], [n], np.zeros((1, 12)))


def _two_radius_batch(c):
    """The speckle nodes with circle templates r = 20 and r = 30 alternating."""
    xy20, _ = circle_template(20)
    xy30, _ = circle_template(30)
    xy = np.vstack([xy20, xy30])
    start = np.array([0, len(xy20), len(xy20) + len(xy30)], dtype=np.int64)
    n = len(c["nodes"])
    tpl = (np.arange(n) % 2).astype(np.int32)
    b = _native.Batch(_native.Engine.get(0), c["nodes"], 1, tpl, start, np.rint(xy).astype(np.int16))
    return b, n


@pytest.mark.parametrize("wps", [4, 8])
def test_mixed_size_list_equals_whole_batch(speckle, wps):
    # a list over two template sizes is dispatched largest first and its
    # results are returned in list order
    c = speckle
    b, n = _two_radius_batch(c)
    rng = np.random.default_rng(7)
    p0 = np.zeros((n, 12))
    p0[:, :2] = rng.uniform(-1, 1, size=(n, 2))
    print(f"mixed-size list vs whole batch: {wps} wavefronts per subset", flush=True)
    b.set_schedule(3, wps)
    for prec in (_native.PREC_PLAIN_F64, _native.PREC_COMP_F32):
        b.set_p0(p0)
        b.solve(c["f"], c["g"], precision=prec, max_iterations=50, tolerance=0.5, max_norm=1e-5)
        whole = b.results()
        idx = rng.permutation(n)[: n // 4]
        part = b.solve_list(c["f"], c["g"], idx, p0[idx], precision=prec, max_iterations=50, tolerance=0.5,
                            max_norm=1e-5)
        np.testing.assert_array_equal(part["n_px"], whole[idx]["n_px"])
        for k in ("p", "c_zncc", "eta_u", "iterations", "converged", "status"):
            np.testing.assert_array_equal(part[k], whole[idx][k])


def test_mixed_size_short_lists_in_latency_mode(speckle):
    # automatic schedule: a short list spreads each size class over more
    # wavefronts (r = 20 through the four-pixel list kernel, r = 30 through
    # the rolled one), which rounds the FP32 sums differently from the whole
    # batch; results agree to the engine's bar
    c = speckle
    b, n = _two_radius_batch(c)
    rng = np.random.default_rng(8)
    p0 = np.zeros((n, 12))
    p0[:, :2] = rng.uniform(-1, 1, size=(n, 2))
    b.set_p0(p0)
    b.solve(c["f"], c["g"], precision=_native.PREC_COMP_F32, max_iterations=50, tolerance=0.5, max_norm=1e-5)
    whole = b.results()
    for size in (1, 7, 60):
        idx = rng.permutation(n)[:size]
        print(f"latency-mode list of {size}", flush=True)
        part = b.solve_list(c["f"], c["g"], idx, p0[idx], precision=_native.PREC_COMP_F32, max_iterations=50,
                            tolerance=0.5, max_norm=1e-5)
        np.testing.assert_array_equal(part["n_px"], whole[idx]["n_px"])
        np.testing.assert_array_equal(part["status"], whole[idx]["status"])
        ok = whole[idx]["converged"] == 1
        assert np.abs(part["p"][ok, :2] - whole[idx]["p"][ok, :2]).max() < 1e-4
        assert np.mean(part["converged"] == whole[idx]["converged"]) >= 0.95


@pytest.mark.parametrize("pre", ["RG", "layer-RG"])
def test_band_fixture_f64_matches_oracle(band, pre):
    st_o, ro, st_g, msg, rg = _both(band, [50.0, 30.0], preconditioning=pre, workers=8)
    assert st_o == 0 and st_g == 0, msg
    _compare(ro, rg)


def test_speckle_layer_rg_f64_matches_oracle(speckle):
    st_o, ro, st_g, msg, rg = _both(speckle, [256.0, 256.0], preconditioning="layer-RG", workers=16,
                                    tolerance=0.75, seed_tolerance=0.9, override=False, max_norm=1e-5)
    assert st_o == 0 and st_g == 0, msg
    assert ro["quality_ok"].all()
    _compare(ro, rg)


def test_speckle_layer_rg_f32_engine_matches_oracle(speckle):
    st_o, ro, st_g, msg, rg = _both(speckle, [256.0, 256.0], precision=_native.PREC_COMP_F32,
                                    preconditioning="layer-RG", workers=16, tolerance=0.75, seed_tolerance=0.9)
    assert st_o == 0 and st_g == 0, msg
    _compare(ro, rg, disp_tol=1e-4, strain_tol=1e-6, zncc_tol=1e-5, iter_frac=0.8)


def test_order2_quadratic_mesh_matches_oracle(speckle):
    nodes, el = gp.structured_mesh(100.0, 100.0, 412.0, 412.0, 26.0)
    nodes, el = gp.quadratic_mesh(nodes, el)
    c = dict(speckle, nodes=nodes, el=el)
    st_o, ro, st_g, msg, rg = _both(c, [256.0, 256.0], subset_order=2, preconditioning="layer-RG", workers=16,
                                    tolerance=0.75, seed_tolerance=0.9)
    assert st_o == 0 and st_g == 0, msg
    # second-order subsets converge more slowly, so more nodes stop within
    # round-off of the max_norm test; one extra (or one fewer) iteration
    # moves a node by ~1e-6 px and its neighbours' initial warps with it
    _compare(ro, rg, iter_frac=0.9)
    np.testing.assert_allclose(rg["warps"], ro["warps"], rtol=0, atol=1e-3)


def test_layer_rg_engine_is_deterministic(speckle):
    runs = [_both(speckle, [256.0, 256.0], precision=_native.PREC_COMP_F32, preconditioning="layer-RG")[4]
            for _ in range(2)]
    for k in ("p", "c_zncc", "iterations"):
        np.testing.assert_array_equal(runs[0][k], runs[1][k])


def test_unsolvable_gate_matches_oracle(band):
    st_o, ro, st_g, msg, rg = _both(band, [50.0, 30.0], tolerance=0.99995, override=False,
                                    preconditioning="layer-RG", workers=8)
    assert st_o == st_g == _native.ERR_INVALID_INPUT
    assert "mesh unsolvable" in msg
    assert rg["n_unsolvable"] == ro["n_bad"] > 0


def test_drop_in_mesh_class(band):
    f = gp.Image(image_gs=band["ref"], border=15)
    g = gp.Image(image_gs=band["tar"], border=15)
    m = gp.Mesh(f_img=f, g_img=g, nodes=band["nodes"], elements=band["el"])
    with pytest.raises(AttributeError):
        m.displacements
    m.solve(gp.Mask("local", shape="circle", size=6), seed_coord=[50.0, 30.0], seed_warp=[0.0] * 6,
            tolerance=0.5, seed_tolerance=0.5, override_active=True,
            solver_options={"preconditioning": "layer-RG", "layer_rg": {"max_workers": 8}})
    _, ro, _, _, _ = _both(band, [50.0, 30.0], preconditioning="layer-RG", workers=8)
    assert m.solved and m.seed_node == ro["seed_node"]
    assert np.abs(m.displacements - ro["p"][:, :2]).max() <= 1e-4
    assert m.warps.shape == (len(band["el"]), 12) and m.areas.shape == (len(band["el"]),)
    with pytest.raises(RuntimeError, match="mesh unsolvable"):
        m.solve(gp.Mask("local", shape="circle", size=6), seed_coord=[50.0, 30.0], tolerance=0.99995,
                seed_tolerance=0.5)


class _Counting:
    """SolveProgress observer counting node_done calls and cancelling once
    stop_after have been reported (mesh.rs:2995-3012)."""

    def __init__(self, stop_after=None):
        self.stop_after = stop_after
        self.started = []
        self.done = 0
        self.polls = 0

    def mesh_start(self, n):
        self.started.append(n)

    def node_done(self):
        self.done += 1

    def cancelled(self):
        self.polls += 1
        return self.stop_after is not None and self.done >= self.stop_after


@pytest.mark.parametrize("pre", ["RG", "layer-RG"])
def test_progress_reports_every_node_and_cancels(band, pre):
    # mesh.rs:3013-3036 through the C ABI: uncancelled, one mesh_start(n)
    # and at least n node_done calls with results equal to a solve without
    # an observer; cancelled after 5 (and, for layer-RG, after 37: inside
    # a colour's list) nodes, status DIC_ERR_CANCELLED, exactly that many
    # node_done calls and the outputs untouched
    c = band
    xy, start = circle_template(c["radius"])
    tpl = np.zeros(len(c["nodes"]), np.int32)
    args = (_native.Engine.get(0), c["f"], c["g"], c["nodes"], c["el"], 1, tpl, start, np.rint(xy).astype(np.int16),
            [50.0, 50.0], np.zeros(6))
    kw = dict(tolerance=0.5, seed_tolerance=0.5, override=True, preconditioning=pre)
    n = len(c["nodes"])
    obs = _Counting()
    st, _, r_obs = _native.mesh_solve(*args, progress=obs, **kw)
    st0, _, r_none = _native.mesh_solve(*args, **kw)
    assert st == 0 and st0 == 0
    assert obs.started == [n] and obs.done >= n and obs.polls >= obs.done
    for k in ("p", "c_zncc", "iterations", "norms", "quality_ok", "warps"):
        np.testing.assert_array_equal(r_obs[k], r_none[k])
    for stop in ([5, 37] if pre == "layer-RG" else [5]):
        sentinel = dict(p=np.full((n, 12), 7.0), c_zncc=np.full(n, 7.0), iterations=np.full(n, 7, np.int32),
                        norms=np.full(n, 7.0), quality_ok=np.full(n, 7, np.int32),
                        warps=np.full((len(c["el"]), 12), 7.0))
        obs = _Counting(stop_after=stop)
        st, msg, _ = _native.mesh_solve(*args, progress=obs, out=sentinel, **kw)
        print(f"{pre}: cancelled after {obs.done} nodes: status {st} ({msg})", flush=True)
        assert st == _native.ERR_CANCELLED and msg == "solve cancelled"
        assert obs.done == stop and obs.started == [n]
        for k, a in sentinel.items():
            assert np.all(a == 7), k


def test_mesh_class_cancel_stores_nothing(speckle):
    c = speckle
    m = gp.Mesh(f_img=gp.Image(image_gs=c["ref"], border=20), g_img=gp.Image(image_gs=c["tar"], border=20),
                nodes=c["nodes"], elements=c["el"])
    with pytest.raises(RuntimeError, match="solve cancelled"):
        m.solve(gp.Mask("local", shape="circle", size=20), seed_coord=[256.0, 256.0], tolerance=0.5,
                seed_tolerance=0.5, progress=_Counting(stop_after=3))
    assert not m.solved
    obs = _Counting()
    m.solve(gp.Mask("local", shape="circle", size=20), seed_coord=[256.0, 256.0], tolerance=0.5, seed_tolerance=0.5,
            progress=obs)
    assert m.solved and obs.started == [len(c["nodes"])] and obs.done >= len(c["nodes"])


def test_solved_mesh_saves_as_pyv(band, tmp_path):
    # a solved Mesh written as .pyv (GPYV, 0x17, bincode v2, zstd) loads back
    # with every array of the solution unchanged
    import os
    f = gp.Image(image_gs=band["ref"], border=15)
    g = gp.Image(image_gs=band["tar"], border=15)
    m = gp.Mesh(f_img=f, g_img=g, nodes=band["nodes"], elements=band["el"])
    m.solve(gp.Mask("local", shape="circle", size=6), seed_coord=[50.0, 30.0], seed_warp=[0.0] * 6,
            tolerance=0.5, seed_tolerance=0.5, override_active=True, solver_options={"preconditioning": "layer-RG"})
    path = os.path.join(tmp_path, "band.pyv")
    m.save(path)
    raw = open(path, "rb").read()
    assert raw[:5] == b"GPYV\x17"
    s = gp.load(path)
    assert np.array_equal(s.displacements, m.displacements) and np.array_equal(s.p, m.p)
    assert np.array_equal(s.c_zncc, m.c_zncc) and np.array_equal(s.iterations, m.iterations)
    assert np.array_equal(s.warps, m.warps) and np.array_equal(s.nodes, m.nodes)
    assert np.array_equal(s.elements, m.elements) and s.seed_node == m.seed_node
    assert s.solve_config["preconditioning"] == "layer-RG" and s.template_shape == "circle"

"""Mesh orchestration on the CPU: the oracle's restatement of mesh.rs pinned
against the reference's own known-answer tests (src/mesh.rs:2515-3246,
src/particle.rs:2616-2642), the library's host-side mesh functions against
the oracle, and the oracle's full mesh solve on the reference's synthetic
fixtures (layer-RG within Tier C of RG, run-to-run identical)."""

import numpy as np
import pytest

import oracle as O
from mesh_fixtures import band_grid, circle_template, shift_band, shift_uniform, texture

TWO = np.array([[0, 1, 2], [1, 3, 2]])
TWO6 = np.array([[0, 1, 2, 4, 5, 6], [1, 3, 2, 7, 8, 5]])
UNIT4 = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]])

# particle.rs:2621 warp_convention_tests::Q and quad_warp
Q = [0.1, -0.2, 0.02, 0.07, 0.03, -0.01, 0.004, 0.008, 0.005, 0.009, 0.006, -0.003]


def quad_warp(x, y):
    q = Q
    return np.array([
        q[0] + q[2] * x + q[4] * y + 0.5 * q[6] * x * x + q[8] * x * y + 0.5 * q[10] * y * y,
        q[1] + q[3] * x + q[5] * y + 0.5 * q[7] * x * x + q[9] * x * y + 0.5 * q[11] * y * y,
        q[2] + q[6] * x + q[8] * y, q[3] + q[7] * x + q[9] * y,
        q[4] + q[8] * x + q[10] * y, q[5] + q[9] * x + q[11] * y,
    ] + q[6:])


def six_node_element():
    c = [[1.0, 2.0], [9.0, 3.5], [3.0, 8.0]]
    mid = lambda a, b: [(c[a][0] + c[b][0]) / 2, (c[a][1] + c[b][1]) / 2]
    return np.array(c + [mid(0, 1), mid(1, 2), mid(2, 0)])


# ---------------------------------------------------------------------------
# connectivity (mesh.rs:2656-2717)
# ---------------------------------------------------------------------------
def test_connectivity_known_answers():
    assert O.mesh_connectivity(TWO, 0, False) == [1, 2]
    assert O.mesh_connectivity(TWO, 1, False) == [0, 2, 3]
    assert O.mesh_connectivity(TWO6, 0, False) == [4, 6]
    assert O.mesh_connectivity(TWO6, 4, False) == [0, 1]


def test_connectivity_brute_force_on_random_meshes():
    """Every node of random 3- and 6-node meshes against a direct reading
    of mesh.rs:1705-1764."""
    rng = np.random.default_rng(4)
    partner = {0: (3, 5), 1: (3, 4), 2: (4, 5), 3: (0, 1), 4: (1, 2), 5: (0, 2)}
    for cols in (3, 6):
        el = np.array([rng.choice(30, cols, replace=False) for _ in range(25)])
        for idx in range(30):
            for full in (False, True):
                want = set()
                for row in el:
                    if idx not in row:
                        continue
                    if cols == 3 or full:
                        want |= {int(v) for v in row if v != idx}
                    else:
                        for col in np.nonzero(row == idx)[0]:
                            want |= {int(row[k]) for k in partner[int(col)] if row[k] != idx}
                assert O.mesh_connectivity(el, idx, full) == sorted(want)


# ---------------------------------------------------------------------------
# find_seed_node, corr_1d, flow_calc, r_stats, project_warp (mesh.rs:2723-2899)
# ---------------------------------------------------------------------------
def test_find_seed_node():
    assert O.mesh_find_seed(UNIT4[:3], [1.0, 0.0]) == 1
    assert O.mesh_find_seed(UNIT4, [0.4, 0.4]) == 0


def test_corr_1d():
    assert O.mesh_corr_1d([0.9] * 10) == []
    assert 19 in O.mesh_corr_1d([0.9] * 19 + [-10.0])


def test_flow_calc():
    d = np.array([[1.0, 0.0]] * 4)
    assert abs(O.mesh_flow_calc(TWO, 4, 0, d) - 1.0) < 1e-10
    assert abs(O.mesh_flow_calc(TWO, 4, 0, d, override=[-1.0, 0.0]) + 1.0) < 1e-10
    assert O.mesh_flow_calc(TWO, 4, 0, d, exclude=[1, 2]) == -1.0


def test_r_stats_ignores_non_finite():
    d = np.array([[0.0, 0.0], [np.nan, np.nan], [0.05, 0.0], [0.05, 0.0]])
    r, flag = O.mesh_r_stats(TWO, d)
    assert np.isnan(r[1]) and not flag[1]
    assert abs(r[2] - 0.05) < 1e-15


def test_r_stats_flags_a_magnitude_outlier():
    """UQ + 4 IQR per element neighbourhood: one node displaced 100 px in a
    field of ~1 px on a 5 x 5 grid is flagged, its neighbours are not."""
    nodes, el = band_grid(5, 0.0, 10.0)
    d = np.tile([1.0, 0.0], (25, 1)) + 0.01 * np.random.default_rng(0).normal(size=(25, 2))
    d[12] = [100.0, 0.0]
    _, flag = O.mesh_r_stats(el, d)
    assert list(np.nonzero(flag)[0]) == [12]


def test_project_warp():
    p = np.zeros(12)
    p[4] = p[8] = 1.0
    assert abs(O.mesh_project_warp(p, 2.0, 3.0)[4] - 3.0) < 1e-12
    cx, cy, dx, dy = 2.0, -1.0, 3.0, 5.0
    np.testing.assert_allclose(O.mesh_project_warp(quad_warp(cx, cy), dx, dy), quad_warp(cx + dx, cy + dy),
                               rtol=0, atol=1e-12)


# ---------------------------------------------------------------------------
# affine_field_fit / precondition_warp (mesh.rs:2818-2875)
# ---------------------------------------------------------------------------
def test_affine_field_fit_exact_linear_field():
    u = lambda x, y: 0.1 * x - 0.05 * y + 2.0
    v = lambda x, y: -0.02 * x + 0.08 * y - 1.0
    nodes = np.array([[0.0, 0.0], [10.0, 0.0], [0.0, 10.0], [10.0, 10.0], [5.0, 2.0]])
    p = np.zeros((5, 6))
    p[:, 0] = u(nodes[:, 0], nodes[:, 1])
    p[:, 1] = v(nodes[:, 0], nodes[:, 1])
    fit = O.mesh_affine_field_fit(nodes, [3.0, 7.0], range(5), p)
    np.testing.assert_allclose(fit, [u(3, 7), v(3, 7), 0.1, -0.02, -0.05, 0.08], rtol=0, atol=1e-9)


def test_affine_field_fit_collinear_is_none():
    nodes = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0]])
    assert O.mesh_affine_field_fit(nodes, [1.0, 5.0], [0, 1, 2], np.zeros((3, 6))) is None


def test_affine_field_fit_exact_quadratic_field():
    """Order 2 with six or more neighbours fits u, v to second order: the
    quadratic field of particle.rs:2621 is recovered in every slot."""
    rng = np.random.default_rng(3)
    nodes = rng.uniform(-10, 10, size=(9, 2))
    p = np.array([quad_warp(x, y) for x, y in nodes])
    fit = O.mesh_affine_field_fit(nodes, [1.5, -2.0], range(9), p)
    np.testing.assert_allclose(fit, quad_warp(1.5, -2.0), rtol=0, atol=1e-10)
    # five neighbours: affine only, second-order slots stay zero
    fit5 = O.mesh_affine_field_fit(nodes, [1.5, -2.0], range(5), p)
    assert np.all(fit5[6:] == 0.0)


def test_precondition_warp_thresholds():
    nodes = np.array([[0.0, 0.0], [1.0, 0.0]])
    cz = np.array([0.9, 0.95])
    p = np.array([[1.0, 2.0, 0, 0, 0, 0], [3.0, 4.0, 0, 0, 0, 0]])
    assert np.all(O.mesh_precondition_warp(nodes, [5.0, 5.0], [], cz, p) == 0.0)
    one = O.mesh_precondition_warp(nodes, [5.0, 5.0], [1], cz, p)
    assert abs(one[0] - 3.0) < 1e-12 and abs(one[1] - 4.0) < 1e-12
    # two trusted: the better correlated one, projected
    two = O.mesh_precondition_warp(nodes, [5.0, 5.0], [0, 1], cz, p)
    assert abs(two[0] - 3.0) < 1e-12
    # equal correlation: the last of the equals (Iterator::max_by)
    tie = O.mesh_precondition_warp(nodes, [5.0, 5.0], [0, 1], np.array([0.9, 0.9]), p)
    assert abs(tie[0] - 3.0) < 1e-12


# ---------------------------------------------------------------------------
# greedy colouring (mesh.rs:3176-3193)
# ---------------------------------------------------------------------------
def test_greedy_colour_proper_and_deterministic():
    nodes, el = band_grid(12, 0.0, 5.0)
    n = len(nodes)
    c1 = O.mesh_greedy_colour(el, n, range(n))
    assert c1 == O.mesh_greedy_colour(el, n, range(n))
    for i in range(n):
        for nb in O.mesh_connectivity(el, i, True):
            assert c1[i] != c1[nb]


# ---------------------------------------------------------------------------
# element_area / element_strains (mesh.rs:2521-2650, 3205-3215)
# ---------------------------------------------------------------------------
def test_element_area():
    assert abs(O.mesh_element_area([[0, 0], [1, 0], [0, 1]], [[0, 1, 2]])[0] - 0.5) < 1e-12
    assert np.allclose(np.abs(O.mesh_element_area(UNIT4, TWO)), 0.5, atol=1e-12)
    assert abs(O.mesh_element_area([[0, 0], [2, 0], [0, 2]], [[0, 1, 2]])[0] - 2.0) < 1e-12


def test_element_strains_known_answers():
    st, w = O.mesh_element_strains(UNIT4, TWO, np.tile([3.5, -2.1], (4, 1)))
    assert st == 0 and w.shape == (2, 12)
    assert np.allclose(w[:, 0], 3.5, atol=1e-10) and np.allclose(w[:, 1], -2.1, atol=1e-10)
    assert np.abs(w[:, 2:6]).max() < 1e-10
    st, w = O.mesh_element_strains([[0, 0], [1, 0], [0, 1]], [[0, 1, 2]], [[0, 0], [1, 0], [0, 0]])
    np.testing.assert_allclose(w[0, :6], [1 / 3, 0, 1, 0, 0, 0], rtol=0, atol=1e-10)
    nodes9 = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0], [0.5, 0.0], [0.5, 0.5], [0.0, 0.5],
                       [1.0, 0.5], [0.5, 1.0]])
    st, w = O.mesh_element_strains(nodes9, TWO6, np.tile([1.0, 2.0], (9, 1)))
    assert st == 0 and np.allclose(w[:, :2], [1.0, 2.0], atol=1e-10) and np.abs(w[:, 2:]).max() < 1e-8


def test_element_strains_order2_every_slot():
    nodes = six_node_element()
    d = np.array([quad_warp(x, y)[:2] for x, y in nodes])
    st, w = O.mesh_element_strains(nodes, [[0, 1, 2, 3, 4, 5]], d)
    c = nodes[:3].mean(axis=0)
    assert st == 0
    np.testing.assert_allclose(w[0], quad_warp(*c), rtol=0, atol=1e-12)


def test_element_strains_degenerate():
    st, _ = O.mesh_element_strains([[0, 0], [1, 0], [2, 0]], [[0, 1, 2]], np.zeros((3, 2)))
    assert st == 4  # Error::MeshGeneration


# ---------------------------------------------------------------------------
# the library's host-side mesh functions against the oracle
# ---------------------------------------------------------------------------
def test_library_element_strains_equal_oracle():
    from geopyv_dev import _native, quadratic_mesh, structured_mesh
    rng = np.random.default_rng(7)
    nodes, el = structured_mesh(0.0, 0.0, 60.0, 45.0, 15.0)
    nodes = nodes + rng.uniform(-3, 3, size=nodes.shape)
    for order in (1, 2):
        nd, e = (nodes, el) if order == 1 else quadratic_mesh(nodes, el)
        d = rng.normal(size=(len(nd), 2))
        w_lib = _native.mesh_element_strains(nd, e, d, order)
        st, w_orc = O.mesh_element_strains(nd, e, d)
        assert st == 0
        np.testing.assert_array_equal(w_lib, w_orc)
    with pytest.raises(_native.GeodicError) as ei:
        _native.mesh_element_strains([[0, 0], [1, 0], [2, 0]], [[0, 1, 2]], np.zeros((3, 2)), 1)
    assert ei.value.status == _native.ERR_MESH_GENERATION


def test_structured_and_quadratic_mesh_layout():
    from geopyv_dev import quadratic_mesh, structured_mesh
    nodes, el = structured_mesh(20.0, 20.0, 80.0, 80.0, 7.5)
    n_ref, el_ref = band_grid()
    np.testing.assert_array_equal(nodes, n_ref)
    np.testing.assert_array_equal(el, el_ref)
    nd6, el6 = quadratic_mesh(nodes, el)
    assert el6.shape == (len(el), 6)
    for row in el6:
        for k, (a, b) in enumerate(((0, 1), (1, 2), (2, 0))):
            np.testing.assert_array_equal(nd6[row[3 + k]], 0.5 * (nd6[row[a]] + nd6[row[b]]))
    # a corner's edge neighbours are midside nodes only
    assert set(O.mesh_connectivity(el6, 40, False)) <= set(range(len(nodes), len(nd6)))


def test_node_templates_equal_mask_update():
    """Per-node pixel lists with a region of interest: the interior shortcut
    and the clipped lists both equal mask_update (masks.rs:199-263) node by
    node, as subset_at builds them (mesh.rs:575-593)."""
    from geopyv_dev import Mask, _native
    from geopyv_dev.mesh import node_templates, roi_mask
    mask = roi_mask((60, 80), exclusions=[[[30, 20], [50, 20], [50, 40], [30, 40]]])
    assert mask[30, 40] == 0 and mask[5, 5] == 1
    rng = np.random.default_rng(1)
    nodes = np.concatenate([rng.uniform(-2, 82, size=(60, 2)) * [1, 0.75], [[10.0, 10.0], [79.5, 59.5]]])
    lm = Mask("local", shape="circle", size=5)
    tpl_index, start, xy = node_templates(lm, nodes, mask)
    for k, c in enumerate(nodes):
        t = tpl_index[k]
        want = _native.mask_update("circle", 5, c, mask)
        if len(want) == 0:
            continue
        np.testing.assert_array_equal(xy[start[t]:start[t + 1]], want)


# ---------------------------------------------------------------------------
# the oracle's mesh solve on the reference's synthetic fixtures
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def band_case():
    tex = texture(100)
    fc = O.get_c(tex, 15)
    gc = O.get_c(shift_band(tex, 50, 1.5, -1.5), 15)
    return fc, gc, tex.shape


@pytest.fixture(scope="module")
def uniform_case():
    tex = texture(100)
    return O.get_c(tex, 15), O.get_c(shift_uniform(tex, 1.5), 15), tex.shape


def _solve(case, radius, **kw):
    fc, gc, shape = case
    nodes, el = band_grid()
    xy, start = circle_template(radius)
    st, r = O.mesh_solve(fc, gc, shape, 15, nodes, el, np.zeros(len(nodes), np.int32), xy, start,
                         kw.pop("seed", [50.0, 30.0]), np.zeros(6), subset_order=1, tolerance=0.5,
                         seed_tolerance=0.5, override=True, **kw)
    assert st == 0
    return r


def test_mesh_solve_recovers_the_band(band_case):
    r = _solve(band_case, 6)
    u = r["p"][:, 0].reshape(9, 9)
    assert r["seed_node"] == 13 and r["quality_ok"].all()
    assert np.abs(u[:4] - 1.5).max() < 0.05 and np.abs(u[5:] + 1.5).max() < 0.05
    v = r["p"][:, 1].reshape(9, 9)
    assert np.abs(np.delete(v, 4, axis=0)).max() < 0.05  # the row on the step straddles it


def test_layer_rg_matches_rg_within_tier_c(uniform_case):
    """mesh.rs:3110-3136 (and batch_factor 1, eps 0: 3153-3172)."""
    a = _solve(uniform_case, 8, seed=[50.0, 50.0])
    for bf, eps in ((4, 0.02), (1, 0.0)):
        b = _solve(uniform_case, 8, seed=[50.0, 50.0], preconditioning="layer-RG", batch_factor=bf,
                   root_rel_eps=eps, workers=8)
        assert a["seed_node"] == b["seed_node"]
        scale = np.maximum(np.abs(a["p"][:, :2]).max(axis=1), 1.0)
        assert (np.abs(a["p"][:, 0] - b["p"][:, 0]) <= 1e-5 * scale).all()
        assert (np.abs(a["p"][:, 1] - b["p"][:, 1]) <= 1e-5 * scale).all()
        assert np.abs(a["c_zncc"] - b["c_zncc"]).max() <= 1e-5
        assert (a["quality_ok"] == b["quality_ok"]).all()
    assert np.abs(a["p"][:, 0] - 1.5).max() < 0.05


def test_layer_rg_is_deterministic(band_case):
    """mesh.rs:3138-3151: byte-identical p and c_zncc over runs."""
    runs = [_solve(band_case, 6, preconditioning="layer-RG", workers=4, threads=t) for t in (1, 3, 8)]
    for r in runs[1:]:
        np.testing.assert_array_equal(r["p"], runs[0]["p"])
        np.testing.assert_array_equal(r["c_zncc"], runs[0]["c_zncc"])
        np.testing.assert_array_equal(r["iterations"], runs[0]["iterations"])


# ---------------------------------------------------------------------------
# the sequence's guide step: warp_increment / element_locator
# (particle.rs:254-320, 968-1013; warp_convention_tests 2648-2669)
# ---------------------------------------------------------------------------
def test_warp_increment_every_slot():
    nodes = six_node_element()
    d = np.array([quad_warp(x, y)[:2] for x, y in nodes])
    np.testing.assert_allclose(O.warp_increment([4.0, 4.0], nodes, d, 2), quad_warp(4.0, 4.0), rtol=0, atol=1e-12)
    lin = lambda x, y: [Q[0] + Q[2] * x + Q[4] * y, Q[1] + Q[3] * x + Q[5] * y]
    d1 = np.array([lin(x, y) for x, y in nodes[:3]])
    np.testing.assert_allclose(O.warp_increment([4.0, 4.0], nodes[:3], d1, 1), lin(4.0, 4.0) + Q[2:6], rtol=0,
                               atol=1e-12)


def test_library_guide_step_equals_oracle():
    from geopyv_dev import sequence as S
    rng = np.random.default_rng(11)
    for order in (1, 2):
        for _ in range(20):
            nodes = six_node_element() + rng.normal(scale=0.3, size=(6, 2))
            d = rng.normal(size=(6, 2))
            c = nodes[:3].mean(axis=0) + rng.normal(scale=0.5, size=2)
            k = 3 if order == 1 else 6
            np.testing.assert_allclose(S.warp_increment(c, nodes[:k], d[:k], order),
                                       O.warp_increment(c, nodes[:k], d[:k], order), rtol=1e-12, atol=1e-12)


def test_element_locator():
    from geopyv_dev import sequence as S
    from geopyv_dev.mesh import element_centroids
    nodes, el = band_grid(6, 0.0, 10.0)
    cen = element_centroids(nodes, el)
    for c in ([12.0, 13.0], [10.0, 10.0], [25.0, 20.0], [49.9, 0.1], [-5.0, 20.0], [70.0, 70.0]):
        got = S.element_locator(np.array(c), nodes, el, cen)
        assert got == O.element_locator(c, nodes, el, cen)
        i, adrift = got
        inside = c[0] >= 0 and c[1] >= 0 and c[0] <= 50 and c[1] <= 50
        assert adrift == (not inside)
    # a point inside: its element's area coordinates are all >= 0
    i, adrift = S.element_locator(np.array([12.0, 13.0]), nodes, el, cen)
    z, e, t, _ = S._local([12.0, 13.0], nodes[el[i]])
    assert min(z, e, t) >= 0 and not adrift


def test_image_files_in_reference_order(tmp_path):
    from geopyv_dev.sequence import image_files
    for name in ("frame_10.png", "frame_2.jpg", "img3_001.JPEG", "notes.txt", "frame_0.png"):
        (tmp_path / name).write_bytes(b"")
    got = [p.split("/")[-1] for p in image_files(str(tmp_path))]
    assert got == ["frame_0.png", "img3_001.JPEG", "frame_2.jpg", "frame_10.png"]

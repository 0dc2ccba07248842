"""Pins the FP64 oracle: the reference's own known-answer tests, closed
forms, invariants and brute force (CPU only)."""

import math

import numpy as np
import pytest

import oracle as O

PAD_LENGTHS = [552, 2088, 4136, 3040, 1064]


# ---------------------------------------------------------------------------
# FFT (rustfft stand-in)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 7, 12, 47, 94, 552, 1064])
def test_fft_matches_direct_dft(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    k = np.arange(n)
    D = np.exp(-2j * np.pi * np.outer(k, k) / n)
    ref = D @ x
    np.testing.assert_allclose(O.fft(x, -1), ref, rtol=0, atol=1e-11 * max(1, np.abs(ref).max()))
    np.testing.assert_allclose(O.fft(x, +1), np.conj(D) @ x, rtol=0, atol=1e-11 * max(1, np.abs(ref).max()))


@pytest.mark.parametrize("n", PAD_LENGTHS)
def test_fft_config_lengths(n):
    x = np.random.default_rng(1).normal(size=n) + 0j
    ref = np.fft.fft(x)
    assert np.abs(O.fft(x) - ref).max() < 1e-12 * np.abs(ref).max()


# ---------------------------------------------------------------------------
# get_c (image.rs:252-327)
# ---------------------------------------------------------------------------
K5 = np.array([1 / 120, 13 / 60, 11 / 20, 13 / 60, 1 / 120])


def _circ_conv_rows(C):
    # zero-phase 5-tap kernel: out[j] = Σ_d k[d+2] C[(j+d) mod n]
    return sum(K5[d + 2] * np.roll(C, -d, axis=1) for d in range(-2, 3))


def test_get_c_constant():
    # image.rs:538-550
    c = O.get_c(np.full((20, 25), 77.0), 5)
    assert c.shape == (30, 35)
    assert np.max(np.abs(c - 77.0)) < 1e-8 * 77.0


def test_pad_replicate_corners():
    # image.rs:602-615
    arr = np.arange(12, dtype=np.float64).reshape(3, 4)
    p = O.pad_replicate(arr, 2)
    assert p.shape == (7, 8)
    assert p[0, 0] == 0.0 and p[-1, -1] == 11.0 and p[2, 2] == arr[0, 0]


@pytest.mark.parametrize("shape,border", [((31, 45), 5), ((64, 64), 20), ((17, 90), 3)])
def test_get_c_reconvolution_invariant(shape, border):
    """Circularly re-convolving C with the kernel along rows and columns
    gives back the padded image: C is the unique solution (cond <= 7.5)."""
    img = np.random.default_rng(2).uniform(0, 255, shape)
    C = O.get_c(img, border)
    P = O.pad_replicate(img, border)
    back = _circ_conv_rows(_circ_conv_rows(C).T).T
    assert np.max(np.abs(back - P)) < 1e-10


def test_get_c_dense_circulant_solve():
    img = np.random.default_rng(4).uniform(0, 255, (9, 13))
    b = 3
    P = O.pad_replicate(img, b)
    hp, wp = P.shape

    def circ(n):
        M = np.zeros((n, n))
        for j in range(n):
            for d in range(-2, 3):
                M[j, (j + d) % n] = K5[d + 2]
        return M

    C_dense = np.linalg.solve(circ(hp), np.linalg.solve(circ(wp), P.T).T)
    assert np.max(np.abs(O.get_c(img, b) - C_dense)) < 1e-11


def _fir_taps(n, half=52):
    """Circular-FIR form of the spectral division: the inverse DFT of 1/K
    (numpy's FFT, independent of the oracle's), taps -half..half, or one
    tap per residue when the line is shorter than 2*half + 1."""
    w = 2.0 * np.pi * np.arange(n) / n
    h = np.real(np.fft.ifft(1.0 / (11 / 20 + (13 / 30) * np.cos(w) + (1 / 60) * np.cos(2 * w))))
    k1 = half if n >= 2 * half + 1 else n // 2
    nt = 2 * half + 1 if n >= 2 * half + 1 else n
    return np.array([h[(t - k1) % n] for t in range(nt)]), k1


def _fir_rows(P):
    n = P.shape[1]
    taps, k1 = _fir_taps(n)
    out = np.zeros_like(P)
    for t, v in enumerate(taps):
        out = out + v * np.roll(P, -(t - k1), axis=1)
    return out


@pytest.mark.parametrize("shape,border", [((64, 65), 20), ((20, 25), 5), ((150, 40), 20)])
def test_get_c_equals_truncated_circular_fir(shape, border):
    """The device's default prefilter applies each line's spectral division
    as a circular FIR with the inverse DFT of 1/K truncated to |k| <= 52
    (taps fall by the pole 0.4306 per step): it matches the oracle's
    FFT division to 1e-10, exactly where a line is shorter than 105 (padded
    lengths here: 104, 105, 30, 35, 190, 80)."""
    img = np.random.default_rng(9).uniform(0, 255, shape)
    P = O.pad_replicate(img, border)
    C = _fir_rows(_fir_rows(P).T).T
    assert np.max(np.abs(C - O.get_c(img, border))) < 1e-10


def test_fir_taps_decay_and_sum():
    taps, k1 = _fir_taps(2088)
    assert k1 == 52 and len(taps) == 105
    assert abs(taps.sum() - 1.0) < 1e-12          # K(0) = 1: the inverse keeps a constant line
    # the last kept tap is at numpy's FFT round-off (~1e-17); the true value
    # is below 1e-18, as 0.4306**52 times the leading coefficient shows
    assert abs(taps[k1 + 52]) < 1e-16 and abs(taps[k1 + 40]) > 1e-15
    ratio = abs(taps[k1 + 21] / taps[k1 + 20])
    assert abs(ratio - 0.430575347) < 1e-7     # the larger pole of the quintic B-spline


# ---------------------------------------------------------------------------
# eval / eval_grad (image.rs:115-230)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def wavy():
    img = np.fromfunction(lambda y, x: 60.0 + np.sin(x * 0.9) * 20.0 + np.cos(y * 0.7) * 15.0, (24, 28))
    return img, O.get_c(img, 6)


def test_q_row_sums():
    from geopyv_dev._bspline import Q
    np.testing.assert_allclose(Q.sum(axis=1), [1, 0, 0, 0, 0, 0], atol=1e-15)


def test_eval_matches_qcqt_blocks(wavy):
    # image.rs:570-600
    img, C = wavy
    p = lambda t: np.array([1, t, t * t, t ** 3, t ** 4, t ** 5])
    d = lambda t: np.array([0, 1, 2 * t, 3 * t * t, 4 * t ** 3, 5 * t ** 4])
    for x, y in [(3.25, 4.75), (10.0, 12.0), (0.0, 0.0), (15.999, 9.001), (7.5, 0.5), (26.3, 22.9)]:
        xf, yf = int(math.floor(x)), int(math.floor(y))
        B = O.qcqt_block(C, img.shape, 6, yf, xf)
        v, gx, gy = O.eval_grad_points(C, img.shape, 6, [[x, y]])[0]
        assert abs(v - p(y - yf) @ B @ p(x - xf)) < 1e-11
        assert abs(gx - p(y - yf) @ B @ d(x - xf)) < 1e-11
        assert abs(gy - d(y - yf) @ B @ p(x - xf)) < 1e-11
        assert O.eval_points(C, img.shape, 6, [[x, y]])[0] == v  # bit-equal to eval


def test_eval_at_integer_pixels_is_the_pixel(wavy):
    img, C = wavy
    ys, xs = np.mgrid[0:24, 0:28]
    v = O.eval_points(C, img.shape, 6, np.column_stack([xs.ravel(), ys.ravel()]).astype(float))
    assert np.max(np.abs(v - img.ravel())) < 1e-10


def test_gradient_is_derivative_at_fractional_position():
    # image.rs:634-647
    a = np.fromfunction(lambda r, c: 120 + 60 * np.sin(2 * np.pi * r / 7 + 0.3) * np.cos(2 * np.pi * c / 9), (60, 60))
    C = O.get_c(a, 8)
    h = 1e-5
    for x, y in [(30.0, 30.0), (30.25, 30.75), (30.5, 29.5), (31.9, 28.1)]:
        _, gx, gy = O.eval_grad_points(C, a.shape, 8, [[x, y]])[0]
        e = lambda xx, yy: O.eval_points(C, a.shape, 8, [[xx, yy]])[0]
        assert abs(gx - (e(x + h, y) - e(x - h, y)) / (2 * h)) < 1e-4
        assert abs(gy - (e(x, y + h) - e(x, y - h)) / (2 * h)) < 1e-4


def test_eval_constant_and_clamp():
    # subset.rs:1813-1821, 1844-1868
    C = O.get_c(np.full((4, 4), 42.0), 5)
    for x, y in [(0.0, 0.0), (1.5, 2.25), (3.0, 3.0)]:
        assert abs(O.eval_points(C, (4, 4), 5, [[x, y]])[0] - 42.0) < 1e-8 * 42
    img = np.fromfunction(lambda y, x: 100 + np.sin(x * 1.3) * 20 + np.cos(y * 0.8) * 10, (8, 9))
    C = O.get_c(img, 5)
    far = O.eval_points(C, img.shape, 5, [[1e6, 0.0]])[0]
    edge = O.eval_points(C, img.shape, 5, [[9.0 - 1e-9, 0.0]])[0]
    assert np.isfinite(far) and abs(far) < 200 and abs(far - edge) < 1e-3
    g = O.eval_grad_points(C, img.shape, 5, [[1e6, -1e6]])[0]
    assert np.all(np.isfinite(g)) and np.all(np.abs(g[1:]) < 200)


# ---------------------------------------------------------------------------
# masks (masks.rs:58-263; tests 274-389; test_templates.py)
# ---------------------------------------------------------------------------
def test_circle_5():
    c, sm = O.mask_local("circle", 5)
    assert len(c) == 81 and list(c[0]) == [0, -5] and list(c[-1]) == [0, 5]
    assert sm.shape == (11, 11) and sm[5, 5] == 1 and sm[0, 5] == 1 and sm[5, 0] == 1
    assert sm[0, 0] == 0 and sm[0, 10] == 0 and sm[10, 0] == 0 and sm[10, 10] == 0


def test_semicircle_and_square_5():
    c, sm = O.mask_local("semicircle", 5)
    assert len(c) == 46 and np.all(c[:, 1] >= 0)
    assert sm[5, 5] == 1 and sm[10, 5] == 1 and sm[0, 5] == 0 and not sm[:5].any()
    c, sm = O.mask_local("square", 5)
    assert len(c) == 121 and list(c[0]) == [-5, -5] and list(c[-1]) == [5, 5] and np.all(sm == 1)


@pytest.mark.parametrize("r,n", [(15, 709), (20, 1257), (25, 1961), (30, 2821), (40, 5025)])
def test_circle_counts_gauss(r, n):
    # n_px(r) = #{x² + y² <= r²}; 1961 for r = 25 is test_subset.py:49
    x = np.arange(-r, r + 1)
    assert len(O.mask_local("circle", r)[0]) == n == int(np.sum(x[:, None] ** 2 + x[None, :] ** 2 <= r * r))


def test_mask_update_full_and_zero():
    assert len(O.mask_update("square", 2, [5.0, 5.0], np.ones((20, 20)))) == 25
    assert len(O.mask_update("square", 2, [5.0, 5.0], np.zeros((20, 20)))) == 0
    c = O.mask_update("square", 1, [3.0, 3.0], np.ones((10, 10)))
    assert c.shape == (9, 2) and list(c[0]) == [-1, -1]


def test_mask_update_edge_known_answers():
    """Vectors derived by hand from masks.rs:204-258.

    square(2) about (2, 2) in a 5x5 all-ones mask: the exclusive ends
    ux = uy = 5 reach the edge, so the last column and row are dropped
    (225-232): 4x4 = 16 pixels, offsets -2..1 in both directions.
    About (2.9, 1.7) the centre truncates to (2, 1): ly = -1 pads one row
    (dy = -2 absent), uy = 4 < 5 keeps dy up to +2, and ux = 5 drops dx = 2.
    """
    c = O.mask_update("square", 2, [2.0, 2.0], np.ones((5, 5)))
    exp = np.array([[x, y] for y in range(-2, 2) for x in range(-2, 2)], dtype=float)
    assert np.array_equal(c, exp)
    c = O.mask_update("square", 2, [2.9, 1.7], np.ones((5, 5)))
    exp = np.array([[x, y] for y in range(-1, 3) for x in range(-2, 2)], dtype=float)
    assert np.array_equal(c, exp)
    # interior window one pixel short of the edge keeps everything
    assert len(O.mask_update("square", 2, [2.0, 2.0], np.ones((6, 6)))) == 25
    # circle(2) at the corner (0, 0) of an all-ones 10x10 mask: x, y >= 0 only
    c = O.mask_update("circle", 2, [0.0, 0.0], np.ones((10, 10)))
    exp = np.array([[x, y] for y in range(0, 3) for x in range(0, 3) if x * x + y * y <= 4], dtype=float)
    assert np.array_equal(c, exp)
    # masked column: a zero column at x = 6 removes dx = 1 about centre 5
    m = np.ones((20, 20))
    m[:, 6] = 0
    c = O.mask_update("square", 1, [5.0, 5.0], m)
    assert np.array_equal(c, np.array([[-1, -1], [0, -1], [-1, 0], [0, 0], [-1, 1], [0, 1]], dtype=float))


# ---------------------------------------------------------------------------
# warp algebra (subset.rs:236-669; tests 1425-1456, 1701-1810, 2211-2324)
# ---------------------------------------------------------------------------
QW = np.array([0.1, -0.2, 0.02, 0.07, 0.03, -0.01, 0.004, 0.008, 0.005, 0.009, 0.006, -0.003])  # particle.rs:2621


def quad_warp(x, y):
    q = QW
    return np.array([q[0] + q[2] * x + q[4] * y + 0.5 * q[6] * x * x + q[8] * x * y + 0.5 * q[10] * y * y,
                     q[1] + q[3] * x + q[5] * y + 0.5 * q[7] * x * x + q[9] * x * y + 0.5 * q[11] * y * y])


PTS = np.array([[13.0, 7.0], [4.0, 15.0], [9.5, 9.5], [16.0, 2.0]])


def test_norms():
    assert abs(O.convergence_norm([3.0, 4.0, 0, 0, 0, 0], 1.0) - 5.0) < 1e-12
    assert O.convergence_norm([0.0] * 6, 44.3) == 0.0
    assert abs(O.convergence_norm([0, 0, 2.0, 0, 0, 0], 5.0) - 10.0) < 1e-12
    assert abs(O.convergence_norm([0.3, 0.4], 123.0) - 0.5) < 1e-12
    d = np.zeros(12)
    d[7] = 2.0
    assert abs(O.convergence_norm(d, 3.0) - 9.0) < 1e-12  # ½·2·3²


def test_compose_golden():
    # subset.rs:1790-1810
    out = O.compose_icgn([1.0, 0.5, 0.01, 0.02, -0.01, 0.005], [0.1, 0.05, 0.001, 0.002, -0.001, 0.0005])
    np.testing.assert_allclose(out, [0.89954843, 0.44797691, 0.00900896, 0.01797300, -0.00898650, 0.00451572],
                               atol=1e-6)


def test_compose_identity_order0_translation():
    for p in (QW[:6], QW):
        np.testing.assert_allclose(O.compose_icgn(p, np.zeros(len(p))), p, atol=1e-14)
    np.testing.assert_allclose(O.compose_icgn([1.0, 2.0], [0.5, 0.3]), [0.5, 1.7], atol=1e-12)
    out = O.compose_icgn(np.zeros(6), [0.5, 0.3, 0, 0, 0, 0])
    np.testing.assert_allclose(out, [-0.5, -0.3, 0, 0, 0, 0], atol=1e-12)
    assert len(O.compose_icgn(np.zeros(12), [0.5, 0.3] + [0.01] * 4 + [0.001] * 6)) == 12


def test_compose_undoes_update():
    c = [10.0, 10.0]
    dp = np.array([0.05, -0.02, 0.004, -0.006, 0.003, 0.008])
    for p in (QW[:6], np.concatenate([QW[:6], np.zeros(6)])):
        d = dp if len(p) == 6 else np.concatenate([dp, np.zeros(6)])
        pn = O.compose_icgn(p, d)
        via = O.apply_warp(c, pn, O.apply_warp(c, d, PTS))
        np.testing.assert_allclose(via, O.apply_warp(c, p, PTS), atol=1e-12)


def test_compose_order2_inverse_is_second_order():
    dp = np.array([0.2, -0.1, 0.02, 0.05, -0.03, 0.01, 0.002, -0.001, 0.0015, 0.001, -0.002, 0.0012])
    inv = O.compose_icgn(np.zeros(12), dp)

    def err(r):
        x = np.array([[0.6 * r, -0.8 * r]])
        back = O.apply_warp([0.0, 0.0], inv, O.apply_warp([0.0, 0.0], dp, x))
        return np.hypot(*(back - x)[0])

    for r in (8.0, 4.0, 2.0):
        assert err(r) / err(r / 2) > 6.5


def test_apply_warp_every_slot():
    c = [10.0, 10.0]
    for order in (1, 2):
        p = QW[:6] if order == 1 else QW
        g = O.apply_warp(c, p, PTS)
        for i, (x, y) in enumerate(PTS):
            dx, dy = x - c[0], y - c[1]
            want = quad_warp(dx, dy)
            if order == 1:
                want = np.array([QW[0] + QW[2] * dx + QW[4] * dy, QW[1] + QW[3] * dx + QW[5] * dy])
            np.testing.assert_allclose(g[i] - PTS[i], want, atol=1e-12)
    np.testing.assert_allclose(O.apply_warp([10, 20], [1.5, -0.5], [[10, 20], [15, 25]]), [[11.5, 19.5], [16.5, 24.5]])


def test_sdi_columns_are_warp_derivatives():
    c = [10.0, 10.0]
    grad = np.array([[0.3, -0.7], [1.1, 0.2], [-0.4, 0.9], [0.6, 0.5]])
    sdi = O.steepest_descent(c, PTS, grad, 2)
    h = 1e-6
    for k in range(12):
        pp = QW.copy()
        pp[k] += h
        d = (O.apply_warp(c, pp, PTS) - O.apply_warp(c, QW, PTS)) / h
        np.testing.assert_allclose(sdi[:, k], np.sum(grad * d, axis=1), atol=1e-6)
    s0 = O.steepest_descent([0, 0], [[3, 4], [-1, 2]], [[0.1, 0.2], [0.3, 0.4]], 0)
    assert s0.shape == (2, 2) and np.allclose(s0, [[0.1, 0.2], [0.3, 0.4]])


def test_hessian_and_inverse():
    sdi = np.cos(np.arange(37 * 6).reshape(37, 6) * 0.13) * (1 + np.arange(6) * 0.1)
    H = O.hessian(sdi)
    np.testing.assert_allclose(H, sdi.T @ sdi, rtol=1e-13)
    for m in (2, 3, 6, 12):
        A = np.random.default_rng(m).normal(size=(m, m)) + m * np.eye(m)
        np.testing.assert_allclose(O.inverse(A) @ A, np.eye(m), atol=1e-12)
    assert O.inverse(np.zeros((6, 6))) is None


# ---------------------------------------------------------------------------
# warp-adequacy diagnostic (subset.rs:1460-1616)
# ---------------------------------------------------------------------------
def test_zn_slice():
    zn = O.zn_slice([1.0, 2.0, 3.0])
    sd = math.sqrt(2 / 3)
    np.testing.assert_allclose(zn, [-1 / sd, 0, 1 / sd], atol=1e-12)
    assert np.all(np.abs(O.zn_slice([5.0] * 4)) < 1e-6)


@pytest.mark.parametrize("modes", [
    [[1, 0, 0], [1, 0, 0], [0, 0, 1], [0, 0, 1]],
    [[1, 0], [-1, 0], [0, 1], [0, -1]],
    [[1, 0, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, -1]],
])
def test_eta_component_hand_computed(modes):
    eta = O.eta_component(np.array(modes, dtype=float), [1.0, -1.0, 1.0, -1.0], [1.0, 0.0, 0.0, 0.0])
    assert abs(eta - math.sqrt(2) / 4) < 1e-12


def test_diagnostic_zero_for_identical_and_none_for_order3():
    fc = np.array([[1.0, 0.0], [-1.0, 0.0], [0.0, 1.0], [0.0, -1.0]])
    gf = np.array([[1.0, 0.5], [-1.0, 0.3], [0.2, 1.0], [-0.2, -1.0]])
    f = np.array([10.0, 12.0, 9.0, 11.0])
    fm = f.mean()
    df = math.sqrt(np.sum((f - fm) ** 2))
    for order in (0, 1, 2):
        eu, ev = O.omitted_mode_diagnostic([0, 0], order, fc, gf, f, fm, df, f, fm, df)
        assert abs(eu) < 1e-10 and abs(ev) < 1e-10
    assert O.omitted_mode_diagnostic([0, 0], 3, fc, gf, f, fm, df, f, fm, df) is None


def test_unrolled_sum():
    x = np.random.default_rng(0).uniform(0, 255, 1961)
    assert abs(O.unrolled_sum(x) - math.fsum(x)) < 1e-9


# ---------------------------------------------------------------------------
# end-to-end solves on synthetic speckle
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def speckle_pair():
    from geopyv_dev import workloads
    sp = workloads.Speckle((128, 128), 400, 2.0, seed=0)
    f = sp.reference(0.0).astype(np.float64)
    g = sp.target(workloads.Translation(0.37, -0.52), 0.0).astype(np.float64)
    return f, g, O.get_c(f, 20), O.get_c(g, 20)


@pytest.mark.parametrize("order", [0, 1, 2])
def test_translation_recovered(speckle_pair, order):
    f, g, Cf, Cg = speckle_pair
    off, _ = O.mask_local("circle", 15)
    r = O.subset_solve(Cf, Cg, f.shape, 20, [64.0, 64.0], off, order, max_norm=1e-5)
    assert r["status"] == 0 and r["converged"] and r["quality_ok"]
    assert abs(r["p"][0] - 0.37) < 0.01 and abs(r["p"][1] + 0.52) < 0.01
    rf = O.subset_solve(Cf, Cg, f.shape, 20, [64.0, 64.0], off, order, method="fagn", max_norm=1e-5)
    # FAGN ~ ICGN (test_subset.py:590-604 pattern)
    assert abs(rf["c_zncc"] - r["c_zncc"]) < 1e-4 and np.max(np.abs(rf["p"][:2] - r["p"][:2])) < 1e-3


def test_identical_images_znssd_zero(speckle_pair):
    f, _, Cf, _ = speckle_pair
    off, _ = O.mask_local("circle", 15)
    r = O.subset_solve(Cf, Cf, f.shape, 20, [64.0, 64.0], off, 1, max_norm=1e-5)
    assert r["c_znssd"] < 1e-20 and r["iterations"] == 1 and np.all(np.abs(r["p"]) < 1e-12)


def test_quality_and_iteration_limits(speckle_pair):
    f, g, Cf, Cg = speckle_pair
    off, _ = O.mask_local("circle", 15)
    r = O.subset_solve(Cf, Cg, f.shape, 20, [64.0, 64.0], off, 1, tolerance=0.0, max_norm=1e-8, max_iterations=1)
    assert r["iterations"] == 1 and not r["converged"] and not r["quality_ok"]
    r = O.subset_solve(Cf, Cg, f.shape, 20, [64.0, 64.0], off, 1, tolerance=2.0)
    assert r["converged"] and not r["quality_ok"]
    r = O.subset_solve(Cf, Cg, f.shape, 20, [64.0, 64.0], off, 1, max_iterations=0)
    assert r["iterations"] == 0 and r["c_zncc"] == 0.0 and r["c_znssd"] == 4.0


def test_featureless_subset_is_invalid_input():
    C = O.get_c(np.zeros((64, 64)), 20)  # C exactly 0, so delta_f == 0 exactly
    off, _ = O.mask_local("circle", 5)
    r = O.subset_solve(C, C, (64, 64), 20, [32.0, 32.0], off, 1)
    assert r["status"] == 3


# ---------------------------------------------------------------------------
# Subset::new statistics (subset.rs:885-899): f_m = mean f, delta_f =
# sqrt(sum (f - f_m)^2), SSSIG = sum 1/2 (gx^2 + gy^2), sigma =
# sqrt(sum (f - f_m)^2 / n), checked against NumPy on the oracle's own f and
# grad_f (themselves pinned through eval / eval_grad above) and against a
# hand-computed four-pixel case (the pattern of test_subset.py:49-53).
# ---------------------------------------------------------------------------
def test_subset_new_statistics_match_numpy(speckle_pair):
    f, g, Cf, Cg = speckle_pair
    for size, centre in ((15, [64.0, 64.0]), (9, [40.3, 71.8])):
        off, _ = O.mask_local("circle", size)
        r = O.subset_solve(Cf, Cg, f.shape, 20, centre, off, 1, max_iterations=0, want_arrays=True)
        fv, gr = r["f"], r["grad_f"]
        n = len(off)
        assert r["n_px"] == n and len(fv) == n
        f_m = np.mean(fv)
        d = fv - f_m
        assert abs(r["f_m"] - f_m) <= 1e-12 * abs(f_m)
        assert abs(r["delta_f"] - np.sqrt(np.sum(d * d))) <= 1e-12 * r["delta_f"]
        assert abs(r["sssig"] - np.sum(0.5 * (gr[:, 0] ** 2 + gr[:, 1] ** 2))) <= 1e-12 * r["sssig"]
        assert abs(r["sigma"] - np.sqrt(np.sum(d * d) / n)) <= 1e-12 * r["sigma"]
        # the f_coords are the offsets about the centre (subset.rs:866-871)
        assert np.array_equal(r["f_coords"], off + np.asarray(centre))


def test_subset_new_statistics_four_pixel_known_answer():
    # I(x, y) = 2x + 3y + 10 on a 128^2 image: a B-spline interpolant
    # reproduces the ramp far from the padded border, so at the four pixels
    # about (64, 64) f = 330, 332, 333, 335 and grad f = (2, 3) exactly:
    # f_m = 332.5, sum (f - f_m)^2 = 13, delta_f = sqrt(13), sigma =
    # sqrt(13 / 4), SSSIG = 4 * 1/2 * (4 + 9) = 26.
    yy, xx = np.mgrid[0:128, 0:128].astype(np.float64)
    img = 2.0 * xx + 3.0 * yy + 10.0
    C = O.get_c(img, 20)
    off = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]])
    r = O.subset_solve(C, C, img.shape, 20, [64.0, 64.0], off, 1, max_iterations=0, want_arrays=True)
    assert r["status"] == 0 and r["n_px"] == 4
    assert np.allclose(r["f"], [330.0, 332.0, 333.0, 335.0], rtol=0, atol=1e-9)
    assert np.allclose(r["grad_f"], [[2.0, 3.0]] * 4, rtol=0, atol=1e-9)
    assert abs(r["f_m"] - 332.5) < 1e-9
    assert abs(r["delta_f"] - np.sqrt(13.0)) < 1e-9
    assert abs(r["sigma"] - np.sqrt(13.0 / 4.0)) < 1e-9
    assert abs(r["sssig"] - 26.0) < 1e-8


# ---------------------------------------------------------------------------
# Subset::residual_map (subset.rs:1350-1392)
# ---------------------------------------------------------------------------
def test_residual_map_zero_for_identical_images(speckle_pair):
    f, _, Cf, _ = speckle_pair
    off, _ = O.mask_local("circle", 15)
    e = O.residual_map(Cf, Cf, f.shape, 20, [64.0, 64.0], off, 1, np.zeros(6))
    assert len(e) == len(off) and np.all(e == 0.0)


@pytest.mark.parametrize("order", [0, 1, 2])
def test_residual_map_sums_match_the_solver(speckle_pair, order):
    # Σe = 0 (both terms are zero-mean) and Σe² is the ZNSSD that a
    # one-iteration solve started at p computes in its own pass
    f, g, Cf, Cg = speckle_pair
    off, _ = O.mask_local("circle", 15)
    m = {0: 2, 1: 6, 2: 12}[order]
    p = np.zeros(m)
    p[:2] = [0.3, -0.45]
    if order >= 1:
        p[2:6] = [0.004, -0.002, 0.001, 0.003]
    e = O.residual_map(Cf, Cg, f.shape, 20, [63.4, 64.7], off, order, p)
    r = O.subset_solve(Cf, Cg, f.shape, 20, [63.4, 64.7], off, order, p0=p, max_iterations=1)
    assert abs(np.sum(e)) < 1e-12
    assert abs(np.sum(e * e) - r["c_znssd"]) < 1e-12 * max(1.0, r["c_znssd"])
    # at the converged warp the residual is smaller than at p = 0
    rs = O.subset_solve(Cf, Cg, f.shape, 20, [63.4, 64.7], off, order, max_norm=1e-8)
    e0 = O.residual_map(Cf, Cg, f.shape, 20, [63.4, 64.7], off, order, np.zeros(m))
    es = O.residual_map(Cf, Cg, f.shape, 20, [63.4, 64.7], off, order, rs["p"])
    assert np.sum(es * es) < 0.5 * np.sum(e0 * e0)

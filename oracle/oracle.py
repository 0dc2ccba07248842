"""ctypes binding of the FP64 CPU oracle (oracle/dic_oracle.cpp and
oracle/mesh_oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this module, and only as the
checker or the timed CPU baseline.  The shipped library never calls it.

Parity status of each oracle function (what pins it without running the
Rust crate, which cannot be built here):

=====================  ==================================================
function               pinned by
=====================  ==================================================
fft                    direct O(n^2) DFT and numpy.fft on every config length
load_convert           BT.709 integer grey known answers; constant frame
                       kept; a single bright pixel gives the rounded
                       outer product of the taps; reflect-101 borders and
                       the interior against an independent numpy reading
                       of image.rs:450-527
get_c                  constant image -> constant C (image.rs:538-550);
                       circular re-convolution of C with the 5-tap kernel
                       reproduces the padded image; dense circulant solve
                       on small images; eval at integer pixels = pixel
eval / eval_grad       QCQT-block formula (image.rs:570-600); value of
                       eval_grad bit-equal to eval; gradient vs central
                       differences (image.rs:634-647); constant image;
                       far out-of-range saturation (subset.rs:1844-1868)
mask_local             reference Rust and Python golden counts and first /
                       last coordinates (masks.rs:274-371,
                       test_templates.py); Gauss circle counts
mask_update            full/zero masks (masks.rs:373-389,
                       test_templates.py:178-199); hand-derived known-answer
                       vectors for edge windows (tests/test_oracle.py)
inverse                closed-form identities (A A^-1 = I), numpy.linalg
compose_icgn           golden vector (subset.rs:1790-1810); identity, order
                       0, translation, undo and third-order tests
                       (subset.rs:1701-1779, 2261-2323)
convergence_norm       closed forms (subset.rs:1425-1456)
apply_warp / sdi       every-slot quad warp (subset.rs:2221-2259)
zn_slice / eta         hand-computed values (subset.rs:1460-1616)
residual_map           zero on identical images at p = 0; Σe = 0; Σe² equal
                       to the ZNSSD a one-iteration solve started at p
                       reports (delta_p_and_znssd, an independent pass);
                       translated pair: Σe² at the solution below Σe² at
                       p = 0
Subset::new statistics f_m, delta_f, SSSIG, sigma against NumPy on the
(f_m, delta_f, sssig,  oracle's own f / grad_f; a hand-computed four-pixel
sigma)                 case on a linear ramp (f_m 332.5, delta_f sqrt 13,
                       sigma sqrt(13/4), SSSIG 26)
subset_solve (ICGN /   synthetic translation recovered to 1e-4 px;
FAGN)                  ZNSSD = 0 on identical images; quality_ok needs
                       convergence; max_iterations=1 not converged;
                       FAGN ~ ICGN (test_subset.py:590-604 pattern)
=====================  ==================================================

The last-bit rounding of rustfft and nalgebra's LU is not reproducible
without the crates; those functions are pinned to round-off (~1e-12 rel).
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class OrcResult(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.c_double * 12),
        ("c_zncc", ctypes.c_double),
        ("c_znssd", ctypes.c_double),
        ("norm", ctypes.c_double),
        ("eta_u", ctypes.c_double),
        ("eta_v", ctypes.c_double),
        ("f_m", ctypes.c_double),
        ("delta_f", ctypes.c_double),
        ("sssig", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("iterations", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("quality_ok", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("n_px", ctypes.c_int32),
        ("order", ctypes.c_int32),
    ]


RESULT_DTYPE = np.dtype([
    ("p", np.float64, (12,)),
    ("c_zncc", np.float64), ("c_znssd", np.float64), ("norm", np.float64),
    ("eta_u", np.float64), ("eta_v", np.float64),
    ("f_m", np.float64), ("delta_f", np.float64), ("sssig", np.float64), ("sigma", np.float64),
    ("iterations", np.int32), ("converged", np.int32), ("quality_ok", np.int32),
    ("status", np.int32), ("n_px", np.int32), ("order", np.int32),
])


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_result_size.restype = ctypes.c_int
        assert L.orc_result_size() == ctypes.sizeof(OrcResult) == RESULT_DTYPE.itemsize
        L.orc_convergence_norm.restype = ctypes.c_double
        L.orc_convergence_norm.argtypes = [_dp, ctypes.c_int, ctypes.c_double]
        L.orc_eta_component.restype = ctypes.c_double
        L.orc_unrolled_sum.restype = ctypes.c_double
        L.orc_unrolled_sum.argtypes = [_dp, ctypes.c_int64]
        L.orc_wall_seconds.restype = ctypes.c_double
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


SHAPES = {"circle": 0, "square": 1, "semicircle": 2}


def fft(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    re = np.ascontiguousarray(x.real)
    im = np.ascontiguousarray(x.imag)
    lib().orc_fft(_d(re), _d(im), ctypes.c_int(len(x)), ctypes.c_int(sign))
    return re + 1j * im


def unrolled_sum(x):
    x = _c64(x)
    return lib().orc_unrolled_sum(_d(x), len(x))


def load_convert(pixels):
    """Image::from_file's conversion of decoded pixels (image.rs:450-527):
    (h, w, 3) RGB8 or (h, w) grey8 -> blurred grey, float64 0..255."""
    src = np.ascontiguousarray(pixels, dtype=np.uint8)
    h, w = src.shape[:2]
    ch = 3 if src.ndim == 3 else 1
    out = np.zeros((h, w))
    lib().orc_load_convert(src.ctypes.data_as(_u8p), h, w, ch, _d(out))
    return out


def pad_replicate(img, border):
    img = _c64(img)
    h, w = img.shape
    out = np.empty((h + 2 * border, w + 2 * border))
    lib().orc_pad_replicate(_d(img), h, w, border, _d(out))
    return out


def get_c(img, border=20, threads=0):
    img = _c64(img)
    h, w = img.shape
    out = np.empty((h + 2 * border, w + 2 * border))
    lib().orc_get_c(_d(img), h, w, border, _d(out), ctypes.c_int(threads))
    return out


def eval_points(coef, shape, border, xy):
    coef = _c64(coef)
    xy = _c64(np.atleast_2d(xy))
    out = np.empty(len(xy))
    lib().orc_eval(_d(coef), shape[0], shape[1], border, _d(xy), ctypes.c_int64(len(xy)), _d(out))
    return out


def eval_grad_points(coef, shape, border, xy):
    coef = _c64(coef)
    xy = _c64(np.atleast_2d(xy))
    out = np.empty((len(xy), 3))
    lib().orc_eval_grad(_d(coef), shape[0], shape[1], border, _d(xy), ctypes.c_int64(len(xy)), _d(out))
    return out


def qcqt_block(coef, shape, border, i, j):
    coef = _c64(coef)
    out = np.empty(36)
    lib().orc_qcqt_block(_d(coef), shape[0], shape[1], border, i, j, _d(out))
    return out.reshape(6, 6)


def mask_local(shape, size):
    sid = SHAPES[shape]
    side = 2 * size + 1
    coords = np.empty((side * side, 2))
    sm = np.empty((side, side), dtype=np.int32)
    n = lib().orc_mask_local(sid, size, _d(coords), sm.ctypes.data_as(_i32p))
    if n < 0:
        raise ValueError("size must be > 0")
    return coords[:n].copy(), sm


def mask_update(shape, size, centre, mask):
    sid = SHAPES[shape]
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    side = 2 * size + 1
    coords = np.empty((side * side, 2))
    n = lib().orc_mask_update(sid, size, ctypes.c_double(centre[0]), ctypes.c_double(centre[1]),
                              mask.ctypes.data_as(_u8p), mask.shape[0], mask.shape[1], _d(coords))
    if n < 0:
        raise ValueError("window outside the mask")
    return coords[:n].copy()


def inverse(M):
    M = _c64(M)
    m = M.shape[0]
    out = np.empty_like(M)
    ok = lib().orc_inverse(_d(M), m, _d(out))
    return out if ok else None


def compose_icgn(p, dp):
    p = _c64(p)
    dp = _c64(dp)
    out = np.empty_like(p)
    st = lib().orc_compose_icgn(_d(p), _d(dp), len(p), _d(out))
    if st:
        raise ArithmeticError("W_delta singular")
    return out


def convergence_norm(dp, size):
    dp = _c64(dp)
    return lib().orc_convergence_norm(_d(dp), len(dp), size)


def apply_warp(coord, p, f_coords):
    coord = _c64(coord)
    p = _c64(p)
    fc = _c64(f_coords)
    out = np.empty_like(fc)
    lib().orc_apply_warp(_d(coord), _d(p), len(p), _d(fc), ctypes.c_int64(len(fc)), _d(out))
    return out


def steepest_descent(coord, coords, grad, order):
    coord = _c64(coord)
    coords = _c64(coords)
    grad = _c64(grad)
    m = {0: 2, 1: 6, 2: 12}[order]
    out = np.empty((len(coords), m))
    lib().orc_steepest_descent(_d(coord), _d(coords), _d(grad), ctypes.c_int64(len(coords)), order, _d(out))
    return out


def hessian(sdi):
    sdi = _c64(sdi)
    n, m = sdi.shape
    out = np.empty((m, m))
    lib().orc_hessian(_d(sdi), ctypes.c_int64(n), m, _d(out))
    return out


def zn_slice(x):
    x = _c64(x)
    out = np.empty_like(x)
    lib().orc_zn_slice(_d(x), ctypes.c_int64(len(x)), _d(out))
    return out


def eta_component(modes, grad_zn, e_centered):
    modes = _c64(modes)
    g = _c64(grad_zn)
    e = _c64(e_centered)
    L = lib()
    L.orc_eta_component.argtypes = [_dp, ctypes.c_int, _dp, _dp, ctypes.c_int64]
    return L.orc_eta_component(_d(modes), modes.shape[1], _d(g), _d(e), len(g))


def omitted_mode_diagnostic(coord, order, f_coords, grad_f, f, f_m, delta_f, g, g_m, delta_g):
    coord = _c64(coord)
    fc = _c64(f_coords)
    gf = _c64(grad_f)
    f = _c64(f)
    g = _c64(g)
    eu = ctypes.c_double()
    ev = ctypes.c_double()
    st = lib().orc_omitted_mode_diagnostic(
        _d(coord), order, _d(fc), _d(gf), _d(f), ctypes.c_double(f_m), ctypes.c_double(delta_f),
        _d(g), ctypes.c_double(g_m), ctypes.c_double(delta_g), ctypes.c_int64(len(f)),
        ctypes.byref(eu), ctypes.byref(ev))
    if st:
        return None
    return eu.value, ev.value


def subset_solve(fcoef, gcoef, shape, border, coord, offsets, order, p0=None, method="icgn",
                 tolerance=0.75, max_norm=1e-3, max_iterations=50, want_arrays=False):
    """Subset::new + solve_{icgn,fagn}_result.  Returns a dict."""
    fcoef = _c64(fcoef)
    gcoef = _c64(gcoef)
    coord = _c64(coord)
    off = _c64(offsets)
    n = len(off)
    m = {0: 2, 1: 6, 2: 12}[order]
    p0a = np.zeros(12)
    p0_len = 0
    if p0 is not None:
        v = np.asarray(p0, dtype=np.float64).ravel()
        p0_len = min(len(v), 12)
        p0a[:p0_len] = v[:p0_len]
    R = OrcResult()
    hist = np.zeros((max(max_iterations, 1), 4))
    f_out = np.zeros(n)
    g_out = np.zeros((n, 2))
    fc_out = np.zeros((n, 2))
    st = lib().orc_subset_solve(
        _d(fcoef), _d(gcoef), shape[0], shape[1], border, _d(coord), _d(off), ctypes.c_int64(n), order,
        _d(p0a), ctypes.c_int(p0_len), ctypes.c_int(0 if method == "icgn" else 1),
        ctypes.c_double(tolerance), ctypes.c_double(max_norm), ctypes.c_int(max_iterations),
        ctypes.byref(R), _d(hist), _d(f_out), _d(g_out), _d(fc_out))
    out = {
        "status": st,
        "p": np.array(R.p[:m]),
        "c_zncc": R.c_zncc, "c_znssd": R.c_znssd, "norm": R.norm,
        "iterations": R.iterations, "converged": bool(R.converged), "quality_ok": bool(R.quality_ok),
        "eta_u": R.eta_u, "eta_v": R.eta_v,
        "f_m": R.f_m, "delta_f": R.delta_f, "sssig": R.sssig, "sigma": R.sigma, "n_px": R.n_px,
        "history": [(int(h[0]), h[1], h[2], h[3]) for h in hist[:R.iterations]],
    }
    if want_arrays:
        out.update(f=f_out, grad_f=g_out, f_coords=fc_out)
    return out


def residual_map(fcoef, gcoef, shape, border, coord, offsets, order, p):
    """Subset::residual_map at warp p (subset.rs:1350-1392): the per-pixel
    zero-normalised residual e_i = zn(f_i) - zn(g_i)."""
    fcoef = _c64(fcoef)
    gcoef = _c64(gcoef)
    coord = _c64(coord)
    off = _c64(offsets)
    m = {0: 2, 1: 6, 2: 12}[order]
    pv = np.zeros(12)
    v = np.asarray(p, dtype=np.float64).ravel()[:m]
    pv[:len(v)] = v
    out = np.zeros(len(off))
    st = lib().orc_residual_map(_d(fcoef), _d(gcoef), shape[0], shape[1], border, _d(coord), _d(off),
                                ctypes.c_int64(len(off)), order, _d(pv), _d(out))
    if st:
        raise RuntimeError(f"residual_map: status {st}")
    return out


def batch_solve(fcoef, gcoef, shape, border, coords, tpl_index, tpl_xy, tpl_start, orders, p0,
                method="icgn", tolerance=0.75, max_norm=1e-5, max_iterations=50, threads=0):
    """One subset per task over a batch; returns a RESULT_DTYPE array."""
    fcoef = _c64(fcoef)
    gcoef = _c64(gcoef)
    coords = _c64(coords)
    tpl_index = np.ascontiguousarray(tpl_index, dtype=np.int32)
    tpl_xy = _c64(tpl_xy)
    tpl_start = np.ascontiguousarray(tpl_start, dtype=np.int64)
    orders = np.ascontiguousarray(orders, dtype=np.int32)
    p0 = _c64(p0).reshape(len(coords), 12)
    out = np.zeros(len(coords), dtype=RESULT_DTYPE)
    lib().orc_batch_solve(
        _d(fcoef), _d(gcoef), shape[0], shape[1], border, ctypes.c_int64(len(coords)), _d(coords),
        tpl_index.ctypes.data_as(_i32p), _d(tpl_xy), tpl_start.ctypes.data_as(_i64p),
        orders.ctypes.data_as(_i32p), _d(p0), ctypes.c_int(0 if method == "icgn" else 1),
        ctypes.c_double(tolerance), ctypes.c_double(max_norm), ctypes.c_int(max_iterations),
        out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(threads))
    return out


def wall():
    return lib().orc_wall_seconds()


# ---------------------------------------------------------------------------
# mesh orchestration (mesh_oracle.cpp; mesh.rs)
# ---------------------------------------------------------------------------
class OrcMeshParams(ctypes.Structure):
    _fields_ = [
        ("subset_order", ctypes.c_int32), ("method", ctypes.c_int32), ("max_iterations", ctypes.c_int32),
        ("preconditioning", ctypes.c_int32), ("override_active", ctypes.c_int32), ("batch_factor", ctypes.c_int32),
        ("workers", ctypes.c_int32), ("threads", ctypes.c_int32),
        ("tolerance", ctypes.c_double), ("max_norm", ctypes.c_double), ("root_rel_eps", ctypes.c_double),
        ("seed_tolerance", ctypes.c_double), ("seed_coord", ctypes.c_double * 2), ("seed_warp", ctypes.c_double * 12),
    ]


class OrcMeshOut(ctypes.Structure):
    _fields_ = [
        ("p", _dp), ("c_zncc", _dp), ("iterations", _i32p), ("norms", _dp), ("quality_ok", _i32p), ("warps", _dp),
        ("seed_node", ctypes.c_int64), ("n_solves", ctypes.c_int64), ("n_bad", ctypes.c_int64),
    ]


def _el(elements):
    el = np.ascontiguousarray(elements, dtype=np.int64)
    return el, el.shape[0], el.shape[1], (1 if el.shape[1] == 3 else 2)


def mesh_connectivity(elements, idx, full, mesh_order=None):
    el, ne, cols, order = _el(elements)
    order = order if mesh_order is None else mesh_order
    out = np.zeros(cols * ne + 1, dtype=np.int64)
    k = lib().orc_mesh_connectivity(el.ctypes.data_as(_i64p), ctypes.c_int64(ne), cols, order, ctypes.c_int64(idx),
                                    int(bool(full)), out.ctypes.data_as(_i64p))
    return list(out[:k])


def mesh_find_seed(nodes, coord):
    nodes = _c64(nodes)
    lib().orc_mesh_find_seed.restype = ctypes.c_int64
    return int(lib().orc_mesh_find_seed(_d(nodes), ctypes.c_int64(len(nodes)), ctypes.c_double(coord[0]),
                                        ctypes.c_double(coord[1])))


def mesh_corr_1d(vals):
    v = _c64(vals)
    out = np.zeros(len(v) + 1, dtype=np.int64)
    lib().orc_mesh_corr_1d.restype = ctypes.c_int64
    k = lib().orc_mesh_corr_1d(_d(v), ctypes.c_int64(len(v)), out.ctypes.data_as(_i64p))
    return list(out[:k])


def mesh_flow_calc(elements, n_nodes, idx, disp, exclude=(), override=None):
    el, ne, cols, order = _el(elements)
    disp = _c64(disp)
    ex = np.zeros(n_nodes, dtype=np.uint8)
    ex[list(exclude)] = 1
    ov = None if override is None else _c64(override)
    L = lib()
    L.orc_mesh_flow_calc.restype = ctypes.c_double
    return L.orc_mesh_flow_calc(el.ctypes.data_as(_i64p), ctypes.c_int64(ne), cols, order, ctypes.c_int64(n_nodes),
                                ctypes.c_int64(idx), _d(disp), ex.ctypes.data_as(_u8p),
                                None if ov is None else _d(ov))


def mesh_r_stats(elements, disp):
    el, ne, cols, order = _el(elements)
    disp = _c64(disp)
    n = len(disp)
    r = np.zeros(n)
    flag = np.zeros(n, dtype=np.uint8)
    lib().orc_mesh_r_stats(el.ctypes.data_as(_i64p), ctypes.c_int64(ne), cols, order, ctypes.c_int64(n), _d(disp),
                           _d(r), flag.ctypes.data_as(_u8p))
    return r, flag.astype(bool)


def mesh_project_warp(p, dx, dy):
    p = _c64(p)
    out = np.zeros(len(p))
    lib().orc_mesh_project_warp(_d(p), len(p), ctypes.c_double(dx), ctypes.c_double(dy), _d(out))
    return out


def mesh_affine_field_fit(nodes, target, trusted, p):
    nodes, p = _c64(nodes), _c64(p)
    t = np.ascontiguousarray(trusted, dtype=np.int64)
    tg = _c64(target)
    out = np.zeros(p.shape[1])
    ok = lib().orc_mesh_affine_field_fit(_d(nodes), _d(tg), t.ctypes.data_as(_i64p), ctypes.c_int64(len(t)), _d(p),
                                         p.shape[1], _d(out))
    return out if ok else None


def mesh_precondition_warp(nodes, target, trusted, c_zncc, p):
    nodes, p, cz = _c64(nodes), _c64(p), _c64(c_zncc)
    t = np.ascontiguousarray(trusted, dtype=np.int64)
    tg = _c64(target)
    out = np.zeros(p.shape[1])
    lib().orc_mesh_precondition_warp(_d(nodes), _d(tg), t.ctypes.data_as(_i64p), ctypes.c_int64(len(t)), _d(cz),
                                     _d(p), p.shape[1], _d(out))
    return out


def mesh_greedy_colour(elements, n_nodes, candidates):
    el, ne, cols, order = _el(elements)
    c = np.ascontiguousarray(candidates, dtype=np.int64)
    out = np.zeros(len(c), dtype=np.int64)
    lib().orc_mesh_greedy_colour(el.ctypes.data_as(_i64p), ctypes.c_int64(ne), cols, order, ctypes.c_int64(n_nodes),
                                 c.ctypes.data_as(_i64p), ctypes.c_int64(len(c)), out.ctypes.data_as(_i64p))
    return list(out)


def mesh_element_area(nodes, elements):
    el, ne, cols, _ = _el(elements)
    nodes = _c64(nodes)
    out = np.zeros(ne)
    lib().orc_mesh_element_area(_d(nodes), el.ctypes.data_as(_i64p), ctypes.c_int64(ne), cols, _d(out))
    return out


def mesh_element_strains(nodes, elements, disp, mesh_order=None):
    el, ne, cols, order = _el(elements)
    order = order if mesh_order is None else mesh_order
    nodes, disp = _c64(nodes), _c64(disp)
    out = np.zeros((ne, 12))
    st = lib().orc_mesh_element_strains(_d(nodes), el.ctypes.data_as(_i64p), ctypes.c_int64(ne), order, _d(disp),
                                        _d(out))
    return st, out


def mesh_solve(fcoef, gcoef, shape, border, nodes, elements, tpl_index, tpl_xy, tpl_start, seed_coord,
               seed_warp=None, subset_order=1, tolerance=0.75, seed_tolerance=0.9, max_norm=1e-5, max_iterations=50,
               method="icgn", preconditioning="RG", override=False, batch_factor=4, root_rel_eps=0.02, workers=0,
               threads=0):
    """Mesh::solve_impl restated (uniform masking, deformable texture).
    Returns (status, dict of per-node and per-element arrays)."""
    fcoef, gcoef, nodes = _c64(fcoef), _c64(gcoef), _c64(nodes)
    el, ne, cols, order = _el(elements)
    tpl_index = np.ascontiguousarray(tpl_index, dtype=np.int32)
    tpl_xy = _c64(tpl_xy)
    tpl_start = np.ascontiguousarray(tpl_start, dtype=np.int64)
    n = len(nodes)
    prm = OrcMeshParams()
    prm.subset_order = subset_order
    prm.method = 0 if method == "icgn" else 1
    prm.max_iterations = max_iterations
    prm.preconditioning = 0 if preconditioning == "RG" else 1
    prm.override_active = int(bool(override))
    prm.batch_factor = batch_factor
    prm.workers = workers
    prm.threads = threads
    prm.tolerance = tolerance
    prm.max_norm = max_norm
    prm.root_rel_eps = root_rel_eps
    prm.seed_tolerance = seed_tolerance
    prm.seed_coord[0], prm.seed_coord[1] = float(seed_coord[0]), float(seed_coord[1])
    sw = np.zeros(12)
    if seed_warp is not None:
        v = np.asarray(seed_warp, dtype=np.float64).ravel()[:12]
        sw[:len(v)] = v
    for k in range(12):
        prm.seed_warp[k] = sw[k]
    res = dict(p=np.zeros((n, 12)), c_zncc=np.zeros(n), iterations=np.zeros(n, dtype=np.int32), norms=np.zeros(n),
               quality_ok=np.zeros(n, dtype=np.int32), warps=np.zeros((ne, 12)))
    out = OrcMeshOut(_d(res["p"]), _d(res["c_zncc"]), res["iterations"].ctypes.data_as(_i32p), _d(res["norms"]),
                     res["quality_ok"].ctypes.data_as(_i32p), _d(res["warps"]), 0, 0, 0)
    st = lib().orc_mesh_solve(_d(fcoef), _d(gcoef), shape[0], shape[1], border, _d(nodes), ctypes.c_int64(n),
                              el.ctypes.data_as(_i64p), ctypes.c_int64(ne), order,
                              tpl_index.ctypes.data_as(_i32p), _d(tpl_xy), tpl_start.ctypes.data_as(_i64p),
                              ctypes.byref(prm), ctypes.byref(out))
    m = 6 * subset_order
    res["p"] = res["p"][:, :m]
    res["quality_ok"] = res["quality_ok"].astype(bool)
    res.update(seed_node=int(out.seed_node), n_solves=int(out.n_solves), n_bad=int(out.n_bad))
    return st, res


def warp_increment(coord, element_nodes, element_disps, mesh_order):
    """particle.rs:254-320: warp of the displacement field at a point of
    one element (6 * mesh_order values)."""
    en, ed = _c64(element_nodes), _c64(element_disps)
    out = np.zeros(6 * mesh_order)
    lib().orc_warp_increment(_d(_c64(coord)), _d(en), _d(ed), mesh_order, _d(out))
    return out


def element_locator(coord, nodes, elements, centroids):
    """particle.rs:968-990: (element index, adrift)."""
    el, ne, cols, _ = _el(elements)
    nodes, cen = _c64(nodes), _c64(centroids)
    adrift = ctypes.c_int()
    L = lib()
    L.orc_element_locator.restype = ctypes.c_int64
    idx = L.orc_element_locator(_d(_c64(coord)), _d(nodes), el.ctypes.data_as(_i64p), ctypes.c_int64(ne), cols,
                                _d(cen), ctypes.byref(adrift))
    return int(idx), bool(adrift.value)

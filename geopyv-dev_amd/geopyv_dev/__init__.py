"""MI355X drop-in for geopyv-dev's subset-correlation path.

Same classes, arguments, defaults, getters and error behaviour as the
reference package's `geopyv_dev.Image`, `Mask` and `Subset`
(python/src/py_image.rs, py_mask.rs, py_subset.rs; geopyv_dev/__init__.py:
110-396), with every B-spline prefilter, interpolation and Gauss-Newton
solve executed by hand-written gfx950 HIP kernels behind the C ABI of
libgeodic.so.  `solve_subsets` exposes the batched engine.
"""

import os
import warnings

import numpy as np

from . import _native
from ._native import GeodicError

__all__ = ["Image", "Mask", "Subset", "Mesh", "Sequence", "structured_mesh", "quadratic_mesh", "solve_subsets",
           "GeodicError", "save", "load"]


def _engine():
    return _native.Engine.get(int(os.environ.get("GEODIC_DEVICE", "0")))


# ---------------------------------------------------------------------------
# Image (python/src/py_image.rs:34-122)
# ---------------------------------------------------------------------------
def decode_frame(path):
    """Image::from_file's decode step (image.rs:450-456): the file's pixels
    as 8-bit RGB, (h, w, 3) uint8.  The grey conversion and the blur run on
    the device (DIC_PIX_RGB8_LOAD)."""
    from PIL import Image as _PIL

    with _PIL.open(path) as im:
        return np.ascontiguousarray(np.asarray(im.convert("RGB"), dtype=np.uint8))


class Image:
    """Greyscale image with its bi-quintic B-spline coefficients resident on
    the GPU.  Exactly one of `filepath` or `image_gs` must be supplied."""

    def __init__(self, filepath=None, image_gs=None, border=20):
        if filepath is not None and image_gs is not None:
            raise ValueError("supply exactly one of `filepath` or `image_gs`, not both")
        if filepath is None and image_gs is None:
            raise ValueError("one of `filepath` or `image_gs` is required")
        if border < 0:
            raise OverflowError("can't convert negative int to unsigned")
        self._border = int(border)
        if filepath is not None:
            fp = os.fspath(filepath)
            if not os.path.exists(fp):
                raise RuntimeError(f"file not found: {fp}")
            try:
                rgb = decode_frame(fp)
            except Exception as exc:  # decoding failure -> Error::ImageLoad
                raise RuntimeError(f"image load error: {exc}") from None
            self._filepath = fp
            # grey conversion, blur and prefilter on the device (image.rs:69-77)
            self._dev = _native.DeviceImage(_engine(), pixels=rgb, border=self._border, load=True)
            self._gs = None
            self._shape = rgb.shape[:2]
        else:
            gs = np.array(image_gs, dtype=np.float64, copy=True)
            if gs.ndim != 2:
                raise TypeError("image_gs must be a 2-D float64 array")
            self._filepath = None
            self._gs = gs
            self._shape = gs.shape
            self._dev = _native.DeviceImage(_engine(), pixels=gs, border=self._border)

    @property
    def image_gs(self):
        if self._gs is None:  # the loader's output, read back once
            self._gs = self._dev.grey()
        return self._gs.copy()

    @property
    def border(self):
        return self._border

    @property
    def filepath(self):
        return self._filepath

    @property
    def coef(self):
        """B-spline coefficient array C, (H + 2b, W + 2b) (image.rs:48-51)."""
        return self._dev.coef()

    @property
    def qcqt(self):
        """Per-pixel Q·C_block·Qᵀ blocks in the (H*6, W*6) layout
        (py_image.rs:68-91), built on request from C."""
        from ._bspline import Q

        C = self.coef
        h, w = self._shape
        b = self._border
        win = np.lib.stride_tricks.sliding_window_view(C, (6, 6))[b - 2:b - 2 + h, b - 2:b - 2 + w]
        blocks = np.einsum("rk,ijkl,cl->ijrc", Q, win, Q)
        return blocks.transpose(0, 2, 1, 3).reshape(h * 6, w * 6)

    def __repr__(self):
        h, w = self._shape
        return f"Image(shape=({h}, {w}), border={self._border})"


# ---------------------------------------------------------------------------
# Mask (python/src/py_mask.rs:58-290)
# ---------------------------------------------------------------------------
def _point_in_polygon(px, py, poly):
    """Winding-number test, vectorised over points (geometry/meshing.rs)."""
    wn = np.zeros(px.shape, dtype=np.int64)
    n = len(poly)
    for i in range(n):
        x0, y0 = poly[i]
        x1, y1 = poly[(i + 1) % n]
        is_left = (x1 - x0) * (py - y0) - (px - x0) * (y1 - y0)
        up = (y0 <= py) & (y1 > py) & (is_left > 0)
        down = (y0 > py) & (y1 <= py) & (is_left < 0)
        wn += up.astype(np.int64) - down.astype(np.int64)
    return wn != 0


class Mask:
    """Local (subset template) or global (image ROI) mask."""

    def __init__(self, mask_type, shape=None, size=25, f_img=None, boundary=None, exclusions=None):
        self._binary = None
        if mask_type == "local":
            if shape is None:
                raise ValueError("shape is required for mask_type='local'")
            if shape not in _native.SHAPE:
                raise ValueError(f"unknown shape '{shape}': expected 'circle', 'square', or 'semicircle'")
            if size < 0:
                raise OverflowError("can't convert negative int to unsigned")
            if size == 0:
                raise RuntimeError("invalid input: " + ("length" if shape == "square" else "radius")
                                   + " must be > 0")
            self._type = "local"
            self._shape = shape
            self._size = int(size)
            self._coords, self._subset_mask = _native.mask_local(shape, self._size)
            self._n_px = len(self._coords)
            self._m_n_px = None
        elif mask_type == "global":
            if f_img is None:
                raise ValueError("f_img is required for mask_type='global'")
            if boundary is None:
                raise ValueError("boundary is required for mask_type='global'")
            h, w = f_img.image_gs.shape
            poly = np.asarray(boundary, dtype=np.float64).reshape(-1, 2)
            cols, rows = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
            binary = _point_in_polygon(cols, rows, poly)
            for exc in exclusions or []:
                e = np.asarray(exc, dtype=np.float64).reshape(-1, 2)
                binary &= ~_point_in_polygon(cols, rows, e)
            self._type = "global"
            self._binary = binary.astype(np.uint8)
        else:
            raise ValueError(f"unknown mask_type '{mask_type}': expected 'local' or 'global'")

    @property
    def mask_type(self):
        return self._type

    def _local(name):
        def get(self):
            return getattr(self, name) if self._type == "local" else None
        return property(get)

    shape = _local("_shape")
    size = _local("_size")
    n_px = _local("_n_px")
    m_n_px = _local("_m_n_px")

    @property
    def dimension(self):
        if self._type != "local":
            return None
        return "length" if self._shape == "square" else "radius"

    @property
    def coords(self):
        return self._coords.copy() if self._type == "local" else None

    @property
    def subset_mask(self):
        return self._subset_mask.copy() if self._type == "local" else None

    @property
    def binary(self):
        return self._binary.copy() if self._type == "global" else None

    def mask(self, centre, mask):
        """Apply a binary image mask (masks.rs:199-263): updates coords and m_n_px."""
        if self._type != "local":
            raise TypeError("mask() can only be called on local masks")
        coords = _native.mask_update(self._shape, self._size, centre, mask)
        self._coords = coords
        self._m_n_px = len(coords)

    def __repr__(self):
        if self._type == "local":
            return f"Mask(mask_type='local', shape='{self._shape}', size={self._size})"
        h, w = self._binary.shape
        return f"Mask(mask_type='global', shape=({h}, {w}))"


# ---------------------------------------------------------------------------
# Subset (python/src/py_subset.rs:28-545, geopyv_dev/__init__.py:299-396)
# ---------------------------------------------------------------------------
_P_LEN = {0: 2, 1: 6, 2: 12}
_SOLVER_OPTION_AXES = {"masking": {"uniform", "adaptive"}, "texture": {"deformable", "rigid"}}
_SOLVER_OPTION_DEFAULTS = {"masking": "uniform", "texture": "deformable"}


def _subset_options(solver_options):
    """Subset.solve's solver_options (reference __init__.py:143-160)."""
    if solver_options is None:
        solver_options = {}
    if not isinstance(solver_options, dict):
        raise TypeError(f"solver_options must be a dict, got {type(solver_options).__name__}")
    unknown = set(solver_options) - {"masking", "texture"}
    if unknown:
        raise ValueError(f"Subset.solve accepts only solver_options['masking'] and ['texture'], "
                         f"got {sorted(unknown)!r}")
    values = []
    for axis in ("masking", "texture"):
        value = solver_options.get(axis, _SOLVER_OPTION_DEFAULTS[axis])
        if value not in _SOLVER_OPTION_AXES[axis]:
            raise ValueError(f"solver_options[{axis!r}] must be one of "
                             f"{sorted(_SOLVER_OPTION_AXES[axis])!r}, got {value!r}")
        values.append(value)
    return tuple(values)


def _resolve_solver(explicit_solver, kwargs, deprecated_name):
    """reference __init__.py:235-262"""
    have_old = deprecated_name in kwargs
    old_val = kwargs.pop(deprecated_name, None)
    if have_old:
        warnings.warn(f"{deprecated_name}= is deprecated, use solver= instead.", DeprecationWarning, stacklevel=3)
    if explicit_solver is not None and have_old and explicit_solver != old_val:
        raise ValueError(f"both solver={explicit_solver!r} and {deprecated_name}={old_val!r} given "
                         "with different values")
    if explicit_solver is not None:
        return explicit_solver
    if have_old:
        return old_val
    return "icgn"


def _checked_solver(solver):
    """reference __init__.py:268-280: unknown -> UserWarning, fall back to icgn."""
    algo = str(solver).lower()
    if algo not in ("icgn", "fagn"):
        warnings.warn(f"Unknown solver '{algo}'; falling back to 'icgn'.", UserWarning, stacklevel=3)
        return "icgn"
    return algo


def _time_arg(time):
    if time is None:
        return None, "none"
    if isinstance(time, bool):
        raise TypeError("time must be a number or a sequence of numbers")
    if np.isscalar(time):
        return (0.0, float(time)), "given"
    try:
        t = [float(v) for v in time]
    except Exception:
        raise TypeError("time must be a time increment (number) or one time per image "
                        "(sequence of numbers)") from None
    if len(t) != 2:
        raise ValueError("a subset takes two times: (f_time, g_time)")
    return (t[0], t[1]), "given"


class Subset:
    """Reference subset for DIC; `solve` runs the FP64 drop-in ICGN/FAGN on
    the GPU and stores the result on the object."""

    def __init__(self, coord, local_mask, f_img, g_img, subset_order=1, global_mask=None, time=None,
                 soil_properties=None):
        if subset_order > 2:
            raise ValueError(f"subset_order must be 0, 1, or 2, got {subset_order}")
        if not isinstance(local_mask, Mask):
            raise TypeError("local_mask must be a Mask")
        if local_mask.mask_type != "local":
            raise TypeError("expected a local mask (mask_type='local'), got a global mask")
        if not isinstance(f_img, Image) or not isinstance(g_img, Image):
            raise TypeError("f_img and g_img must be Image")
        self._coord = [float(coord[0]), float(coord[1])]
        self._order = int(subset_order)
        self._local_mask = local_mask
        self._global_mask = global_mask
        self._f_img, self._g_img = f_img, g_img
        self._time, self._time_source = _time_arg(time)
        self._soil = soil_properties
        if global_mask is not None and global_mask.mask_type == "global":
            offsets = _native.mask_update(local_mask.shape, local_mask.size, self._coord, global_mask.binary)
        else:
            offsets = local_mask.coords
        self._offsets = offsets
        n = len(offsets)
        if n == 0:
            raise RuntimeError("invalid input: effective mask coords is empty")
        self._result = None
        self._history = None
        # Subset::new runs on the device with the first solve; the reference
        # quantities are computed here with a zero-iteration solve so that the
        # getters and the featureless check behave as at construction.
        self._batch = self._make_batch()
        self._run(None, "icgn", 1e-3, 0, 0.75, keep=False)

    def _make_batch(self):
        xy = np.ascontiguousarray(self._offsets, dtype=np.float64)
        if np.any(xy != np.rint(xy)) or np.any(np.abs(xy) > 32767):
            raise ValueError("mask offsets must be integers within ±32767")
        return _native.Batch(_engine(), np.array([self._coord]), self._order, np.array([0], dtype=np.int32),
                             np.array([0, len(xy)], dtype=np.int64), np.rint(xy).astype(np.int16))

    def _run(self, p_0, method, max_norm, max_iterations, tolerance, keep=True):
        m = _P_LEN[self._order]
        p0 = np.zeros(12)
        if p_0 is not None:
            v = np.asarray(p_0, dtype=np.float64).ravel()
            k = min(len(v), m)
            p0[:k] = v[:k]
        self._batch.set_p0(p0[None, :])
        if max_iterations < 0:
            raise OverflowError("can't convert negative int to unsigned")
        self._batch.solve(self._f_img._dev, self._g_img._dev, method=method, precision=_native.PREC_PLAIN_F64,
                          max_iterations=max_iterations, tolerance=tolerance, max_norm=max_norm, history=True)
        r = self._batch.results()[0]
        st = int(r["status"])
        self._f_m = float(r["f_m"])
        self._delta_f = float(r["delta_f"])
        self._sssig = float(r["sssig"])
        self._sigma = float(r["sigma"])
        if st == _native.ERR_INVALID_INPUT:
            raise RuntimeError("invalid input: reference subset is featureless (delta_f = 0)")
        if st == _native.ERR_SINGULAR_HESSIAN:
            raise RuntimeError("ICGN Hessian singular" if method == "icgn" else "FAGN Hessian singular")
        if st == _native.ERR_SINGULAR_WARP:
            raise RuntimeError(f"W_delta singular (order-{self._order})")
        if st != 0:
            raise GeodicError(st, f"solve failed with status {st}")
        if keep:
            it = int(r["iterations"])
            h = self._batch.history()[0][:it]
            self._result = dict(
                p=[float(v) for v in r["p"][:m]], c_zncc=float(r["c_zncc"]), c_znssd=float(r["c_znssd"]),
                converged=bool(r["converged"]), iterations=it, quality_ok=bool(r["quality_ok"]),
                history=[(int(a), float(b), float(c), float(d)) for a, b, c, d in h],
                max_norm=float(max_norm), tolerance=float(tolerance), eta_u=float(r["eta_u"]),
                eta_v=float(r["eta_v"]), method=method, norm=float(r["norm"]))
        else:
            f, g = self._batch.subset_arrays(0)
            self._f, self._grad_f = f, g

    def solve(self, p_0=None, solver=None, max_norm=1e-3, max_iterations=50, tolerance=0.75, solver_options=None,
              **extra_kwargs):
        """Run the DIC solver; mutates the object and returns None
        (reference __init__.py:338-385)."""
        algo = _checked_solver(_resolve_solver(solver, extra_kwargs, "algorithm"))
        masking, texture = _subset_options(solver_options)
        if masking != "uniform" or texture != "deformable":
            raise NotImplementedError("only solver_options masking='uniform', texture='deformable' run on the GPU "
                                      "engine (adaptive masking and rigid texture are outside the hot path)")
        self._run(p_0, algo, max_norm, max_iterations, tolerance)

    def solve_icgn(self, p_0=None, max_norm=1e-3, max_iterations=50, tolerance=0.75, masking="uniform",
                   texture="deformable"):
        self.solve(p_0=p_0, solver="icgn", max_norm=max_norm, max_iterations=max_iterations, tolerance=tolerance,
                   solver_options={"masking": masking, "texture": texture})

    def solve_fagn(self, p_0=None, max_norm=1e-3, max_iterations=50, tolerance=0.75, masking="uniform",
                   texture="deformable"):
        self.solve(p_0=p_0, solver="fagn", max_norm=max_norm, max_iterations=max_iterations, tolerance=tolerance,
                   solver_options={"masking": masking, "texture": texture})

    def residual_map(self):
        """Per-pixel zero-normalised residual at the stored solution
        (subset.rs:1350-1392)."""
        if self._result is None:
            raise RuntimeError("invalid input: residual_map: subset has not been solved")
        p = np.zeros((1, 12))
        p[0, :len(self._result["p"])] = self._result["p"]
        return self._batch.residual_map(self._f_img._dev, self._g_img._dev, [0], p)[0]

    # always-available getters
    coord = property(lambda self: list(self._coord))
    n_px = property(lambda self: len(self._offsets))
    subset_order = property(lambda self: self._order)
    sssig = property(lambda self: self._sssig)
    time = property(lambda self: self._time)
    time_source = property(lambda self: self._time_source)
    solved = property(lambda self: self._result is not None)
    template_shape = property(lambda self: self._local_mask.shape)
    template_size = property(lambda self: self._local_mask.size)
    template_n_px = property(lambda self: len(self._offsets))
    f_img = property(lambda self: self._f_img)
    g_img = property(lambda self: self._g_img)
    local_mask = property(lambda self: self._local_mask)
    f_img_path = property(lambda self: self._f_img.filepath)
    g_img_path = property(lambda self: self._g_img.filepath)
    soil_properties = property(lambda self: self._soil)

    # image-dependent getters
    @property
    def f_coords(self):
        return np.asarray(self._offsets, dtype=np.float64) + np.array(self._coord)

    f = property(lambda self: self._f.copy())
    f_m = property(lambda self: self._f_m)
    delta_f = property(lambda self: self._delta_f)
    grad_f = property(lambda self: self._grad_f.copy())
    sigma_intensity = property(lambda self: self._sigma)

    def _r(key):
        return property(lambda self: None if self._result is None else self._result[key])

    p = _r("p")
    c_zncc = _r("c_zncc")
    c_znssd = _r("c_znssd")
    converged = _r("converged")
    iterations = _r("iterations")
    history = _r("history")
    max_norm = _r("max_norm")
    tolerance = _r("tolerance")
    eta_u = _r("eta_u")
    eta_v = _r("eta_v")
    quality_ok = _r("quality_ok")

    @property
    def texture(self):
        return None if self._result is None else "deformable"

    @property
    def seams(self):
        return None

    def __repr__(self):
        tmpl = f"{self._local_mask.shape}({self._local_mask.size})"
        head = (f"Subset(coord=[{self._coord[0]:.2f}, {self._coord[1]:.2f}], n_px={self.n_px}, template={tmpl}, "
                f"sssig={self._sssig:.1f}")
        if self._result is None:
            return head + ", solved=False)"
        p = ", ".join(f"{v:.6f}" for v in self._result["p"])
        return head + f",\n       solved=True, c_zncc={self._result['c_zncc']:.5f}, p=[{p}])"


# ---------------------------------------------------------------------------
# batched engine
# ---------------------------------------------------------------------------
def templates_from_masks(offset_lists):
    """Deduplicate pixel lists into (tpl_index, tpl_start, tpl_xy)."""
    keys = {}
    starts = [0]
    xy = []
    index = np.empty(len(offset_lists), dtype=np.int32)
    for s, off in enumerate(offset_lists):
        a = np.rint(np.asarray(off, dtype=np.float64)).astype(np.int16)
        k = a.tobytes()
        t = keys.get(k)
        if t is None:
            t = len(keys)
            keys[k] = t
            xy.append(a)
            starts.append(starts[-1] + len(a))
        index[s] = t
    return index, np.array(starts, dtype=np.int64), np.concatenate(xy).reshape(-1, 2)


def solve_subsets(f_img, g_img, coords, local_mask, p0=None, subset_order=1, precision="f32", method="icgn",
                  max_norm=1e-5, max_iterations=50, tolerance=0.75, global_mask=None):
    """Solve many subsets in one launch (Subset::new + solve per subset).

    local_mask is one Mask for every subset or a sequence of Masks, one per
    subset (uneven radii); subset_order is one warp order or one per subset
    (orders may mix).  precision "f64" is the plain FP64 drop-in, "f32" the
    batched engine.  Returns a structured array of results (fields as
    dic_subset_result)."""
    coords = np.asarray(coords, dtype=np.float64).reshape(-1, 2)
    if isinstance(local_mask, (list, tuple)):
        if len(local_mask) != len(coords):
            raise ValueError("one local mask per subset is required")
        if global_mask is not None:
            lists = [_native.mask_update(m.shape, m.size, c, global_mask.binary) for m, c in zip(local_mask, coords)]
        else:
            lists = [m.coords for m in local_mask]
        index, start, xy = templates_from_masks(lists)
    elif global_mask is not None:
        lists = [_native.mask_update(local_mask.shape, local_mask.size, c, global_mask.binary) for c in coords]
        index, start, xy = templates_from_masks(lists)
    else:
        xy = np.rint(local_mask.coords).astype(np.int16)
        index = np.zeros(len(coords), dtype=np.int32)
        start = np.array([0, len(xy)], dtype=np.int64)
    b = _native.Batch(_engine(), coords, subset_order, index, start, xy)
    if p0 is not None:
        b.set_p0(np.asarray(p0, dtype=np.float64).reshape(len(coords), 12))
    prec = _native.PREC_PLAIN_F64 if precision == "f64" else _native.PREC_COMP_F32
    b.solve(f_img._dev, g_img._dev, method=method, precision=prec, max_iterations=max_iterations,
            tolerance=tolerance, max_norm=max_norm)
    return b.results()


def solve_sequence(frames, coords, local_mask, subset_order=1, reference_every=20, p0=None, border=20,
                   max_norm=1e-5, max_iterations=50, tolerance=0.75, use_graph=True):
    """Incremental correlation over a sequence of grey frames, the frame loop
    of sequence.rs:624-968 in batch form: frames[0] is the first reference;
    each later frame is prefiltered on the device and its subsets solved,
    warm-started from the previous frame's warps; every `reference_every`
    frames the current frame becomes the reference (its coefficients are
    recomputed on the device, sequence.rs:914) and the warps restart from
    zero.  The per-frame steps run as one captured HIP graph (use_graph) or
    as eager calls; both give identical results.

    Returns one result array per frame after the first (fields as
    dic_subset_result, warps relative to that frame's reference)."""
    frames = list(frames)
    if len(frames) < 2:
        raise ValueError("a sequence needs at least two frames")
    coords = np.asarray(coords, dtype=np.float64).reshape(-1, 2)
    n = len(coords)
    if isinstance(local_mask, (list, tuple)):
        index, start, xy = templates_from_masks([m.coords for m in local_mask])
    else:
        xy = np.rint(local_mask.coords).astype(np.int16)
        index = np.zeros(n, dtype=np.int32)
        start = np.array([0, len(xy)], dtype=np.int64)
    eng = _engine()
    first = np.asarray(frames[0])
    f = _native.DeviceImage(eng, pixels=first, border=border)
    g = _native.DeviceImage(eng, pixels=np.asarray(frames[1]), border=border)
    b = _native.Batch(eng, coords, subset_order, index, start, xy)
    b.reset_warps(np.zeros((n, 12)) if p0 is None else np.asarray(p0, dtype=np.float64).reshape(n, 12))
    out = []
    graph = None
    for k in range(1, len(frames)):
        frame = np.asarray(frames[k])
        if use_graph:
            g.stage(frame)
            if graph is None:
                graph = _native.FrameGraph(b, f, g, warm_start=True, max_iterations=max_iterations,
                                           tolerance=tolerance, max_norm=max_norm)  # runs this frame
            else:
                graph.launch()
        else:
            if k > 1:
                b.set_p0_device(b.device_p())
            g.update(frame)
            b.solve(f, g, method="icgn", precision=_native.PREC_COMP_F32, max_iterations=max_iterations,
                    tolerance=tolerance, max_norm=max_norm)
        eng.sync()
        out.append(b.results())
        if reference_every and k % reference_every == 0 and k + 1 < len(frames):
            f.update(frame)
            b.reset_warps(np.zeros((n, 12)))
    return out


from .mesh import Mesh, quadratic_mesh, structured_mesh  # noqa: E402  (uses Image, Mask and the solver helpers above)
from .sequence import Sequence  # noqa: E402


def save(path, obj):
    """Serialise a solved mesh (a Mesh or its MeshSolution) to a `.pyv` file
    (geopyv_dev/__init__.py:81-86 of the reference; other solution types
    are outside the correlation engine)."""
    from .pyv import MeshSolution, save_mesh
    if isinstance(obj, Mesh):
        obj = obj.solution()
    if not isinstance(obj, MeshSolution):
        raise TypeError("save: only mesh solutions are written by this package")
    save_mesh(path, obj)


def load(path):
    """Load a `.pyv` mesh solution (MeshSolution record)."""
    from .pyv import load_mesh
    return load_mesh(path)

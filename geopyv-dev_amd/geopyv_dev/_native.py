"""ctypes binding of libgeodic.so (the C ABI declared in include/geodic.h).

The shared library is the product: every solve in this package goes through
it.  There is no CPU fallback; importing a GPU entry point without the built
library raises.
"""

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GEODIC_LIB selects another build of the same library (same C ABI)
LIB_PATH = os.environ.get("GEODIC_LIB") or os.path.join(os.path.dirname(_HERE), "lib", "libgeodic.so")

# status codes (include/geodic.h)
OK = 0
ERR_IMAGE_LOAD = 1
ERR_FILE_NOT_FOUND = 2
ERR_INVALID_INPUT = 3
ERR_NOT_SOLVED = 8
ERR_CANCELLED = 9
ERR_SINGULAR_HESSIAN = 10
ERR_SINGULAR_WARP = 11
ERR_DEVICE = 20
ERR_OUT_OF_MEMORY = 21
ERR_BUSY = 22
PENDING = 23

PIX_U8, PIX_F64, PIX_RGB8_LOAD, PIX_GREY8_LOAD = 0, 1, 2, 3
MEM_HOST, MEM_DEVICE = 0, 1
SHAPE = {"circle": 0, "square": 1, "semicircle": 2}
METHOD = {"icgn": 0, "fagn": 1}
PREC_PLAIN_F64, PREC_COMP_F32 = 0, 1
PREFILTER_FIR, PREFILTER_FFT = 0, 1
PRECOND = {"RG": 0, "layer-RG": 1}
ERR_MESH_GENERATION = 4

EXPORTED = [
    "dic_engine_create", "dic_engine_destroy", "dic_engine_sync", "dic_engine_timing", "dic_last_error",
    "dic_engine_set_prefilter",
    "dic_engine_bytes_in_use", "dic_image_create", "dic_image_update", "dic_image_coef", "dic_image_info",
    "dic_image_destroy", "dic_mask_local", "dic_mask_update", "dic_batch_create", "dic_batch_set_p0",
    "dic_batch_solve", "dic_batch_results", "dic_batch_history", "dic_batch_subset_arrays",
    "dic_batch_set_schedule", "dic_batch_autotune", "dic_batch_device_p", "dic_batch_destroy",
    "dic_image_stage", "dic_batch_reset_warps", "dic_frame_graph_create", "dic_frame_graph_launch",
    "dic_frame_graph_destroy", "dic_batch_solve_list", "dic_mesh_solve", "dic_mesh_element_strains",
    "dic_image_grey", "dic_batch_solve_noalloc", "dic_batch_get_schedule", "dic_batch_residual_map",
    "dic_server_create", "dic_submit", "dic_poll", "dic_server_stats", "dic_server_destroy",
]


class SolveParams(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32),
        ("want_history", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("max_norm", ctypes.c_double),
    ]


class ServerParams(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("n_slots", ctypes.c_int32), ("max_queued", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("hbm_budget", ctypes.c_int64)]


class JobStruct(ctypes.Structure):
    """dic_job."""
    _fields_ = [
        ("ref_key", ctypes.c_uint64), ("ref_pixels", ctypes.c_void_p), ("tar_pixels", ctypes.c_void_p),
        ("pix_type", ctypes.c_int32), ("border", ctypes.c_int32), ("h", ctypes.c_int64), ("w", ctypes.c_int64),
        ("n_subsets", ctypes.c_int64), ("coords", ctypes.c_void_p), ("orders", ctypes.c_void_p),
        ("tpl_index", ctypes.c_void_p), ("n_templates", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("tpl_start", ctypes.c_void_p), ("tpl_xy", ctypes.c_void_p), ("p0", ctypes.c_void_p),
        ("params", SolveParams),
    ]


class MeshParams(ctypes.Structure):
    _fields_ = [
        ("solve", SolveParams),
        ("subset_order", ctypes.c_int32),
        ("preconditioning", ctypes.c_int32),
        ("override_active", ctypes.c_int32),
        ("batch_factor", ctypes.c_int32),
        ("workers", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("root_rel_eps", ctypes.c_double),
        ("seed_coord", ctypes.c_double * 2),
        ("seed_warp", ctypes.c_double * 12),
        ("seed_tolerance", ctypes.c_double),
    ]


class MeshOutput(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.c_void_p), ("c_zncc", ctypes.c_void_p), ("iterations", ctypes.c_void_p),
        ("norms", ctypes.c_void_p), ("quality_ok", ctypes.c_void_p), ("warps", ctypes.c_void_p),
        ("seed_node", ctypes.c_int64), ("n_node_solves", ctypes.c_int64), ("n_lists", ctypes.c_int64),
        ("n_unsolvable", ctypes.c_int64), ("list_seconds", ctypes.c_double),
    ]


_MESH_START = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int64)
_NODE_DONE = ctypes.CFUNCTYPE(None, ctypes.c_void_p)
_CANCELLED = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p)


class Progress(ctypes.Structure):
    """dic_progress: the SolveProgress observer (src/progress.rs:16-38)."""
    _fields_ = [("user", ctypes.c_void_p), ("mesh_start", _MESH_START), ("node_done", _NODE_DONE),
                ("cancelled", _CANCELLED)]


def progress_struct(observer):
    """A dic_progress calling observer.mesh_start(n), observer.node_done()
    and observer.cancelled() (each optional, as the trait's no-op
    defaults).  Returns (struct, keepalive): the callbacks must outlive the
    call."""
    if observer is None:
        return None, None
    cbs = []

    def wrap(name, ctype, fn):
        if not callable(getattr(observer, name, None)):
            return ctype()
        c = ctype(fn)
        cbs.append(c)
        return c

    pr = Progress(None,
                  wrap("mesh_start", _MESH_START, lambda _u, n: observer.mesh_start(int(n))),
                  wrap("node_done", _NODE_DONE, lambda _u: observer.node_done()),
                  wrap("cancelled", _CANCELLED, lambda _u: 1 if observer.cancelled() else 0))
    return pr, cbs


RESULT_DTYPE = np.dtype([
    ("p", np.float64, (12,)),
    ("c_zncc", np.float64), ("c_znssd", np.float64), ("norm", np.float64),
    ("eta_u", np.float64), ("eta_v", np.float64),
    ("f_m", np.float64), ("delta_f", np.float64), ("sssig", np.float64), ("sigma", np.float64),
    ("iterations", np.int32), ("converged", np.int32), ("quality_ok", np.int32),
    ("status", np.int32), ("n_px", np.int32), ("order", np.int32),
])

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)

_lib = None
_lock = threading.Lock()


class GeodicError(RuntimeError):
    """A failed engine call; `status` is the dic_status code."""

    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


class SolveCancelled(RuntimeError):
    """A solve stopped by its progress observer (Error::Cancelled,
    error.rs:38-40; raised as RuntimeError("solve cancelled") like every
    Error the PyO3 layer surfaces)."""

    def __init__(self, msg="solve cancelled"):
        super().__init__(msg)


def load():
    """Load libgeodic.so and declare its prototypes."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libgeodic.so is not built ({LIB_PATH}); run __graft_entry__.build() "
                              "or `make -C geopyv-dev_amd`")
        L = ctypes.CDLL(LIB_PATH)
        L.dic_last_error.restype = ctypes.c_char_p
        L.dic_engine_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
        L.dic_engine_destroy.argtypes = [_vp]
        L.dic_engine_destroy.restype = None
        L.dic_engine_sync.argtypes = [_vp]
        L.dic_engine_timing.argtypes = [_vp, _dp, _dp]
        L.dic_engine_set_prefilter.argtypes = [_vp, ctypes.c_int]
        L.dic_engine_bytes_in_use.argtypes = [_vp]
        L.dic_engine_bytes_in_use.restype = ctypes.c_int64
        L.dic_image_create.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                       ctypes.c_int, ctypes.POINTER(_vp)]
        L.dic_image_update.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int]
        L.dic_image_coef.argtypes = [_vp, _vp]
        L.dic_image_grey.argtypes = [_vp, _vp]
        L.dic_image_info.argtypes = [_vp, _i64p, _i64p, ctypes.POINTER(ctypes.c_int)]
        L.dic_image_destroy.argtypes = [_vp]
        L.dic_image_destroy.restype = None
        L.dic_mask_local.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp, _i64p]
        L.dic_mask_update.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp,
                                      ctypes.c_int64, ctypes.c_int64, _vp, _i64p]
        L.dic_batch_create.argtypes = [_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int32, _vp, _vp,
                                       ctypes.POINTER(_vp)]
        L.dic_batch_set_p0.argtypes = [_vp, _vp, ctypes.c_int]
        L.dic_batch_solve.argtypes = [_vp, _vp, _vp, ctypes.POINTER(SolveParams)]
        L.dic_batch_solve_noalloc.argtypes = [_vp, _vp, _vp, ctypes.POINTER(SolveParams)]
        L.dic_batch_results.argtypes = [_vp, _vp]
        L.dic_batch_history.argtypes = [_vp, _vp]
        L.dic_batch_subset_arrays.argtypes = [_vp, ctypes.c_int64, _vp, _vp]
        L.dic_batch_set_schedule.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.dic_batch_autotune.argtypes = [_vp, _vp, _vp, ctypes.POINTER(SolveParams)]
        L.dic_batch_get_schedule.argtypes = [_vp, _vp, _vp]
        L.dic_batch_residual_map.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int64, _vp, _vp]
        L.dic_image_stage.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int]
        L.dic_batch_reset_warps.argtypes = [_vp, _vp, ctypes.c_int]
        L.dic_frame_graph_create.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.POINTER(SolveParams),
                                             ctypes.POINTER(_vp)]
        L.dic_frame_graph_launch.argtypes = [_vp]
        L.dic_frame_graph_destroy.argtypes = [_vp]
        L.dic_frame_graph_destroy.restype = None
        L.dic_batch_device_p.argtypes = [_vp, ctypes.POINTER(_vp)]
        L.dic_batch_destroy.argtypes = [_vp]
        L.dic_batch_destroy.restype = None
        L.dic_batch_solve_list.argtypes = [_vp, _vp, _vp, ctypes.POINTER(SolveParams), _vp, ctypes.c_int64, _vp,
                                           _vp]
        L.dic_mesh_solve.argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_int32, _vp, _vp,
                                     ctypes.c_int32, _vp, _vp, ctypes.POINTER(MeshParams),
                                     ctypes.POINTER(Progress), ctypes.POINTER(MeshOutput)]
        L.dic_mesh_element_strains.argtypes = [ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_int32, _vp, _vp, _vp]
        L.dic_server_create.argtypes = [ctypes.POINTER(ServerParams), ctypes.POINTER(_vp)]
        L.dic_submit.argtypes = [_vp, ctypes.POINTER(JobStruct), ctypes.POINTER(ctypes.c_uint64)]
        L.dic_poll.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int32, _vp, _dp]
        L.dic_server_stats.argtypes = [_vp, _i64p, _i64p, _i64p, _i64p]
        L.dic_server_destroy.argtypes = [_vp]
        L.dic_server_destroy.restype = None
        _lib = L
        return L


def load_lib():
    return load()


def check(status):
    if status != OK:
        msg = load().dic_last_error().decode(errors="replace")
        raise GeodicError(status, msg or f"geodic status {status}")


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


# ---------------------------------------------------------------------------
# masks (host, no device needed)
# ---------------------------------------------------------------------------
def mask_local(shape, size):
    L = load()
    side = 2 * size + 1
    coords = np.empty((side * side, 2), dtype=np.float64)
    sm = np.empty((side, side), dtype=np.int32)
    n = ctypes.c_int64()
    st = L.dic_mask_local(SHAPE[shape], int(size), ptr(coords), ptr(sm), ctypes.byref(n))
    if st:
        raise RuntimeError(L.dic_last_error().decode())
    return coords[:n.value].copy(), sm


def mask_update(shape, size, centre, mask):
    L = load()
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    side = 2 * size + 1
    coords = np.empty((side * side, 2), dtype=np.float64)
    n = ctypes.c_int64()
    st = L.dic_mask_update(SHAPE[shape], int(size), float(centre[0]), float(centre[1]), ptr(mask),
                           mask.shape[0], mask.shape[1], ptr(coords), ctypes.byref(n))
    if st:
        raise RuntimeError(L.dic_last_error().decode())
    return coords[:n.value].copy()


# ---------------------------------------------------------------------------
# device objects
# ---------------------------------------------------------------------------
class Engine:
    """One engine (HIP stream + caches) per device, shared by the package."""

    _engines = {}

    def __init__(self, device=0):
        L = load()
        h = _vp()
        check(L.dic_engine_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device

    @classmethod
    def get(cls, device=0):
        with _lock:
            e = cls._engines.get(device)
        if e is None:
            e = cls(device)
            with _lock:
                cls._engines[device] = e
        return e

    def sync(self):
        check(load().dic_engine_sync(self.handle))

    def set_prefilter(self, method):
        """PREFILTER_FIR (default) or PREFILTER_FFT: how get_c is evaluated."""
        check(load().dic_engine_set_prefilter(self.handle, int(method)))

    def timing(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(load().dic_engine_timing(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def bytes_in_use(self):
        return load().dic_engine_bytes_in_use(self.handle)


def _pix_type(pixels, load):
    """Pixel type of a host array: decoded file pixels when `load` (the
    loader's grey conversion and blur run on the device), else grey."""
    if load:
        if pixels.dtype != np.uint8 or pixels.ndim not in (2, 3) or (pixels.ndim == 3 and pixels.shape[2] != 3):
            raise ValueError("decoded file pixels must be (h, w, 3) or (h, w) uint8")
        return PIX_RGB8_LOAD if pixels.ndim == 3 else PIX_GREY8_LOAD
    return PIX_U8 if pixels.dtype == np.uint8 else PIX_F64


class DeviceImage:
    """Device-resident B-spline coefficients of one image."""

    def __init__(self, engine, pixels=None, device_ptr=None, shape=None, pix_type=None, border=20, load=False):
        L = load_lib()
        self.engine = engine
        h = _vp()
        if device_ptr is not None:
            hh, ww = shape
            check(L.dic_image_create(engine.handle, ctypes.c_void_p(device_ptr), pix_type, MEM_DEVICE, hh, ww,
                                     int(border), ctypes.byref(h)))
        else:
            pixels = np.ascontiguousarray(pixels)
            pt = _pix_type(pixels, load)
            if pt == PIX_F64:
                pixels = np.ascontiguousarray(pixels, dtype=np.float64)
            hh, ww = pixels.shape[:2]
            check(L.dic_image_create(engine.handle, ptr(pixels), pt, MEM_HOST, hh, ww, int(border),
                                     ctypes.byref(h)))
        self.handle = h
        self.shape = (int(hh), int(ww))
        self.border = int(border)

    def update(self, pixels=None, device_ptr=None, pix_type=None, load=False):
        L = load_lib()
        if device_ptr is not None:
            check(L.dic_image_update(self.handle, ctypes.c_void_p(device_ptr), pix_type, MEM_DEVICE))
        else:
            pixels = np.ascontiguousarray(pixels)
            pt = _pix_type(pixels, load)
            if pt == PIX_F64:
                pixels = np.ascontiguousarray(pixels, dtype=np.float64)
            check(L.dic_image_update(self.handle, ptr(pixels), pt, MEM_HOST))

    def stage(self, pixels=None, device_ptr=None, pix_type=None, load=False):
        """Copy a frame's pixels into the image's device pixel buffer (the
        source its frame graph prefilters)."""
        L = load_lib()
        if device_ptr is not None:
            check(L.dic_image_stage(self.handle, ctypes.c_void_p(device_ptr), pix_type, MEM_DEVICE))
        else:
            pixels = np.ascontiguousarray(pixels)
            pt = _pix_type(pixels, load)
            if pt == PIX_F64:
                pixels = np.ascontiguousarray(pixels, dtype=np.float64)
            check(L.dic_image_stage(self.handle, ptr(pixels), pt, MEM_HOST))

    def grey(self):
        """The grey frame an image loaded from file pixels was prefiltered
        from (h, w) float64."""
        out = np.empty(self.shape, dtype=np.float64)
        check(load_lib().dic_image_grey(self.handle, ptr(out)))
        return out

    def coef(self):
        h, w = self.shape
        b = self.border
        out = np.empty((h + 2 * b, w + 2 * b), dtype=np.float64)
        check(load().dic_image_coef(self.handle, ptr(out)))
        return out

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().dic_image_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class NativeServer:
    """The library's serving loop (dic_server_create / dic_submit / dic_poll):
    submit() queues a job and returns its ticket; poll() returns (records,
    latency_ms) once it has finished, None while it runs (wait_ms 0)."""

    def __init__(self, n_slots=4, hbm_budget=64 << 30, device=0, max_queued=0):
        prm = ServerParams(int(device), int(n_slots), int(max_queued), 0, int(hbm_budget))
        h = _vp()
        check(load().dic_server_create(ctypes.byref(prm), ctypes.byref(h)))
        self.handle = h
        self._n = {}

    def submit(self, ref_key, ref, tar, coords, orders, tpl_index, tpl_start, tpl_xy, p0=None, border=20,
               method="icgn", precision=PREC_COMP_F32, max_iterations=50, tolerance=0.75, max_norm=1e-5):
        ref = np.asarray(ref)
        if ref.dtype == np.uint8:
            pix = PIX_U8
            ref = np.ascontiguousarray(ref)
            tar = np.ascontiguousarray(tar, dtype=np.uint8)
        else:
            pix = PIX_F64
            ref = np.ascontiguousarray(ref, dtype=np.float64)
            tar = np.ascontiguousarray(tar, dtype=np.float64)
        if tar.shape != ref.shape or ref.ndim != 2:
            raise ValueError("reference and target must be 2-D images of one shape")
        coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 2)
        n = len(coords)
        orders = (np.full(n, int(orders), np.int32) if np.ndim(orders) == 0
                  else np.ascontiguousarray(orders, dtype=np.int32).reshape(n))
        tpl_index = np.ascontiguousarray(tpl_index, dtype=np.int32).reshape(n)
        tpl_start = np.ascontiguousarray(tpl_start, dtype=np.int64)
        tpl_xy = np.ascontiguousarray(tpl_xy, dtype=np.int16).reshape(-1, 2)
        p0a = None if p0 is None else np.ascontiguousarray(p0, dtype=np.float64).reshape(n, 12)
        prm = SolveParams(METHOD[method], int(precision), int(max_iterations), 0, float(tolerance), float(max_norm))
        j = JobStruct(int(ref_key), ref.ctypes.data, tar.ctypes.data, pix, int(border), ref.shape[0], ref.shape[1],
                      n, coords.ctypes.data, orders.ctypes.data, tpl_index.ctypes.data, len(tpl_start) - 1, 0,
                      tpl_start.ctypes.data, tpl_xy.ctypes.data, None if p0a is None else p0a.ctypes.data, prm)
        t = ctypes.c_uint64()
        check(load().dic_submit(self.handle, ctypes.byref(j), ctypes.byref(t)))  # copies the arrays
        self._n[t.value] = n
        return t.value

    def poll(self, ticket, wait_ms=-1):
        n = self._n[ticket]
        out = np.zeros(n, dtype=RESULT_DTYPE)
        lat = ctypes.c_double()
        st = load().dic_poll(self.handle, ctypes.c_uint64(ticket), int(wait_ms), ptr(out), ctypes.byref(lat))
        if st == PENDING:
            return None
        del self._n[ticket]
        check(st)
        return out, lat.value

    def stats(self):
        v = [ctypes.c_int64() for _ in range(4)]
        check(load().dic_server_stats(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(("queued", "running", "finished", "cache_bytes"), (x.value for x in v)))

    def close(self):
        if getattr(self, "handle", None):
            load().dic_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A set of subsets solved together (dic_batch_*).  `order` is one warp
    order for every subset, or one per subset (orders may mix)."""

    def __init__(self, engine, coords, order, tpl_index, tpl_start, tpl_xy):
        L = load()
        self.engine = engine
        coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 2)
        n = len(coords)
        if np.ndim(order) == 0:
            orders = np.full(n, int(order), dtype=np.int32)
        else:
            orders = np.ascontiguousarray(order, dtype=np.int32).reshape(n)
        tpl_index = np.ascontiguousarray(tpl_index, dtype=np.int32)
        tpl_start = np.ascontiguousarray(tpl_start, dtype=np.int64)
        tpl_xy = np.ascontiguousarray(tpl_xy, dtype=np.int16).reshape(-1, 2)
        h = _vp()
        check(L.dic_batch_create(engine.handle, n, ptr(coords), ptr(orders), ptr(tpl_index),
                                 len(tpl_start) - 1, ptr(tpl_start), ptr(tpl_xy), ctypes.byref(h)))
        self.handle = h
        self.n = n
        self.orders = orders
        self.order = int(orders.max())  # the highest warp order present
        self.tpl_index = tpl_index
        self.tpl_start = tpl_start
        self._keep = (coords, orders, tpl_xy)
        self.max_iterations = 0

    def set_p0(self, p0):
        p0 = np.ascontiguousarray(p0, dtype=np.float64).reshape(self.n, 12)
        check(load().dic_batch_set_p0(self.handle, ptr(p0), MEM_HOST))

    def set_p0_device(self, device_ptr):
        check(load().dic_batch_set_p0(self.handle, ctypes.c_void_p(device_ptr), MEM_DEVICE))

    def set_schedule(self, win_slack, waves_per_subset):
        check(load().dic_batch_set_schedule(self.handle, int(win_slack), int(waves_per_subset)))

    def reset_warps(self, p):
        """Set both the initial and the final warps (n x 12)."""
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(self.n, 12))
        check(load().dic_batch_reset_warps(self.handle, ptr(p), MEM_HOST))

    def autotune(self, f_img, g_img, max_iterations=50, tolerance=0.75, max_norm=1e-5):
        """Time the batched engine's candidate schedules per size class and
        keep the fastest (results are unspecified until the next solve)."""
        prm = SolveParams(METHOD["icgn"], PREC_COMP_F32, int(max_iterations), 0, float(tolerance), float(max_norm))
        check(load().dic_batch_autotune(self.handle, f_img.handle, g_img.handle, ctypes.byref(prm)))

    def schedule(self):
        """{fewest wavefronts of a size class: (wavefronts per subset, window
        slack)} the batched engine runs each class with."""
        w = np.zeros(4, np.int32)
        s = np.zeros(4, np.int32)
        check(load().dic_batch_get_schedule(self.handle, ptr(w), ptr(s)))
        return {1 << k: (int(w[k]), int(s[k])) for k in range(4)}

    def solve(self, f_img, g_img, method="icgn", precision=PREC_PLAIN_F64, max_iterations=50, tolerance=0.75,
              max_norm=1e-5, history=False):
        prm = SolveParams(METHOD[method], int(precision), int(max_iterations), 1 if history else 0,
                          float(tolerance), float(max_norm))
        self.max_iterations = int(max_iterations)
        check(load().dic_batch_solve(self.handle, f_img.handle, g_img.handle, ctypes.byref(prm)))

    def solve_noalloc(self, f_img, g_img, method="icgn", max_iterations=50, tolerance=0.75, max_norm=1e-5):
        """The plain FP64 solve through dic_batch_solve_noalloc: its per-pixel
        scratch is the library's fixed device array, so it allocates nothing
        (batches beyond that array's size raise GeodicError)."""
        prm = SolveParams(METHOD[method], PREC_PLAIN_F64, int(max_iterations), 0, float(tolerance), float(max_norm))
        self.max_iterations = int(max_iterations)
        check(load().dic_batch_solve_noalloc(self.handle, f_img.handle, g_img.handle, ctypes.byref(prm)))

    def solve_list(self, f_img, g_img, idx, p0, method="icgn", precision=PREC_PLAIN_F64, max_iterations=50,
                   tolerance=0.75, max_norm=1e-5):
        """Solve only the subsets idx (each once) from p0 (len(idx) x 12);
        returns their results in list order."""
        idx = np.ascontiguousarray(idx, dtype=np.int32).ravel()
        p0 = np.ascontiguousarray(np.asarray(p0, dtype=np.float64).reshape(len(idx), 12))
        out = np.zeros(len(idx), dtype=RESULT_DTYPE)
        prm = SolveParams(METHOD[method], int(precision), int(max_iterations), 0, float(tolerance), float(max_norm))
        check(load().dic_batch_solve_list(self.handle, f_img.handle, g_img.handle, ctypes.byref(prm), ptr(idx),
                                          len(idx), ptr(p0), ptr(out)))
        return out

    def results(self):
        out = np.zeros(self.n, dtype=RESULT_DTYPE)
        check(load().dic_batch_results(self.handle, ptr(out)))
        return out

    def history(self):
        cap = max(self.max_iterations, 1)
        out = np.zeros((self.n, cap, 4), dtype=np.float64)
        check(load().dic_batch_history(self.handle, ptr(out)))
        return out

    def subset_arrays(self, s):
        t = self.tpl_index[s]
        n = int(self.tpl_start[t + 1] - self.tpl_start[t])
        f = np.empty(n)
        g = np.empty((n, 2))
        check(load().dic_batch_subset_arrays(self.handle, int(s), ptr(f), ptr(g)))
        return f, g

    def residual_map(self, f_img, g_img, idx, p=None):
        """Subset::residual_map of the listed subsets at warps p (len(idx) x
        12), or at the stored final warps; a list of per-pixel arrays."""
        idx = np.ascontiguousarray(idx, dtype=np.int32).ravel()
        sizes = [int(self.tpl_start[self.tpl_index[s] + 1] - self.tpl_start[self.tpl_index[s]]) for s in idx]
        out = np.zeros(int(sum(sizes)))
        pp = None
        if p is not None:
            pp = np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(len(idx), 12))
        check(load().dic_batch_residual_map(self.handle, f_img.handle, g_img.handle, ptr(idx), len(idx),
                                            None if pp is None else ptr(pp), ptr(out)))
        return np.split(out, np.cumsum(sizes)[:-1])

    def device_p(self):
        p = _vp()
        check(load().dic_batch_device_p(self.handle, ctypes.byref(p)))
        return p.value

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().dic_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class FrameGraph:
    """The per-frame steps of an incremental sequence as one HIP graph
    (dic_frame_graph_*): prefilter of the staged target frame, warm start
    p0 := p (if asked), batched ICGN solve.  Creating it processes the frame
    staged at that moment."""

    def __init__(self, batch, f_img, g_img, warm_start=True, max_iterations=50, tolerance=0.75, max_norm=1e-5):
        self.batch, self.f_img, self.g_img = batch, f_img, g_img  # the graph holds their device buffers
        prm = SolveParams(METHOD["icgn"], PREC_COMP_F32, int(max_iterations), 0, float(tolerance), float(max_norm))
        h = _vp()
        check(load().dic_frame_graph_create(batch.handle, f_img.handle, g_img.handle, 1 if warm_start else 0,
                                            ctypes.byref(prm), ctypes.byref(h)))
        self.handle = h

    def launch(self):
        check(load().dic_frame_graph_launch(self.handle))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().dic_frame_graph_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# mesh solve
# ---------------------------------------------------------------------------
def mesh_element_strains(nodes, elements, displacements, mesh_order):
    """element_strains (host): warps (n_elements x 12)."""
    nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 2)
    el = np.ascontiguousarray(elements, dtype=np.int64)
    d = np.ascontiguousarray(displacements, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((len(el), 12))
    check(load().dic_mesh_element_strains(len(nodes), ptr(nodes), len(el), int(mesh_order), ptr(el), ptr(d),
                                          ptr(out)))
    return out


def mesh_solve(engine, f_img, g_img, nodes, elements, mesh_order, tpl_index, tpl_start, tpl_xy, seed_coord,
               seed_warp=None, subset_order=1, method="icgn", precision=PREC_PLAIN_F64, max_iterations=50,
               tolerance=0.75, max_norm=1e-5, seed_tolerance=0.9, preconditioning="RG", override=False,
               batch_factor=4, root_rel_eps=0.02, workers=0, progress=None, out=None):
    """dic_mesh_solve.  Returns (status, message, outputs dict); outputs
    are filled when the propagation completed (status 0, or the quality
    gate, a degenerate or a folded element).  progress: an observer with
    any of mesh_start(n), node_done(), cancelled() (SolveProgress); out: a
    dict of caller-owned output arrays to fill instead of new ones."""
    L = load()
    nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 2)
    el = np.ascontiguousarray(elements, dtype=np.int64)
    tpl_index = np.ascontiguousarray(tpl_index, dtype=np.int32)
    tpl_start = np.ascontiguousarray(tpl_start, dtype=np.int64)
    tpl_xy = np.ascontiguousarray(tpl_xy, dtype=np.int16).reshape(-1, 2)
    n, ne = len(nodes), len(el)
    prm = MeshParams()
    prm.solve = SolveParams(METHOD[method], int(precision), int(max_iterations), 0, float(tolerance),
                            float(max_norm))
    prm.subset_order = int(subset_order)
    prm.preconditioning = PRECOND[preconditioning]
    prm.override_active = 1 if override else 0
    prm.batch_factor = int(batch_factor)
    prm.workers = int(workers)
    prm.root_rel_eps = float(root_rel_eps)
    prm.seed_coord[0], prm.seed_coord[1] = float(seed_coord[0]), float(seed_coord[1])
    sw = np.zeros(12)
    if seed_warp is not None:
        v = np.asarray(seed_warp, dtype=np.float64).ravel()[:12]
        sw[:len(v)] = v
    for k in range(12):
        prm.seed_warp[k] = sw[k]
    prm.seed_tolerance = float(seed_tolerance)
    if out is None:
        res = dict(p=np.zeros((n, 12)), c_zncc=np.zeros(n), iterations=np.zeros(n, dtype=np.int32),
                   norms=np.zeros(n), quality_ok=np.zeros(n, dtype=np.int32), warps=np.zeros((ne, 12)))
    else:
        res = out
        for k, shape, dt in (("p", (n, 12), np.float64), ("c_zncc", (n,), np.float64),
                             ("iterations", (n,), np.int32), ("norms", (n,), np.float64),
                             ("quality_ok", (n,), np.int32), ("warps", (ne, 12), np.float64)):
            a = res[k]
            if a.shape != shape or a.dtype != dt or not a.flags.c_contiguous:
                raise ValueError(f"out[{k!r}] must be a C-contiguous {dt.__name__} array of shape {shape}")
    mo = MeshOutput(res["p"].ctypes.data, res["c_zncc"].ctypes.data, res["iterations"].ctypes.data,
                    res["norms"].ctypes.data, res["quality_ok"].ctypes.data, res["warps"].ctypes.data, -1, 0, 0, 0,
                    0.0)
    pr, keep = progress_struct(progress)
    st = L.dic_mesh_solve(engine.handle, f_img.handle, g_img.handle, n, ptr(nodes), ne, int(mesh_order), ptr(el),
                          ptr(tpl_index), len(tpl_start) - 1, ptr(tpl_start), ptr(tpl_xy), ctypes.byref(prm),
                          ctypes.byref(pr) if pr is not None else None, ctypes.byref(mo))
    del keep
    out = mo
    msg = L.dic_last_error().decode(errors="replace") if st else ""
    res = dict(res)
    res["p"] = res["p"][:, :6 * int(subset_order)]
    res["quality_ok"] = res["quality_ok"].astype(bool)
    res.update(seed_node=int(out.seed_node), n_node_solves=int(out.n_node_solves), n_lists=int(out.n_lists),
               n_unsolvable=int(out.n_unsolvable), list_seconds=float(out.list_seconds))
    return st, msg, res

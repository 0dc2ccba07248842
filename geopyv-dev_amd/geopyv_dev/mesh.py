"""Mesh: reliability-guided correlation over the nodes of a triangular mesh.

The drop-in for the reference's ``Mesh`` solve (geopyv_dev/__init__.py:398-508,
python/src/py_mesh.rs:200-500, src/mesh.rs:609-857) on a mesh given by its
nodes and elements.  Every node solve runs on the engine (``dic_mesh_solve``):
serial RG one node per launch, layer-RG one launch per colour of each
frontier round.  Generating the mesh from a boundary and a target node
count (geometry/meshing.rs) is not part of the correlation path; pass
``nodes=`` and ``elements=`` (``structured_mesh`` builds a regular one).
"""

import numpy as np

from . import _native

_P_LEN = {1: 6, 2: 12}


def structured_mesh(x0, y0, x1, y1, pitch):
    """Nodes on a regular grid over [x0, x1] x [y0, y1] (row-major, x
    fastest) and two triangles per cell, both counter-clockwise in image
    axes: (c, r), (c+1, r), (c, r+1) and (c+1, r), (c+1, r+1), (c, r+1)."""
    xs = np.arange(x0, x1 + 1e-9, pitch, dtype=np.float64)
    ys = np.arange(y0, y1 + 1e-9, pitch, dtype=np.float64)
    nx, ny = len(xs), len(ys)
    gx, gy = np.meshgrid(xs, ys)
    nodes = np.stack([gx.ravel(), gy.ravel()], axis=1)
    c, r = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1))
    c, r = c.ravel(), r.ravel()
    a = r * nx + c
    el = np.empty((2 * len(a), 3), dtype=np.int64)
    el[0::2] = np.stack([a, a + 1, a + nx], axis=1)
    el[1::2] = np.stack([a + 1, a + nx + 1, a + nx], axis=1)
    return nodes, el


def quadratic_mesh(nodes, elements):
    """6-node elements from 3-node ones: one new node at the midpoint of
    each distinct edge, appended after the corners in first-seen order;
    element rows become [c0, c1, c2, m01, m12, m20]."""
    nodes = np.asarray(nodes, dtype=np.float64).reshape(-1, 2)
    el = np.asarray(elements, dtype=np.int64)
    mids = {}
    extra = []
    out = np.empty((len(el), 6), dtype=np.int64)
    out[:, :3] = el
    for e, (a, b, c) in enumerate(el):
        for k, (i, j) in enumerate(((a, b), (b, c), (c, a))):
            key = (min(i, j), max(i, j))
            m = mids.get(key)
            if m is None:
                m = mids[key] = len(nodes) + len(extra)
                extra.append(0.5 * (nodes[i] + nodes[j]))
            out[e, 3 + k] = m
    return np.concatenate([nodes, np.array(extra).reshape(-1, 2)]), out


def _polygon(region):
    if region is None:
        return None
    pts = getattr(region, "nodes", region)
    return np.asarray(pts, dtype=np.float64).reshape(-1, 2)


def roi_mask(shape, boundary=None, boundary_hard=False, exclusions=(), exclusions_hard=None):
    """Binary region-of-interest mask (mask_image, geometry/meshing.rs:58-103):
    pixel centres inside the boundary when it is hard (every pixel when it
    is soft), less the hard exclusions."""
    from . import _point_in_polygon
    h, w = shape
    cols, rows = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
    if boundary is not None and boundary_hard:
        m = _point_in_polygon(cols, rows, _polygon(boundary))
    else:
        m = np.ones((h, w), dtype=bool)
    exclusions = list(exclusions or [])
    hard = [True] * len(exclusions) if exclusions_hard is None else list(exclusions_hard)
    for exc, hd in zip(exclusions, hard):
        if hd:
            m &= ~_point_in_polygon(cols, rows, _polygon(exc))
    return m.astype(np.uint8)


def node_templates(local_mask, nodes, mask):
    """Each node's pixel list as subset_at builds it (mesh.rs:575-593): the
    local mask clipped by ``mask`` with mask_update (masks.rs:199-263).
    Nodes whose window lies inside the image over unmasked pixels keep the
    whole local template; the others are clipped one by one.  Returns
    (tpl_index, tpl_start, tpl_xy) with identical lists shared."""
    size = int(local_mask.size)
    base = np.rint(local_mask.coords).astype(np.int16)
    mh, mw = mask.shape
    zero = np.zeros((mh + 1, mw + 1), dtype=np.int64)
    zero[1:, 1:] = np.cumsum(np.cumsum(mask == 0, axis=0), axis=1)
    cx = np.trunc(nodes[:, 0]).astype(np.int64)  # `as isize` truncates (masks.rs:204-205)
    cy = np.trunc(nodes[:, 1]).astype(np.int64)
    lx, ux = cx - size, cx + size + 1
    ly, uy = cy - size, cy + size + 1
    inside = (lx >= 0) & (ly >= 0) & (ux < mw) & (uy < mh)
    full = np.zeros(len(nodes), dtype=bool)
    i = np.nonzero(inside)[0]
    holes = zero[uy[i], ux[i]] - zero[ly[i], ux[i]] - zero[uy[i], lx[i]] + zero[ly[i], lx[i]]
    full[i[holes == 0]] = True
    lists = [base]
    keys = {base.tobytes(): 0}
    tpl_index = np.zeros(len(nodes), dtype=np.int32)
    for k in np.nonzero(~full)[0]:
        xy = np.rint(_native.mask_update(local_mask.shape, size, nodes[k], mask)).astype(np.int16)
        if len(xy) == 0:
            raise RuntimeError("invalid input: effective mask coords is empty")
        key = xy.tobytes()
        t = keys.get(key)
        if t is None:
            t = keys[key] = len(lists)
            lists.append(xy)
        tpl_index[k] = t
    start = np.concatenate([[0], np.cumsum([len(a) for a in lists])]).astype(np.int64)
    return tpl_index, start, np.concatenate(lists)


def element_area(nodes, elements):
    """Signed element areas (mesh.rs:1468-1481)."""
    e = np.asarray(elements)[:, :3]
    x, y = nodes[:, 0], nodes[:, 1]
    return 0.5 * ((x[e[:, 1]] - x[e[:, 0]]) * (y[e[:, 2]] - y[e[:, 0]])
                  - (x[e[:, 2]] - x[e[:, 0]]) * (y[e[:, 1]] - y[e[:, 0]]))


def element_centroids(nodes, elements):
    """Mean of the corner nodes (compute_centroids, mesh.rs:1668-1680)."""
    e = np.asarray(elements)[:, :3]
    return (nodes[e[:, 0]] + nodes[e[:, 1]] + nodes[e[:, 2]]) / 3.0


class Mesh:
    """DIC mesh over given nodes and elements; ``solve()`` mutates in place
    and returns None.  mesh_order 1: elements (M, 3); 2: (M, 6) with the
    midside nodes m01, m12, m20 after the corners."""

    def __init__(self, boundary=None, target_nodes=None, f_img=None, g_img=None, size=(1, 1000),
                 exclusions=None, exclusions_hard=None, mesh_order=1, time=None, noise=None,
                 soil_properties=None, *, nodes=None, elements=None, boundary_hard=False):
        from . import Image
        if nodes is None or elements is None:
            raise NotImplementedError("automatic mesh generation (geometry/meshing.rs) is outside the "
                                      "correlation engine: pass nodes= and elements= (see structured_mesh)")
        if not isinstance(f_img, Image) or not isinstance(g_img, Image):
            raise TypeError("f_img and g_img must be Image")
        if mesh_order not in (1, 2):
            raise ValueError("mesh_order must be 1 or 2")
        self._nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 2)
        self._elements = np.ascontiguousarray(elements, dtype=np.int64)
        if self._elements.ndim != 2 or self._elements.shape[1] != (3 if mesh_order == 1 else 6):
            raise ValueError(f"elements must be (M, {3 if mesh_order == 1 else 6}) for mesh_order {mesh_order}")
        self._mesh_order = int(mesh_order)
        self._f_img, self._g_img = f_img, g_img
        self._mask = roi_mask(f_img._shape, boundary, boundary_hard, exclusions, exclusions_hard)
        self._time = time
        self._soil = soil_properties
        self._solution = None
        self._templates = None  # (local mask key, tpl_index, tpl_start, tpl_xy)

    def _use_images(self, f_img, g_img):
        """Another image pair of the same shape on the same geometry (a
        sequence's next pair): the per-node pixel lists are kept."""
        if f_img._shape != self._f_img._shape or g_img._shape != self._f_img._shape:
            raise ValueError("images of another shape need another mesh")
        self._f_img, self._g_img = f_img, g_img
        self._solution = None

    # ---- geometry ----
    @property
    def nodes(self):
        return self._nodes.copy()

    @property
    def elements(self):
        return self._elements.copy()

    @property
    def mesh_order(self):
        return self._mesh_order

    @property
    def f_img(self):
        return self._f_img

    @property
    def g_img(self):
        return self._g_img

    @property
    def solved(self):
        return self._solution is not None

    def _get(self, key):
        if self._solution is None:
            raise AttributeError("mesh not solved: call solve() first")
        return self._solution[key]

    areas = property(lambda self: self._get("areas"))
    centroids = property(lambda self: self._get("centroids"))
    warps = property(lambda self: self._get("warps"))
    displacements = property(lambda self: self._get("displacements"))
    c_zncc = property(lambda self: self._get("c_zncc"))
    p = property(lambda self: self._get("p"))
    seed_node = property(lambda self: self._get("seed_node"))
    subset_order = property(lambda self: self._get("subset_order"))
    iterations = property(lambda self: self._get("iterations"))
    norms = property(lambda self: self._get("norms"))
    quality_ok = property(lambda self: self._get("quality_ok"))
    template_shape = property(lambda self: self._get("template_shape"))
    template_sizes = property(lambda self: self._get("template_sizes"))
    stats = property(lambda self: self._get("stats"))

    def solve(self, local_mask, seed_coord, seed_warp=None, solver=None, solver_options=None, max_norm=1e-5,
              max_iterations=50, subset_order=1, tolerance=0.75, seed_tolerance=0.9, override_active=False,
              precision="f64", progress=None, **kwargs):
        """Seed, RG or layer-RG propagation, corrections, the quality gate,
        element strains and the compatibility check (mesh.rs:628-857).
        ``precision`` "f64" runs the FP64 drop-in solver at every node, "f32"
        the batched engine (FP32 interpolation, FP64 accumulation).
        ``progress`` is Mesh::solve's observer (mesh.rs:610-616): an object
        with any of ``mesh_start(n)``, ``node_done()`` and ``cancelled()``;
        a cancelled solve raises RuntimeError("cancelled") and stores
        nothing."""
        from . import Mask, _checked_solver, _resolve_solver
        if not isinstance(local_mask, Mask) or local_mask.mask_type != "local":
            raise TypeError("local_mask must be a Mask")
        method = _checked_solver(_resolve_solver(solver, kwargs, "method"))
        if kwargs:
            raise TypeError(f"unexpected keyword arguments: {sorted(kwargs)}")
        opts = dict(solver_options or {})
        pre = opts.pop("preconditioning", "RG")
        if pre not in ("RG", "layer-RG"):
            raise ValueError(f"preconditioning must be 'RG' or 'layer-RG', got {pre!r}")
        lrg = dict(opts.pop("layer_rg", None) or {})
        for axis, default in (("topology", "uniform"), ("masking", "uniform"), ("texture", "deformable")):
            v = opts.pop(axis, default)
            if v != default:
                raise NotImplementedError(f"solver_options['{axis}']={v!r} is outside the correlation engine "
                                          f"(only '{default}' runs on the GPU)")
        if opts:
            raise ValueError(f"unknown solver_options keys: {sorted(opts)}")
        batch_factor = int(lrg.get("batch_factor", 4))
        if batch_factor < 1:
            raise ValueError("layer_rg_batch_factor must be >= 1")
        if subset_order not in _P_LEN:
            raise ValueError("a mesh solves subset_order 1 or 2")
        if precision not in ("f64", "f32"):
            raise ValueError("precision must be 'f64' or 'f32'")
        key = (local_mask.shape, local_mask.size, local_mask.coords.tobytes())
        if self._templates is None or self._templates[0] != key:
            self._templates = (key,) + node_templates(local_mask, self._nodes, self._mask)
        _, tpl_index, tpl_start, tpl_xy = self._templates
        m = _P_LEN[subset_order]
        warp = np.zeros(m) if seed_warp is None else np.asarray(seed_warp, dtype=np.float64).ravel()
        eng = _native.Engine.get(0)
        st, msg, r = _native.mesh_solve(
            eng, self._f_img._dev, self._g_img._dev, self._nodes, self._elements, self._mesh_order, tpl_index,
            tpl_start, tpl_xy, seed_coord, warp, subset_order=subset_order, method=method,
            precision=_native.PREC_PLAIN_F64 if precision == "f64" else _native.PREC_COMP_F32,
            max_iterations=max_iterations, tolerance=tolerance, max_norm=max_norm, seed_tolerance=seed_tolerance,
            preconditioning=pre, override=override_active, batch_factor=batch_factor,
            root_rel_eps=float(lrg.get("root_rel_eps", 0.02)), workers=int(lrg.get("max_workers") or 0),
            progress=progress)
        if st == _native.ERR_CANCELLED:
            raise _native.SolveCancelled(msg or "solve cancelled")
        if st:
            raise RuntimeError(msg or f"mesh solve failed with status {st}")
        p = r["p"]
        self._solution = dict(
            areas=element_area(self._nodes, self._elements), centroids=element_centroids(self._nodes, self._elements),
            warps=r["warps"], displacements=p[:, :2].copy(), c_zncc=r["c_zncc"], p=p, seed_node=r["seed_node"],
            subset_order=int(subset_order), iterations=r["iterations"].astype(np.uint32), norms=r["norms"],
            quality_ok=r["quality_ok"], template_shape=local_mask.shape,
            template_sizes=np.full(len(self._nodes), local_mask.size, dtype=np.uint32),
            stats=dict(node_solves=r["n_node_solves"], lists=r["n_lists"]),
            solve_config=dict(max_norm=float(max_norm), max_iterations=int(max_iterations),
                              subset_order=int(subset_order), tolerance=float(tolerance), method=method,
                              override_active=bool(override_active), preconditioning=pre,
                              batch_factor=batch_factor, root_rel_eps=float(lrg.get("root_rel_eps", 0.02)),
                              max_workers=lrg.get("max_workers")),
            seed=dict(coord=[float(seed_coord[0]), float(seed_coord[1])], warp=np.resize(warp, m).copy(),
                      tolerance=float(seed_tolerance)))

    def solution(self):
        """The solved mesh as the MeshSolution record of a `.pyv` file
        (mesh.rs:54-113)."""
        from .pyv import MeshSolution, boundary_nodes
        if self._solution is None:
            raise RuntimeError("Mesh has not been solved; call solve() first")
        s = self._solution
        t = self._time
        return MeshSolution(
            nodes=self._nodes, elements=self._elements, boundary=boundary_nodes(self._elements, self._mesh_order),
            exclusions=[], centroids=s["centroids"], areas=s["areas"], warps=s["warps"],
            displacements=s["displacements"], c_zncc=s["c_zncc"], p=s["p"], seed_node=s["seed_node"],
            mesh_order=self._mesh_order, subset_order=s["subset_order"], iterations=s["iterations"],
            norms=s["norms"], f_img_path=self._f_img.filepath or "", g_img_path=self._g_img.filepath or "",
            solve_config=s["solve_config"], seed=s["seed"], template_shape=s["template_shape"],
            template_sizes=s["template_sizes"], time=None if t is None else (float(t[0]), float(t[1])))

    def save(self, path):
        """Save the solved mesh to a `.pyv` file (GPYV, version 0x17,
        bincode v2, zstd level 3: io.rs:2796-2816)."""
        from .pyv import save_mesh
        save_mesh(path, self.solution())

    def __repr__(self):
        state = "solved" if self.solved else "unsolved"
        return (f"Mesh({len(self._nodes)} nodes, {len(self._elements)} elements, order {self._mesh_order}, "
                f"{state})")

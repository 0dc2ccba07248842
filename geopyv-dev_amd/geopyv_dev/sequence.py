"""Sequence: correlation over a series of frames, one mesh solve per pair.

The drop-in for the reference's ``Sequence`` (geopyv_dev/__init__.py:536-650,
python/src/py_sequence.rs, src/sequence.rs:624-968) on a fixed mesh given by
its nodes and elements (the reference's ``sync`` geometry).  Each new target
frame is decoded on the host and converted (BT.709 grey, 5x5 blur) and
prefiltered on the device; each pair is one ``dic_mesh_solve``; between pairs
the seed moves with the displacement interpolated at it and takes the
interpolated warp as its next first guess (``guide``,
sequence.rs:1062-1097).  A pair that fails falls back to the previous frame
as its reference (``sequential`` does so after every pair), optionally with
the tolerance override (sequence.rs:790-830).
"""

import os
import re

import numpy as np

from . import _native
from .mesh import element_centroids

_IMAGE_EXTS = (".jpg", ".jpeg", ".png")


def _last_number(name):
    """last_number_in_name (sequence.rs:98-104): the last run of digits."""
    runs = re.findall(r"\d+", name)
    return int(runs[-1]) if runs else 0


def image_files(image_dir):
    """The frames of a directory in the reference's order (from_dir,
    sequence.rs:587-610): .jpg/.jpeg/.png by the last number in the stem
    (ties by name)."""
    try:
        names = sorted(os.listdir(image_dir))
    except OSError as exc:
        raise RuntimeError(f"file not found: {image_dir}: {exc}") from None
    files = [n for n in names if os.path.splitext(n)[1].lower() in _IMAGE_EXTS]
    files.sort(key=lambda n: _last_number(os.path.splitext(n)[0]))
    return [os.path.join(image_dir, n) for n in files]


def _local(coord, en):
    """Area coordinates (zeta, eta, theta) of a point in a corner triangle and
    the denominator determinant (local_coordinates, particle.rs:137-183)."""
    def det3(a, b, c, d, e, f, g, h, i):
        return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g)
    px, py = coord
    (n0x, n0y), (n1x, n1y), (n2x, n2y) = en[0], en[1], en[2]
    den = det3(1.0, 1.0, 1.0, n0x, n1x, n2x, n0y, n1y, n2y)
    z = det3(1.0, 1.0, 1.0, px, n1x, n2x, py, n1y, n2y) / den
    e = det3(1.0, 1.0, 1.0, px, n2x, n0x, py, n2y, n0y) / den
    return z, e, 1.0 - z - e, den


def element_locator(coord, nodes, elements, centroids):
    """The element holding a point (particle.rs:968-990): elements in order of
    squared centroid distance, the first whose corner triangle holds it;
    else the nearest, adrift.  Returns (index, adrift)."""
    d2 = (centroids[:, 0] - coord[0]) ** 2 + (centroids[:, 1] - coord[1]) ** 2
    for i in np.argsort(d2, kind="stable"):
        z, e, t, _ = _local(coord, nodes[elements[i, :3]])
        if z >= 0.0 and e >= 0.0 and t >= 0.0:
            return int(i), False
    return int(np.argmin(d2)), True


def warp_increment(coord, en, ed, mesh_order):
    """The warp of the displacement field at a point of one element
    (warp_increment, particle.rs:254-320): [u, v, u_x, v_x, u_y, v_y] and for
    6-node elements the second gradients, from the shape functions at the
    point's area coordinates."""
    z, e, t, den = _local(coord, en)
    if mesh_order == 1:
        N = np.array([z, e, t])
        dN = np.array([[1.0, 0.0, -1.0], [0.0, 1.0, -1.0]])
    else:
        N = np.array([z * (2.0 * z - 1.0), e * (2.0 * e - 1.0), t * (2.0 * t - 1.0), 4.0 * z * e, 4.0 * e * t,
                      4.0 * t * z])
        dN = np.array([[4.0 * z - 1.0, 0.0, 1.0 - 4.0 * t, 4.0 * e, -4.0 * e, 4.0 * (t - z)],
                       [0.0, 4.0 * e - 1.0, 1.0 - 4.0 * t, 4.0 * z, 4.0 * (t - e), -4.0 * z]])
    out = np.zeros(6 * mesh_order)
    out[:2] = N @ ed
    jx, ju = dN @ en, dN @ ed
    idet = 1.0 / (jx[0, 0] * jx[1, 1] - jx[0, 1] * jx[1, 0])
    ij = np.array([[jx[1, 1] * idet, -jx[0, 1] * idet], [-jx[1, 0] * idet, jx[0, 0] * idet]])
    out[2:6] = (ij @ ju).ravel()  # [du/dx, dv/dx, du/dy, dv/dy]
    if mesh_order == 2:
        d2N = np.array([[4.0, 0.0, 4.0, 0.0, 0.0, -8.0], [0.0, 0.0, 4.0, 4.0, -4.0, -4.0],
                        [0.0, 4.0, 4.0, 0.0, -8.0, 0.0]])
        ku = d2N @ ed
        g = np.array([[en[1, 1] - en[2, 1], en[2, 1] - en[0, 1]], [en[2, 0] - en[1, 0], en[0, 0] - en[2, 0]]]) / den
        kx = np.array([[g[0, 0] ** 2, 2.0 * g[0, 0] * g[0, 1], g[0, 1] ** 2],
                       [g[0, 0] * g[1, 0], g[0, 0] * g[1, 1] + g[0, 1] * g[1, 0], g[0, 1] * g[1, 1]],
                       [g[1, 0] ** 2, 2.0 * g[1, 0] * g[1, 1], g[1, 1] ** 2]])
        out[6:] = (kx @ ku).ravel()
    return out


def deformation_preconditioning(nodes, elements, displacements, mesh_order, seed_coord, subset_order):
    """The seed's move and next first guess from a solved pair
    (sequence.rs:1062-1097): the warp increment at the seed in the element
    holding it, strains clipped to [-0.99, 0.99] as a particle stores them
    (particle.rs:1428-1437); the first 6 * min(mesh_order, subset_order)
    terms become the warp, the rest zero."""
    cen = element_centroids(nodes, elements)
    idx, _ = element_locator(seed_coord, nodes, elements, cen)
    row = elements[idx]
    inc = warp_increment(seed_coord, nodes[row], displacements[row], mesh_order)
    inc[2:] = np.clip(inc[2:], -0.99, 0.99)
    warp = np.zeros(6 * subset_order)
    n_copy = 6 * min(mesh_order, subset_order)
    warp[:n_copy] = inc[:n_copy]
    return inc[:2].copy(), warp


class MeshSolution:
    """One solved pair: the per-node and per-element arrays of MeshSolution
    (mesh.rs:54-113) plus the pair's frame indices."""

    def __init__(self, f_index, g_index, **arrays):
        self.f_index, self.g_index = f_index, g_index
        for k, v in arrays.items():
            setattr(self, k, v)


class Sequence:
    """Multi-pair DIC over frames (file paths, a directory of images, or
    grey arrays) on a fixed mesh; ``solve()`` mutates in place and returns
    None."""

    def __init__(self, image_dir=None, boundary=None, target_nodes=None, size=(1.0, 1000.0), exclusions=None,
                 mesh_order=1, time=None, soil_properties=None, *, frames=None, nodes=None, elements=None,
                 exclusions_hard=None, boundary_hard=False):
        if nodes is None or elements is None:
            raise NotImplementedError("automatic mesh generation (geometry/meshing.rs) is outside the "
                                      "correlation engine: pass nodes= and elements= (see structured_mesh)")
        if frames is None:
            if image_dir is None:
                raise ValueError("one of image_dir or frames is required")
            frames = image_files(image_dir)
        self._frames = list(frames)
        if len(self._frames) < 2:
            raise RuntimeError("invalid input: a sequence needs at least 2 images")
        if mesh_order not in (1, 2):
            raise ValueError("mesh_order must be 1 or 2")
        self._nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 2)
        self._elements = np.ascontiguousarray(elements, dtype=np.int64)
        self._mesh_order = int(mesh_order)
        self._boundary, self._exclusions = boundary, exclusions
        self._exclusions_hard, self._boundary_hard = exclusions_hard, boundary_hard
        self._time = time
        self._soil = soil_properties
        self._solution = None

    def _image(self, k, border):
        from . import Image
        fr = self._frames[k]
        if isinstance(fr, (str, os.PathLike)):
            return Image(filepath=fr, border=border)
        return Image(image_gs=np.asarray(fr, dtype=np.float64), border=border)

    def _label(self, k):
        """A frame's name in pair labels: its file stem, else its index."""
        fr = self._frames[k]
        if isinstance(fr, (str, os.PathLike)):
            return os.path.splitext(os.path.basename(os.fspath(fr)))[0]
        return str(k)

    def solve(self, local_mask, seed_coord, seed_warp=None, solver=None, solver_options=None, max_norm=1e-5,
              max_iterations=50, subset_order=1, tolerance=0.75, seed_tolerance=0.9, guide=True, sequential=False,
              sync=True, override=False, border=20, precision="f64", progress=None, **kwargs):
        """The frame loop of sequence.rs:624-968 for uniform masking,
        deformable texture and uniform topology, on this sequence's mesh
        (``sync`` is implied: the geometry is reused for every pair).
        ``progress`` is Sequence::solve's observer: ``pair_start(pair,
        n_pairs, label)`` before every pair (sequence.rs:788-791), then the
        mesh solve's ``mesh_start``/``node_done``/``cancelled``; a
        cancellation raises RuntimeError("solve cancelled") and stores
        nothing (it is not a failed pair to fall back from)."""
        from . import Mask, _checked_solver, _resolve_solver
        from .mesh import Mesh
        if not isinstance(local_mask, Mask) or local_mask.mask_type != "local":
            raise TypeError("local_mask must be a Mask")
        method = _checked_solver(_resolve_solver(solver, kwargs, "method"))
        if kwargs:
            raise TypeError(f"unexpected keyword arguments: {sorted(kwargs)}")
        n = len(self._frames)
        n_pairs = n - 1
        m = 6 * int(subset_order)
        warp = np.zeros(m)
        if seed_warp is not None:
            v = np.asarray(seed_warp, dtype=np.float64).ravel()[:m]
            warp[:len(v)] = v
        coord = np.array([float(seed_coord[0]), float(seed_coord[1])])
        solutions, override_log = [], []
        reference_updates = [False] * n_pairs
        f_index, g_index = 0, 1
        f_img, g_img = self._image(0, border), self._image(1, border)
        mesh_override = pending_ref = False
        unsolvable = all_solved = False
        mesh = Mesh(f_img=f_img, g_img=g_img, nodes=self._nodes, elements=self._elements,
                    mesh_order=self._mesh_order, boundary=self._boundary, boundary_hard=self._boundary_hard,
                    exclusions=self._exclusions, exclusions_hard=self._exclusions_hard)
        while True:
            mesh._use_images(f_img, g_img)
            if progress is not None:
                if callable(getattr(progress, "pair_start", None)):
                    progress.pair_start(len(solutions), n_pairs, f"{self._label(f_index)}→{self._label(g_index)}")
                if callable(getattr(progress, "cancelled", None)) and progress.cancelled():
                    raise _native.SolveCancelled()
            # override: tolerance 0 and the quality gate bypassed (sequence.rs:751-766)
            tol = 0.0 if mesh_override else tolerance
            try:
                mesh.solve(local_mask, seed_coord=coord, seed_warp=warp, solver=method,
                           solver_options=solver_options, max_norm=max_norm, max_iterations=max_iterations,
                           subset_order=subset_order, tolerance=tol, seed_tolerance=seed_tolerance,
                           override_active=mesh_override, precision=precision, progress=progress)
            except _native.SolveCancelled:
                raise
            except RuntimeError:
                if f_index + 1 < g_index:
                    # a non-consecutive pair failed: the previous frame becomes
                    # the reference (sequence.rs:802-815)
                    f_index = g_index - 1
                    f_img = self._image(f_index, border)
                    pending_ref = True
                    mesh_override = bool(override)
                    continue
                unsolvable = True  # a consecutive pair cannot be solved: curtail
                break
            sol = MeshSolution(f_index, g_index, nodes=mesh.nodes, elements=mesh.elements,
                               displacements=mesh.displacements, p=mesh.p, c_zncc=mesh.c_zncc,
                               iterations=mesh.iterations, norms=mesh.norms, quality_ok=mesh.quality_ok,
                               warps=mesh.warps, areas=mesh.areas, centroids=mesh.centroids,
                               seed_node=mesh.seed_node, seed_coord=coord.copy(), seed_warp=warp.copy())
            if mesh_override:
                if np.any(sol.c_zncc < tolerance):
                    override_log.append(g_index)
                mesh_override = False
            if pending_ref:
                if len(solutions) < n_pairs:
                    reference_updates[len(solutions)] = True
                pending_ref = False
            solutions.append(sol)
            prev_g = g_img
            g_index += 1
            if g_index >= n:
                all_solved = True
                break
            g_img = self._image(g_index, border)
            if guide:
                disp, new_warp = deformation_preconditioning(sol.nodes, sol.elements, sol.displacements,
                                                             self._mesh_order, coord, int(subset_order))
                coord = coord + disp
                n_copy = 6 * min(self._mesh_order, int(subset_order))
                warp[:n_copy] = new_warp[:n_copy]
                warp[n_copy:] = 0.0
            if sequential:
                f_index = g_index - 1
                f_img = prev_g
                if g_index - 1 < n_pairs:
                    reference_updates[g_index - 1] = True
        self._solution = dict(mesh_solutions=solutions, all_converged=all_solved, unsolvable=unsolvable,
                              override_log=override_log, reference_updates=reference_updates)

    def _get(self, key):
        if self._solution is None:
            raise AttributeError("sequence not solved: call solve() first")
        return self._solution[key]

    @property
    def solved(self):
        return self._solution is not None

    mesh_solutions = property(lambda self: self._get("mesh_solutions"))
    meshes = mesh_solutions
    all_converged = property(lambda self: self._get("all_converged"))
    unsolvable = property(lambda self: self._get("unsolvable"))
    override_log = property(lambda self: self._get("override_log"))
    reference_updates = property(lambda self: self._get("reference_updates"))

    def mesh_solution_at(self, idx):
        return self.mesh_solutions[idx]

    def all_c_zncc(self):
        return [s.c_zncc for s in self.mesh_solutions]

    def __repr__(self):
        return f"Sequence({len(self._frames)} frames, {len(self._nodes)} nodes, " \
               f"{'solved' if self.solved else 'unsolved'})"

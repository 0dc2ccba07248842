"""Synthetic mesh fixtures restated from the reference's mesh tests
(src/mesh.rs:2906-3096): a smooth texture, rigid bilinear shifts (uniform,
and a step in u at a given row) and the fixed 9 x 9 structured grid of the
band test.  Shared by the CPU and GPU mesh tests."""

import numpy as np


def texture(size):
    """make_synthetic_texture_mesh (mesh.rs:2906-2920): 128 + 40 v(x, y)."""
    y, x = np.mgrid[0:size, 0:size].astype(np.float64)
    v = np.sin(0.7 * x) * np.cos(0.5 * y) + np.sin(0.31 * x + 0.2 * y) + np.cos(0.13 * x - 0.42 * y)
    return 128.0 + 40.0 * v


def _shift_rows(src, u_of_row):
    h, w = src.shape
    out = np.empty_like(src)
    x = np.arange(w, dtype=np.float64)
    for y in range(h):
        sx = x - u_of_row(y)
        x0 = np.floor(sx)
        frac = sx - x0
        x0i = np.minimum(np.maximum(x0, 0.0).astype(np.int64), w - 1)
        x1i = np.minimum(np.maximum(x0 + 1.0, 0.0).astype(np.int64), w - 1)
        out[y] = src[y, x0i] * (1.0 - frac) + src[y, x1i] * frac
    return out


def shift_uniform(src, u):
    """shift_image_uniform (mesh.rs:2926-2941)."""
    return _shift_rows(src, lambda y: u)


def shift_band(src, split_row, u_top, u_bottom):
    """shift_image_band (mesh.rs:2978-2994)."""
    return _shift_rows(src, lambda y: u_top if y < split_row else u_bottom)


def band_grid(n_side=9, at0=20.0, step=7.5):
    """The fixed grid of make_band_test_mesh_and_solve_inputs
    (mesh.rs:3035-3050): nodes row-major, two triangles per cell."""
    at = lambda k: at0 + step * k
    nodes = np.array([[at(i % n_side), at(i // n_side)] for i in range(n_side * n_side)])
    idx = lambda c, r: r * n_side + c
    el = []
    for r in range(n_side - 1):
        for c in range(n_side - 1):
            el.append([idx(c, r), idx(c + 1, r), idx(c, r + 1)])
            el.append([idx(c + 1, r), idx(c + 1, r + 1), idx(c, r + 1)])
    return nodes, np.array(el, dtype=np.int64)


def circle_template(radius):
    """circle local mask offsets as one template (tpl_index, tpl_start, xy)."""
    import oracle as O
    xy, _ = O.mask_local("circle", radius)
    return xy, np.array([0, len(xy)], dtype=np.int64)

"""Seeded synthetic DIC workloads: Gaussian speckle images, analytic
displacement fields, and the subset layouts solved on them.

Every image is I(x) = Σ_k a_k·exp(-|x - c_k|²/(2σ_k²)), scaled to 0-255 by
one factor fixed on the reference image, plus Gaussian grey-level noise,
rounded and clipped to 8 bits.  Targets are the same speckle sum evaluated
at the inverse-mapped position X(x) (x = X + u(X)), so the displacement
field is known exactly.  Pixel (i, j) sits at x = j, y = i.

Random streams: numpy Generator(PCG64(seed)) spawned into independent
children for centres, amplitudes, widths, reference noise and target noise.
"""

import math

import numpy as np


def _streams(seed, n=6):
    return [np.random.Generator(np.random.PCG64(s)) for s in np.random.SeedSequence(seed).spawn(n)]


def _render(shape, centres, amps, sigmas, inv_map=None, fwd_centre=None, max_stretch=1.0, chunk=4096):
    """Sum of Gaussian blobs on an (h, w) grid.

    inv_map(xs, ys) -> (X, Y) maps target pixel positions to reference
    positions; fwd_centre(c) gives where a reference centre lands in the
    target, which centres the evaluation window.
    """
    h, w = shape
    out = np.zeros(h * w, dtype=np.float64)
    smax = float(np.max(sigmas))
    rad = int(math.ceil(4.5 * smax * max_stretch)) + 1
    offs = np.arange(-rad, rad + 1)
    oy, ox = np.meshgrid(offs, offs, indexing="ij")
    oy = oy.ravel()
    ox = ox.ravel()
    cen_t = centres if fwd_centre is None else fwd_centre(centres)
    for s in range(0, len(centres), chunk):
        c = centres[s:s + chunk]
        ct = cen_t[s:s + chunk]
        a = amps[s:s + chunk]
        sg = sigmas[s:s + chunk]
        px = np.floor(ct[:, 0])[:, None].astype(np.int64) + ox[None, :]
        py = np.floor(ct[:, 1])[:, None].astype(np.int64) + oy[None, :]
        ok = (px >= 0) & (px < w) & (py >= 0) & (py < h)
        xs = px.astype(np.float64)
        ys = py.astype(np.float64)
        if inv_map is not None:
            xs, ys = inv_map(xs, ys)
        d2 = (xs - c[:, 0:1]) ** 2 + (ys - c[:, 1:2]) ** 2
        val = a[:, None] * np.exp(-d2 / (2.0 * sg[:, None] ** 2))
        idx = (py * w + px)[ok]
        out += np.bincount(idx, weights=val[ok], minlength=h * w)
    return out.reshape(h, w)


def _quantise(img, scale, noise_sigma, rng):
    v = img * scale + rng.normal(0.0, noise_sigma, img.shape)
    return np.clip(np.rint(v), 0, 255).astype(np.uint8)


class Speckle:
    """A seeded speckle pattern that can be rendered under a warp."""

    def __init__(self, shape, n_speckles, sigma, seed, amp=(0.5, 1.0)):
        self.shape = tuple(shape)
        h, w = self.shape
        rc, ra, rs, self.rn_ref, self.rn_tar, _ = _streams(seed)
        self.centres = np.column_stack([rc.uniform(0, w, n_speckles), rc.uniform(0, h, n_speckles)])
        self.amps = ra.uniform(amp[0], amp[1], n_speckles)
        if np.isscalar(sigma):
            self.sigmas = np.full(n_speckles, float(sigma))
        else:
            self.sigmas = rs.uniform(sigma[0], sigma[1], n_speckles)
        clean = _render(self.shape, self.centres, self.amps, self.sigmas)
        self.scale = 255.0 / clean.max()
        self.clean_ref = clean

    def reference(self, noise_sigma):
        return _quantise(self.clean_ref, self.scale, noise_sigma, self.rn_ref)

    def target(self, field, noise_sigma, rng=None):
        clean = _render(self.shape, self.centres, self.amps, self.sigmas, inv_map=field.inverse,
                        fwd_centre=field.forward, max_stretch=field.max_stretch)
        return _quantise(clean, self.scale, noise_sigma, rng if rng is not None else self.rn_tar)


class Translation:
    max_stretch = 1.0

    def __init__(self, u, v):
        self.u, self.v = float(u), float(v)

    def displacement(self, X, Y):
        return np.full_like(X, self.u), np.full_like(Y, self.v)

    def forward(self, c):
        return c + np.array([self.u, self.v])

    def inverse(self, xs, ys):
        return xs - self.u, ys - self.v

    def warp(self, X, Y, order=1):
        n = len(X)
        p = np.zeros((n, 12))
        p[:, 0] = self.u
        p[:, 1] = self.v
        return p


class Affine:
    """u(X) = G (X - c) with G = [[u_x, u_y], [v_x, v_y]]."""

    def __init__(self, centre, u_x, u_y, v_x, v_y):
        self.c = np.asarray(centre, dtype=np.float64)
        self.G = np.array([[u_x, u_y], [v_x, v_y]], dtype=np.float64)
        self.F = np.eye(2) + self.G
        self.Finv = np.linalg.inv(self.F)
        self.max_stretch = float(np.linalg.norm(self.F, 2))

    def displacement(self, X, Y):
        dx, dy = X - self.c[0], Y - self.c[1]
        return self.G[0, 0] * dx + self.G[0, 1] * dy, self.G[1, 0] * dx + self.G[1, 1] * dy

    def forward(self, cen):
        d = cen - self.c
        return self.c + d @ self.F.T

    def inverse(self, xs, ys):
        dx, dy = xs - self.c[0], ys - self.c[1]
        return (self.c[0] + self.Finv[0, 0] * dx + self.Finv[0, 1] * dy,
                self.c[1] + self.Finv[1, 0] * dx + self.Finv[1, 1] * dy)

    def warp(self, X, Y, order=1):
        """True warp vector [u, v, u_x, v_x, u_y, v_y, 0...] at reference points."""
        u, v = self.displacement(X, Y)
        p = np.zeros((len(X), 12))
        p[:, 0], p[:, 1] = u, v
        p[:, 2], p[:, 3], p[:, 4], p[:, 5] = self.G[0, 0], self.G[1, 0], self.G[0, 1], self.G[1, 1]
        return p


def grid_coords(shape, pitch, margin, n_side=None):
    """Subset centres on a regular grid (x = column, y = row)."""
    h, w = shape
    xs = np.arange(margin, w - margin + 1, pitch, dtype=np.float64)
    ys = np.arange(margin, h - margin + 1, pitch, dtype=np.float64)
    if n_side is not None:
        xs = xs[:n_side]
        ys = ys[:n_side]
    X, Y = np.meshgrid(xs, ys)
    return np.column_stack([X.ravel(), Y.ravel()])


def single_subset_pair(seed=0):
    """One circle subset r = 25 at the centre of a 512² speckle pair under
    u = (0.37, -0.52) px: 3000 speckles σ = 2 px, a ~ U[0.5, 1], noise σ = 2
    grey levels."""
    sp = Speckle((512, 512), 3000, 2.0, seed)
    field = Translation(0.37, -0.52)
    return dict(ref=sp.reference(2.0), tar=sp.target(field, 2.0), field=field,
                coords=np.array([[256.0, 256.0]]), radius=25, order=1)


def uniform_mesh_pair(seed=1, n_side=200, shape=(2048, 2048)):
    """40 000 circle subsets r = 20 on a 10 px grid (x, y in 29..2019) of a
    2048² pair: 50 000 speckles σ = 2.5 px, noise σ = 2 grey levels, affine
    field about the image centre with u_x = 0.010, v_y = -0.003 and simple
    shear u_y = 0.005 (v_x = 0).  p0 = the true displacement rounded to whole
    pixels with zero gradients; the truth is returned for error reporting."""
    h, w = shape
    sp = Speckle(shape, int(round(50000 * h * w / 2048 ** 2)), 2.5, seed)
    field = Affine(((w - 1) / 2.0, (h - 1) / 2.0), u_x=0.010, u_y=0.005, v_x=0.0, v_y=-0.003)
    coords = grid_coords(shape, 10, 29, n_side)
    truth = field.warp(coords[:, 0], coords[:, 1])
    p0 = np.zeros((len(coords), 12))
    p0[:, 0] = np.rint(truth[:, 0])
    p0[:, 1] = np.rint(truth[:, 1])
    return dict(ref=sp.reference(2.0), tar=sp.target(field, 2.0), field=field, coords=coords, radius=20,
                order=1, p0=p0, truth=truth)

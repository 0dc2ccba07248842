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
        inv = field.inverse
        if hasattr(field, "inverse_grid"):
            # the map is evaluated at integer pixels only: invert once per pixel
            h, w = self.shape
            Xg, Yg = field.inverse_grid(self.shape)

            def inv(xs, ys):
                ix = np.clip(xs.astype(np.int64), 0, w - 1)
                iy = np.clip(ys.astype(np.int64), 0, h - 1)
                return Xg[iy, ix], Yg[iy, ix]
        clean = _render(self.shape, self.centres, self.amps, self.sigmas, inv_map=inv,
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


class ShearBand:
    """A Gaussian shear band plus a vertical sinusoid.

    The band's centre line runs through `centre` at `angle` degrees; with
    t = (cos a, sin a) along it, n = (-sin a, cos a) across it and
    d = (X - c)·n, the tangential slip is s(d) = slip·exp(-d²/(2σ²)) with
    σ = width / (2√(2 ln 2)) (full width at half maximum = width).  Then
        u = s(d)·t_x,    v = s(d)·t_y + amp·sin(2π X / period).
    """

    def __init__(self, centre, angle=30.0, width=60.0, slip=4.0, amp=0.8, period=512.0):
        a = math.radians(angle)
        self.c = np.asarray(centre, dtype=np.float64)
        self.t = np.array([math.cos(a), math.sin(a)])
        self.n = np.array([-math.sin(a), math.cos(a)])
        self.sigma = width / (2.0 * math.sqrt(2.0 * math.log(2.0)))
        self.slip, self.amp, self.k = float(slip), float(amp), 2.0 * math.pi / float(period)
        self.max_stretch = 1.0 + self.slip / self.sigma + self.amp * self.k

    def _s(self, X, Y):
        d = (X - self.c[0]) * self.n[0] + (Y - self.c[1]) * self.n[1]
        s = self.slip * np.exp(-0.5 * (d / self.sigma) ** 2)
        return d, s

    def displacement(self, X, Y):
        _, s = self._s(X, Y)
        return s * self.t[0], s * self.t[1] + self.amp * np.sin(self.k * X)

    def forward(self, cen):
        u, v = self.displacement(cen[:, 0], cen[:, 1])
        return cen + np.column_stack([u, v])

    def inverse(self, xs, ys):
        """X with X + u(X) = x, by fixed-point iteration to 1e-10 px."""
        X, Y = np.array(xs, dtype=np.float64), np.array(ys, dtype=np.float64)
        for _ in range(200):
            u, v = self.displacement(X, Y)
            Xn, Yn = xs - u, ys - v
            delta = max(float(np.max(np.abs(Xn - X), initial=0.0)), float(np.max(np.abs(Yn - Y), initial=0.0)))
            X, Y = Xn, Yn
            if delta < 1e-10:
                break
        return X, Y

    def inverse_grid(self, shape):
        """The inverse map at every pixel of an (h, w) grid, row by row."""
        h, w = shape
        X = np.empty((h, w))
        Y = np.empty((h, w))
        xs = np.arange(w, dtype=np.float64)
        for i in range(0, h, 256):
            ys = np.arange(i, min(h, i + 256), dtype=np.float64)
            gx, gy = np.meshgrid(xs, ys)
            a, b = self.inverse(gx, gy)
            X[i:i + len(ys)] = a
            Y[i:i + len(ys)] = b
        return X, Y

    def warp(self, X, Y, order=2):
        """True warp vector [u, v, u_x, v_x, u_y, v_y, u_xx, v_xx, u_xy, v_xy, u_yy, v_yy]."""
        d, s = self._s(X, Y)
        s1 = -s * d / self.sigma ** 2                          # ds/dd
        s2 = s * (d * d / self.sigma ** 4 - 1.0 / self.sigma ** 2)  # d²s/dd²
        nx, ny = self.n
        tx, ty = self.t
        kx = self.k * X
        p = np.zeros((len(X), 12))
        p[:, 0] = s * tx
        p[:, 1] = s * ty + self.amp * np.sin(kx)
        p[:, 2] = s1 * nx * tx
        p[:, 3] = s1 * nx * ty + self.amp * self.k * np.cos(kx)
        p[:, 4] = s1 * ny * tx
        p[:, 5] = s1 * ny * ty
        if order >= 2:
            p[:, 6] = s2 * nx * nx * tx
            p[:, 7] = s2 * nx * nx * ty - self.amp * self.k ** 2 * np.sin(kx)
            p[:, 8] = s2 * nx * ny * tx
            p[:, 9] = s2 * nx * ny * ty
            p[:, 10] = s2 * ny * ny * tx
            p[:, 11] = s2 * ny * ny * ty
        return p


def band_nodes(shape, field, n_target, margin, near=100.0):
    """Adaptive node layout: a square grid over the image plus a second grid
    offset by half a pitch within `near` px of the band's centre line (twice
    the density there), pitch chosen so that about n_target nodes result."""
    h, w = shape
    area = (w - 2 * margin) * (h - 2 * margin)

    def layout(pitch):
        xs = np.arange(margin, w - margin + 1e-9, pitch)
        ys = np.arange(margin, h - margin + 1e-9, pitch)
        X, Y = np.meshgrid(xs, ys)
        base = np.column_stack([X.ravel(), Y.ravel()])
        off = base + 0.5 * pitch
        off = off[(off[:, 0] <= w - margin) & (off[:, 1] <= h - margin)]
        d = (off[:, 0] - field.c[0]) * field.n[0] + (off[:, 1] - field.c[1]) * field.n[1]
        return np.vstack([base, off[np.abs(d) <= near]])

    pitch = math.sqrt(area / n_target)
    for _ in range(30):
        pts = layout(pitch)
        if abs(len(pts) - n_target) <= 0.002 * n_target:
            break
        pitch *= math.sqrt(len(pts) / n_target)
    return np.rint(pts * 4.0) / 4.0  # quarter-pixel centres


def shear_band_pair(seed=2, shape=(3000, 4096), n_speckles=150000, n_subsets=150000, radii=(15, 40)):
    """Uneven subsets, second-order warp: a 4096 x 3000 pair with 150 000
    speckles (σ ~ U[1.5, 3] px per speckle, noise σ = 2 grey levels) under a
    Gaussian shear band (full width 60 px, 30°, peak tangential slip 4 px)
    plus v += 0.8 sin(2πx/512); about 150 000 subsets on adaptive nodes
    (twice the density within 100 px of the band) with radii uniform in
    15..40 px; p0 = the true displacement rounded to whole pixels, zero
    gradients.  Smaller shapes scale the speckle and subset counts."""
    h, w = shape
    scale = (h * w) / (3000.0 * 4096.0)
    sp = Speckle(shape, max(1, int(round(n_speckles * scale))), (1.5, 3.0), seed)
    field = ShearBand(((w - 1) / 2.0, (h - 1) / 2.0))
    nsub = max(1, int(round(n_subsets * scale)))
    rmax = radii[1]
    coords = band_nodes(shape, field, nsub, margin=rmax + 6)
    rr = _streams(seed, 7)[6]
    r = rr.integers(radii[0], radii[1] + 1, len(coords))
    truth = field.warp(coords[:, 0], coords[:, 1], order=2)
    p0 = np.zeros((len(coords), 12))
    p0[:, 0] = np.rint(truth[:, 0])
    p0[:, 1] = np.rint(truth[:, 1])
    return dict(ref=sp.reference(2.0), tar=sp.target(field, 2.0), field=field, coords=coords, radii=r,
                order=2, p0=p0, truth=truth)


_SEQ_PATTERNS = {}


def _sequence_frame(spec):
    """Frame k of sequence_frames (a worker of its process pool; the
    pattern is rendered once per process)."""
    seed, shape, n_frames, n_speckles, sigma, slip_max, k = spec
    key = (seed, shape, n_speckles, sigma)
    sp = _SEQ_PATTERNS.get(key)
    if sp is None:
        sp = _SEQ_PATTERNS[key] = Speckle(shape, n_speckles, sigma, seed)
    h, w = shape
    slip = slip_max * k / max(1, n_frames - 1)
    field = ShearBand(((w - 1) / 2.0, (h - 1) / 2.0), slip=slip)
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, 1000 + k])))
    return sp.target(field, 2.0, rng=rng), field


def sequence_frames(seed=3, shape=(2048, 2048), n_frames=200, n_speckles=80000, sigma=2.0, slip_max=8.0,
                    progress=None, workers=1):
    """Frames of an incremental sequence: one speckle pattern (σ = 2 px)
    under a shear band whose slip grows linearly from 0 (frame 0) to
    slip_max (last frame); grey-level noise σ = 2 drawn per frame from
    seed(frame).  Returns (frames, fields); frames are 8-bit arrays.  With
    workers > 1 the frames are rendered by that many spawned processes (the
    same frames: each one's noise comes from its own seed)."""
    specs = [(seed, tuple(shape), n_frames, n_speckles, sigma, slip_max, k) for k in range(n_frames)]
    frames, fields = [], []
    if workers > 1:
        import multiprocessing as mp
        with mp.get_context("spawn").Pool(workers) as pool:
            for k, (fr, fd) in enumerate(pool.imap(_sequence_frame, specs)):
                frames.append(fr)
                fields.append(fd)
                if progress is not None:
                    progress(k)
        return frames, fields
    for k, spec in enumerate(specs):
        fr, fd = _sequence_frame(spec)
        frames.append(fr)
        fields.append(fd)
        if progress is not None:
            progress(k)
    return frames, fields

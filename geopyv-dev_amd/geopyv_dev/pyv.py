"""The `.pyv` file format of geopyv-dev for mesh solutions (src/io.rs:1-25,
168, 2796-2816): the magic ``b"GPYV"``, one version byte (0x17, the current
schema), then the bincode v2 encoding (standard configuration) of
``GeopyvObject::Mesh(MeshSolution)`` compressed with zstd at level 3.

bincode's standard configuration writes integers wider than a byte as
variable-length integers (below 251: one byte; else a marker byte 251 / 252
/ 253 followed by the value as a little-endian u16 / u32 / u64), ``u8`` and
``bool`` as one byte, floats as little-endian IEEE words, strings and
sequences as their length followed by the items, ``Option`` as a 0 / 1 byte
followed by the value, enum variants as their index, and structs and
tuples as their fields in order.  ndarray arrays serialise as the struct
``{v: u8 = 1, dim: (usize, ...), data: [items in row-major order]}``.  The
MeshSolution field order is mesh.rs:54-113, SolveConfig and SeedConfig
mesh.rs:221-276, LayerRgConfig mesh.rs:341-351, MaskShape masks.rs:21-31.

zstd comes from the system's ``libzstd.so.1`` through ctypes.
"""

import ctypes
import ctypes.util
import os
import struct

import numpy as np

MAGIC = b"GPYV"
VERSION = 0x17          # io.rs:168, CURRENT_VERSION (schema v18)
ZSTD_LEVEL = 3          # io.rs:2800
VARIANT_MESH = 1        # GeopyvObject::{Subset, Mesh, Field, Sequence, Particle, Speckle, Calibration}
SHAPES = ("circle", "square", "semicircle")  # MaskShape variant order
METHODS = ("icgn", "fagn")
PRECONDITIONING = ("RG", "layer-RG")

_zstd = None


def _lib():
    global _zstd
    if _zstd is None:
        name = ctypes.util.find_library("zstd") or "libzstd.so.1"
        L = ctypes.CDLL(name)
        L.ZSTD_compress.restype = ctypes.c_size_t
        L.ZSTD_compress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.ZSTD_compressBound.restype = ctypes.c_size_t
        L.ZSTD_compressBound.argtypes = [ctypes.c_size_t]
        L.ZSTD_decompress.restype = ctypes.c_size_t
        L.ZSTD_decompress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
        L.ZSTD_getFrameContentSize.restype = ctypes.c_ulonglong
        L.ZSTD_getFrameContentSize.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        L.ZSTD_isError.restype = ctypes.c_uint
        L.ZSTD_isError.argtypes = [ctypes.c_size_t]
        _zstd = L
    return _zstd


def zstd_compress(data, level=ZSTD_LEVEL):
    L = _lib()
    cap = L.ZSTD_compressBound(len(data))
    out = ctypes.create_string_buffer(cap)
    n = L.ZSTD_compress(out, cap, data, len(data), level)
    if L.ZSTD_isError(n):
        raise OSError("zstd compression failed")
    return out.raw[:n]


def zstd_decompress(data):
    L = _lib()
    size = L.ZSTD_getFrameContentSize(data, len(data))
    cap = size if size < (1 << 40) else max(64 << 20, 32 * len(data))
    while True:
        out = ctypes.create_string_buffer(max(int(cap), 1))
        n = L.ZSTD_decompress(out, len(out), data, len(data))
        if not L.ZSTD_isError(n):
            return out.raw[:n]
        if size < (1 << 40) or cap > (1 << 34):
            raise OSError("zstd decompression failed")
        cap *= 4  # frame of unknown size: grow the buffer


# ---------------------------------------------------------------------------
# bincode v2, standard configuration
# ---------------------------------------------------------------------------
class Writer:
    def __init__(self):
        self.parts = []

    def bytes(self):
        return b"".join(self.parts)

    def u8(self, v):
        self.parts.append(struct.pack("<B", int(v)))

    def bool(self, v):
        self.u8(1 if v else 0)

    def varint(self, v):
        v = int(v)
        if v < 0:
            raise ValueError("unsigned integer expected")
        if v < 251:
            self.parts.append(struct.pack("<B", v))
        elif v < 1 << 16:
            self.parts.append(b"\xfb" + struct.pack("<H", v))
        elif v < 1 << 32:
            self.parts.append(b"\xfc" + struct.pack("<I", v))
        else:
            self.parts.append(b"\xfd" + struct.pack("<Q", v))

    def f64(self, v):
        self.parts.append(struct.pack("<d", float(v)))

    def string(self, s):
        b = os.fsencode(s) if not isinstance(s, bytes) else s
        self.varint(len(b))
        self.parts.append(b)

    def f64s(self, a):
        a = np.ascontiguousarray(a, dtype="<f8").ravel()
        self.varint(a.size)
        self.parts.append(a.tobytes())

    def varints(self, a):
        a = np.asarray(a).ravel()
        self.varint(a.size)
        for v in a:
            self.varint(v)

    def ndarray_f64(self, a):
        a = np.asarray(a, dtype=np.float64)
        self.u8(1)
        for d in a.shape:
            self.varint(d)
        self.f64s(a)

    def ndarray_uint(self, a):
        a = np.asarray(a)
        self.u8(1)
        for d in a.shape:
            self.varint(d)
        self.varints(a)


class Reader:
    def __init__(self, data):
        self.b = memoryview(data)
        self.i = 0

    def done(self):
        return self.i == len(self.b)

    def take(self, n):
        if self.i + n > len(self.b):
            raise ValueError("truncated bincode stream")
        out = self.b[self.i:self.i + n]
        self.i += n
        return out

    def u8(self):
        return self.take(1)[0]

    def bool(self):
        v = self.u8()
        if v > 1:
            raise ValueError("invalid bool")
        return v == 1

    def varint(self):
        m = self.u8()
        if m < 251:
            return m
        if m == 251:
            return struct.unpack("<H", self.take(2))[0]
        if m == 252:
            return struct.unpack("<I", self.take(4))[0]
        if m == 253:
            return struct.unpack("<Q", self.take(8))[0]
        raise ValueError("unsupported integer marker")

    def f64(self):
        return struct.unpack("<d", self.take(8))[0]

    def string(self):
        return bytes(self.take(self.varint())).decode("utf-8")

    def f64s(self):
        n = self.varint()
        return np.frombuffer(bytes(self.take(8 * n)), dtype="<f8").astype(np.float64)

    def varints(self):
        n = self.varint()
        return np.array([self.varint() for _ in range(n)], dtype=np.int64)

    def option(self):
        v = self.u8()
        if v > 1:
            raise ValueError("invalid Option tag")
        return v == 1

    def ndarray(self, ndim, floats=True):
        if self.u8() != 1:
            raise ValueError("unknown ndarray serialisation version")
        shape = tuple(self.varint() for _ in range(ndim))
        data = self.f64s() if floats else self.varints()
        if data.size != int(np.prod(shape)):
            raise ValueError("ndarray data length does not match its shape")
        return data.reshape(shape)


# ---------------------------------------------------------------------------
# MeshSolution (mesh.rs:54-113)
# ---------------------------------------------------------------------------
class MeshSolution:
    """The arrays and settings a `.pyv` mesh file holds."""

    FIELDS = ("nodes", "elements", "boundary", "exclusions", "centroids", "areas", "warps", "displacements",
              "c_zncc", "p", "seed_node", "mesh_order", "subset_order", "iterations", "norms", "f_img_path",
              "g_img_path", "solve_config", "seed", "template_shape", "template_sizes", "time")

    def __init__(self, **kw):
        for k in self.FIELDS:
            setattr(self, k, kw.get(k))


def boundary_nodes(elements, mesh_order):
    """Boundary node indices of a triangulation, in order along each
    boundary loop (corner nodes, with the midside nodes of quadratic
    elements between them), starting from the lowest corner index."""
    el = np.asarray(elements, dtype=np.int64)
    edges = {}
    for row in el:
        c = row[:3]
        mids = row[3:6] if mesh_order == 2 and len(row) >= 6 else (None, None, None)
        for k in range(3):
            a, b = int(c[k]), int(c[(k + 1) % 3])
            key = (min(a, b), max(a, b))
            edges.setdefault(key, []).append((a, b, None if mids[k] is None else int(mids[k])))
    nxt = {}
    for key, uses in edges.items():
        if len(uses) == 1:
            a, b, m = uses[0]
            nxt[a] = (b, m)
    out, seen = [], set()
    for start in sorted(nxt):
        if start in seen:
            continue
        cur = start
        while cur not in seen and cur in nxt:
            seen.add(cur)
            out.append(cur)
            b, m = nxt[cur]
            if m is not None:
                out.append(m)
            cur = b
    return np.array(out, dtype=np.int64)


def encode_mesh(sol):
    """bincode of GeopyvObject::Mesh(sol)."""
    w = Writer()
    w.varint(VARIANT_MESH)
    w.ndarray_f64(sol.nodes)
    w.ndarray_uint(np.asarray(sol.elements, dtype=np.int64))
    w.varints(np.asarray(sol.boundary if sol.boundary is not None else [], dtype=np.int64))
    ex = sol.exclusions or []
    w.varint(len(ex))
    for group in ex:
        w.varints(np.asarray(group, dtype=np.int64))
    w.ndarray_f64(sol.centroids)
    w.ndarray_f64(np.asarray(sol.areas, dtype=np.float64).reshape(-1))
    w.ndarray_f64(sol.warps)
    w.ndarray_f64(sol.displacements)
    w.ndarray_f64(np.asarray(sol.c_zncc, dtype=np.float64).reshape(-1))
    w.ndarray_f64(sol.p)
    w.varint(sol.seed_node)
    w.u8(sol.mesh_order)
    w.u8(sol.subset_order)
    w.ndarray_uint(np.asarray(sol.iterations, dtype=np.int64).reshape(-1))
    w.ndarray_f64(np.asarray(sol.norms, dtype=np.float64).reshape(-1))
    w.string(sol.f_img_path or "")
    w.string(sol.g_img_path or "")
    sc = sol.solve_config
    if sc is None:
        w.u8(0)
    else:  # SolveConfig (mesh.rs:221-247)
        w.u8(1)
        w.f64(sc["max_norm"])
        w.varint(sc["max_iterations"])
        w.varint(sc["subset_order"])
        w.f64(sc["tolerance"])
        w.varint(METHODS.index(sc["method"]))
        w.bool(sc.get("override_active", False))
        w.varint(PRECONDITIONING.index(sc.get("preconditioning", "RG")))
        w.varint(sc.get("batch_factor", 4))           # LayerRgConfig (mesh.rs:341-351)
        w.f64(sc.get("root_rel_eps", 0.02))
        mw = sc.get("max_workers")
        if mw is None:
            w.u8(0)
        else:
            w.u8(1)
            w.varint(mw)
        w.varint(0)  # Masking::Uniform
        w.varint(0)  # Texture::Deformable
    sd = sol.seed
    if sd is None:
        w.u8(0)
    else:  # SeedConfig (mesh.rs:266-276)
        w.u8(1)
        w.f64(sd["coord"][0])
        w.f64(sd["coord"][1])
        w.f64s(np.asarray(sd["warp"], dtype=np.float64))
        w.f64(sd["tolerance"])
    if sol.template_shape is None:
        w.u8(0)
    else:
        w.u8(1)
        w.varint(SHAPES.index(sol.template_shape))
    if sol.template_sizes is None:
        w.u8(0)
    else:
        w.u8(1)
        w.ndarray_uint(np.asarray(sol.template_sizes, dtype=np.int64).reshape(-1))
    if sol.time is None:
        w.u8(0)
    else:
        w.u8(1)
        w.f64(sol.time[0])
        w.f64(sol.time[1])
    w.u8(0)  # seams: None (uniform masking)
    w.u8(0)  # tracks: None (outside tracked sequences)
    return w.bytes()


def decode_mesh(data):
    """GeopyvObject::Mesh payload (current schema) -> MeshSolution."""
    r = Reader(data)
    variant = r.varint()
    if variant != VARIANT_MESH:
        raise ValueError(f"not a mesh solution (GeopyvObject variant {variant})")
    s = MeshSolution()
    s.nodes = r.ndarray(2)
    s.elements = r.ndarray(2, floats=False)
    s.boundary = r.varints()
    s.exclusions = [r.varints() for _ in range(r.varint())]
    s.centroids = r.ndarray(2)
    s.areas = r.ndarray(1)
    s.warps = r.ndarray(2)
    s.displacements = r.ndarray(2)
    s.c_zncc = r.ndarray(1)
    s.p = r.ndarray(2)
    s.seed_node = r.varint()
    s.mesh_order = r.u8()
    s.subset_order = r.u8()
    s.iterations = r.ndarray(1, floats=False)
    s.norms = r.ndarray(1)
    s.f_img_path = r.string()
    s.g_img_path = r.string()
    if r.option():
        sc = {"max_norm": r.f64(), "max_iterations": r.varint(), "subset_order": r.varint(), "tolerance": r.f64(),
              "method": METHODS[r.varint()], "override_active": r.bool(),
              "preconditioning": PRECONDITIONING[r.varint()], "batch_factor": r.varint(), "root_rel_eps": r.f64()}
        sc["max_workers"] = r.varint() if r.option() else None
        sc["masking"] = ("uniform", "adaptive")[r.varint()]
        sc["texture"] = ("deformable", "rigid")[r.varint()]
        s.solve_config = sc
    if r.option():
        s.seed = {"coord": [r.f64(), r.f64()], "warp": r.f64s(), "tolerance": r.f64()}
    if r.option():
        s.template_shape = SHAPES[r.varint()]
    if r.option():
        s.template_sizes = r.ndarray(1, floats=False)
    if r.option():
        s.time = (r.f64(), r.f64())
    if r.option() or r.option():
        raise ValueError("adaptive-masking seams and sequence tracks are outside this reader")
    if not r.done():
        raise ValueError("trailing bytes after the mesh solution")
    return s


class PyvError(RuntimeError):
    """Error::InvalidMagic / Error::UnsupportedVersion / Error::Io of
    error.rs, as the PyO3 layer raises them (RuntimeError)."""


def write_pyv(path, payload):
    with open(path, "wb") as fh:
        fh.write(MAGIC)
        fh.write(bytes([VERSION]))
        fh.write(zstd_compress(payload))


def read_pyv(path):
    """(version byte, decompressed payload) of a `.pyv` file."""
    with open(path, "rb") as fh:
        raw = fh.read()
    if raw[:4] != MAGIC:
        raise PyvError("invalid magic bytes: not a .pyv file")
    if len(raw) < 5:
        raise PyvError("truncated .pyv file")
    version = raw[4]
    if version == 0x01:
        return version, raw[5:]
    if version in (0x02, 0x03) or 0x07 <= version <= VERSION:
        return version, zstd_decompress(raw[5:])
    raise PyvError(f"unsupported .pyv format version {version}")


def save_mesh(path, sol):
    write_pyv(path, encode_mesh(sol))


def load_mesh(path):
    version, payload = read_pyv(path)
    if version != VERSION:
        raise PyvError(f"mesh solutions are read at the current schema (version {VERSION:#04x}); "
                       f"this file is {version:#04x}")
    return decode_mesh(payload)

"""The `.pyv` writer and reader (geopyv_dev/pyv.py) against the format of
src/io.rs: the bincode primitives are pinned by decoding the reference's
own `.pyv` file (images/shear/speckle_shear.pyv, read as data when the
reference tree is present) field by field to its last byte; a small
MeshSolution's encoding is compared with bytes assembled by hand from the
bincode v2 rules; files round-trip through zstd; bad magic and unknown
versions are refused as Error::InvalidMagic / UnsupportedVersion are."""

import os
import struct

import numpy as np
import pytest

from geopyv_dev import pyv

REF_PYV = "/root/reference/images/shear/speckle_shear.pyv"


@pytest.mark.skipif(not os.path.exists(REF_PYV), reason="reference tree not present")
def test_reference_speckle_file_decodes_to_its_last_byte():
    # LegacySpeckleV1 (io.rs:738-755), the shape of version-0x03 files
    version, payload = pyv.read_pyv(REF_PYV)
    assert version == 0x03
    r = pyv.Reader(payload)
    assert r.varint() == 5  # GeopyvObject::Speckle
    assert r.string() == "images/shear" and r.string() == "shear" and r.string() == ".jpg"
    assert (r.varint(), r.varint()) == (1001, 1001)  # image_size: the 251-marker u16 form
    assert r.f64() == 10.0 and r.varint() == 7000  # speckle_size, speckle_number
    assert r.varint() in (0, 1)  # progression
    comp = [r.f64() for _ in range(12)]
    origin = [r.f64(), r.f64()]
    assert all(np.isfinite(comp)) and all(np.isfinite(origin))
    mode = r.varint()  # WarpMode: None, Rotation, ShearBand{option, width}, CircularShear{r1, r2, tau}
    if mode == 2:
        if r.varint() == 3:  # ShearBandOption::Smooth{tau}
            r.f64()
        assert r.f64() > 0.0
    elif mode == 3:
        r.f64(), r.f64()
        if r.option():
            r.f64()
    noise_pos, noise_int = r.f64(), r.f64()
    assert noise_pos >= 0.0 and noise_int >= 0.0
    assert r.varint() >= 1  # image_no
    if r.varint() == 1:  # ScaleType::Log{min}
        r.f64()
    assert r.option()
    pos = r.ndarray(2)  # speckle_positions: ndarray's {v, dim, data}
    assert pos.shape == (7000, 2) and np.all((pos >= -200) & (pos <= 1201))
    r.bool()  # solved
    assert r.done()


def _f(*v):
    return b"".join(struct.pack("<d", x) for x in v)


def _tiny():
    nodes = np.array([[10.0, 10.0], [30.0, 10.0], [10.0, 30.0]])
    return pyv.MeshSolution(
        nodes=nodes, elements=np.array([[0, 1, 2]]), boundary=np.array([0, 1, 2]), exclusions=[],
        centroids=np.array([[50.0 / 3, 50.0 / 3]]), areas=np.array([200.0]), warps=np.arange(12.0)[None, :] / 100,
        displacements=np.array([[0.1, 0.2], [0.3, 0.4], [0.5, 0.6]]), c_zncc=np.array([0.9, 0.95, 0.99]),
        p=np.arange(18.0).reshape(3, 6) / 10, seed_node=2, mesh_order=1, subset_order=1,
        iterations=np.array([3, 300, 70000]), norms=np.array([1e-6, 2e-6, 3e-6]), f_img_path="a.png",
        g_img_path="", solve_config=dict(max_norm=1e-5, max_iterations=50, subset_order=1, tolerance=0.75,
                                         method="icgn", override_active=False, preconditioning="layer-RG",
                                         batch_factor=4, root_rel_eps=0.02, max_workers=None),
        seed=dict(coord=[10.0, 30.0], warp=np.zeros(6), tolerance=0.9), template_shape="circle",
        template_sizes=np.array([20, 20, 20]), time=None)


def test_mesh_encoding_matches_hand_assembled_bytes():
    s = _tiny()
    want = b"".join([
        b"\x01",                                               # GeopyvObject::Mesh
        b"\x01\x03\x02\x06" + _f(10, 10, 30, 10, 10, 30),      # nodes: v, dim, len, data
        b"\x01\x01\x03\x03\x00\x01\x02",                       # elements (usize)
        b"\x03\x00\x01\x02",                                   # boundary
        b"\x00",                                               # exclusions
        b"\x01\x01\x02\x02" + _f(50.0 / 3, 50.0 / 3),          # centroids
        b"\x01\x01\x01" + _f(200.0),                           # areas (1-D dim)
        b"\x01\x01\x0c\x0c" + _f(*(np.arange(12.0) / 100)),    # warps
        b"\x01\x03\x02\x06" + _f(0.1, 0.2, 0.3, 0.4, 0.5, 0.6),
        b"\x01\x03\x03" + _f(0.9, 0.95, 0.99),                 # c_zncc
        b"\x01\x03\x06\x12" + _f(*(np.arange(18.0) / 10)),     # p
        b"\x02\x01\x01",                                       # seed_node, mesh_order u8, subset_order u8
        b"\x01\x03\x03" + b"\x03" + b"\xfb\x2c\x01" + b"\xfc\x70\x11\x01\x00",  # iterations: 3, 300, 70000
        b"\x01\x03\x03" + _f(1e-6, 2e-6, 3e-6),                # norms
        b"\x05a.png" + b"\x00",                                # f_img_path, g_img_path
        b"\x01" + _f(1e-5) + b"\x32\x01" + _f(0.75) + b"\x00\x00\x01\x04" + _f(0.02) + b"\x00\x00\x00",
        b"\x01" + _f(10.0, 30.0) + b"\x06" + _f(*([0.0] * 6)) + _f(0.9),  # seed
        b"\x01\x00",                                           # template_shape: Some(Circle)
        b"\x01\x01\x03\x03\x14\x14\x14",                       # template_sizes
        b"\x00\x00\x00",                                       # time, seams, tracks: None
    ])
    assert pyv.encode_mesh(s) == want


def test_mesh_file_round_trip(tmp_path):
    s = _tiny()
    path = os.path.join(tmp_path, "m.pyv")
    pyv.save_mesh(path, s)
    raw = open(path, "rb").read()
    assert raw[:4] == b"GPYV" and raw[4] == 0x17
    assert pyv.zstd_decompress(raw[5:]) == pyv.encode_mesh(s)
    t = pyv.load_mesh(path)
    for k in ("nodes", "elements", "boundary", "centroids", "areas", "warps", "displacements", "c_zncc", "p",
              "iterations", "norms", "template_sizes"):
        assert np.array_equal(np.asarray(getattr(t, k)), np.asarray(getattr(s, k))), k
    assert (t.seed_node, t.mesh_order, t.subset_order) == (2, 1, 1)
    assert (t.f_img_path, t.g_img_path, t.template_shape, t.time) == ("a.png", "", "circle", None)
    assert t.solve_config["preconditioning"] == "layer-RG" and t.solve_config["max_workers"] is None
    assert t.seed["tolerance"] == 0.9


def test_bad_magic_and_unknown_version(tmp_path):
    p = os.path.join(tmp_path, "x.pyv")
    open(p, "wb").write(b"NOPE\x17abc")
    with pytest.raises(pyv.PyvError, match="magic"):
        pyv.read_pyv(p)
    open(p, "wb").write(b"GPYV\x05abc")
    with pytest.raises(pyv.PyvError, match="unsupported"):
        pyv.read_pyv(p)


def test_varint_forms():
    w = pyv.Writer()
    for v in (0, 250, 251, 65535, 65536, 2 ** 32 - 1, 2 ** 32):
        w.varint(v)
    b = w.bytes()
    assert b == (b"\x00\xfa\xfb\xfb\x00\xfb\xff\xff\xfc\x00\x00\x01\x00\xfc\xff\xff\xff\xff"
                 b"\xfd\x00\x00\x00\x00\x01\x00\x00\x00")
    r = pyv.Reader(b)
    assert [r.varint() for _ in range(7)] == [0, 250, 251, 65535, 65536, 2 ** 32 - 1, 2 ** 32]


def test_boundary_nodes_of_structured_mesh():
    import geopyv_dev as gp
    nodes, el = gp.structured_mesh(0.0, 0.0, 40.0, 30.0, 10.0)
    b = pyv.boundary_nodes(el, 1)
    on_edge = np.flatnonzero((nodes[:, 0] == nodes[:, 0].min()) | (nodes[:, 0] == nodes[:, 0].max()) |
                             (nodes[:, 1] == nodes[:, 1].min()) | (nodes[:, 1] == nodes[:, 1].max()))
    assert sorted(b.tolist()) == sorted(on_edge.tolist())
    # consecutive boundary nodes are joined by an element edge
    edges = {frozenset((int(e[i]), int(e[(i + 1) % 3]))) for e in el for i in range(3)}
    assert all(frozenset((int(b[k]), int(b[(k + 1) % len(b)]))) in edges for k in range(len(b)))

"""GPU tests of the mesh sequence loop (gp.Sequence, sequence.rs:624-968):
which pairs are solved and when the reference moves (known schedules of
the reference's loop), every pair against the oracle's mesh solve from the
same seed, the guide step against the oracle's warp_increment, and frames
read from files through the device loader.

Bars: each pair's node displacements within 1e-4 px of the oracle's mesh
solve (FP64 drop-in) with equal quality flags; the recorded seed and warp of
each pair equal to the oracle's guide step from the previous pair to 1e-9.
"""

import os

import numpy as np
import pytest

import oracle as O
import geopyv_dev as gp
from geopyv_dev import workloads

pytestmark = pytest.mark.gpu

RADIUS = 12


def _mesh():
    return gp.structured_mesh(50.0, 50.0, 206.0, 206.0, 13.0)


def _mixtures(alphas, shifts=None):
    """Frames (1-a) A + a B of two independent speckle patterns, optionally
    translated (x shift per frame): the correlation with frame 0 falls as a
    grows while neighbouring frames stay alike."""
    a = workloads.Speckle((256, 256), 2500, 2.0, 21)
    b = workloads.Speckle((256, 256), 2500, 2.0, 22)
    frames = []
    for k, al in enumerate(alphas):
        if shifts is None or shifts[k] == 0.0:
            fa, fb = a.reference(0.0).astype(np.float64), b.reference(0.0).astype(np.float64)
        else:
            tr = workloads.Translation(shifts[k], 0.0)
            fa, fb = a.target(tr, 0.0).astype(np.float64), b.target(tr, 0.0).astype(np.float64)
        frames.append(np.rint((1.0 - al) * fa + al * fb))
    return frames


def _oracle_pair(f, g, nodes, el, seed, warp, **kw):
    xy, _ = O.mask_local("circle", RADIUS)
    st, r = O.mesh_solve(O.get_c(f, 20), O.get_c(g, 20), f.shape, 20, nodes, el, np.zeros(len(nodes), np.int32),
                         xy, np.array([0, len(xy)]), seed, warp, preconditioning="RG", **kw)
    return st, r


def _oracle_guide(sol, seed):
    """sequence.rs:1062-1097 with the oracle's element_locator and
    warp_increment (mesh_order 1, subset_order 1)."""
    cen = (sol.nodes[sol.elements[:, 0]] + sol.nodes[sol.elements[:, 1]] + sol.nodes[sol.elements[:, 2]]) / 3.0
    i, _ = O.element_locator(seed, sol.nodes, sol.elements, cen)
    row = sol.elements[i]
    inc = O.warp_increment(seed, sol.nodes[row], sol.displacements[row], 1)
    inc[2:] = np.clip(inc[2:], -0.99, 0.99)
    return seed + inc[:2], inc[:6]


def test_guided_sequence_pairs_match_oracle():
    frames = _mixtures([0.0, 0.1, 0.2, 0.3], shifts=[0.0, 0.4, 0.8, 1.2])
    nodes, el = _mesh()
    seq = gp.Sequence(frames=frames, nodes=nodes, elements=el)
    seq.solve(gp.Mask("local", shape="circle", size=RADIUS), seed_coord=[128.0, 128.0], seed_warp=np.zeros(6),
              tolerance=0.75, seed_tolerance=0.75, max_norm=1e-5)
    sols = seq.mesh_solutions
    assert [(s.f_index, s.g_index) for s in sols] == [(0, 1), (0, 2), (0, 3)]
    assert seq.all_converged and not seq.unsolvable and seq.reference_updates == [False] * 3
    seed = np.array([128.0, 128.0])
    warp = np.zeros(6)
    for k, s in enumerate(sols):
        print(f"pair {s.f_index}->{s.g_index}", flush=True)
        np.testing.assert_allclose(s.seed_coord, seed, rtol=0, atol=1e-9)
        np.testing.assert_allclose(s.seed_warp, warp, rtol=0, atol=1e-9)
        st, r = _oracle_pair(frames[s.f_index], frames[s.g_index], nodes, el, seed, warp, tolerance=0.75,
                             seed_tolerance=0.75)
        assert st == 0
        np.testing.assert_array_equal(s.quality_ok, r["quality_ok"])
        assert np.abs(s.p - r["p"]).max() <= 1e-4
        # the frames move 0.4 px per frame (their mixing leaves a few
        # poorly textured nodes further off, in the oracle as here)
        assert abs(np.median(s.displacements[:, 0]) - 0.4 * (k + 1)) < 0.02
        seed, warp = _oracle_guide(s, seed)


def test_sequential_reference_updates():
    frames = _mixtures([0.0, 0.1, 0.2, 0.3])
    nodes, el = _mesh()
    seq = gp.Sequence(frames=frames, nodes=nodes, elements=el)
    seq.solve(gp.Mask("local", shape="circle", size=RADIUS), seed_coord=[128.0, 128.0], tolerance=0.75,
              seed_tolerance=0.75, sequential=True)
    assert [(s.f_index, s.g_index) for s in seq.mesh_solutions] == [(0, 1), (1, 2), (2, 3)]
    assert seq.reference_updates == [False, True, True]


def test_failed_pair_falls_back_to_the_previous_frame():
    # pair (0, 3) leaves 34 nodes below C_ZNCC 0.75 (the gate fails); pair
    # (2, 3) keeps every node above 0.85
    frames = _mixtures([0.0, 0.1, 0.2, 0.4])
    nodes, el = _mesh()
    seq = gp.Sequence(frames=frames, nodes=nodes, elements=el)
    seq.solve(gp.Mask("local", shape="circle", size=RADIUS), seed_coord=[128.0, 128.0], tolerance=0.75,
              seed_tolerance=0.75, guide=False)
    assert [(s.f_index, s.g_index) for s in seq.mesh_solutions] == [(0, 1), (0, 2), (2, 3)]
    assert seq.reference_updates == [False, False, True]
    assert seq.all_converged and not seq.unsolvable
    st, r = _oracle_pair(frames[0], frames[3], nodes, el, [128.0, 128.0], np.zeros(6), tolerance=0.75,
                         seed_tolerance=0.75)
    assert st == 3  # the reference's gate rejects the pair the loop fell back from


def test_unsolvable_consecutive_pair_curtails():
    frames = _mixtures([0.0, 0.2, 0.95])
    nodes, el = _mesh()
    seq = gp.Sequence(frames=frames, nodes=nodes, elements=el)
    seq.solve(gp.Mask("local", shape="circle", size=RADIUS), seed_coord=[128.0, 128.0], tolerance=0.75,
              seed_tolerance=0.75)
    assert [(s.f_index, s.g_index) for s in seq.mesh_solutions] == [(0, 1)]
    assert seq.unsolvable and not seq.all_converged


def test_sequence_from_files_through_the_device_loader(tmp_path):
    from PIL import Image as PILImage
    frames = _mixtures([0.0, 0.1, 0.2], shifts=[0.0, 0.5, 1.0])
    for k, name in enumerate(["f_1.png", "f_2.png", "f_10.png"]):
        PILImage.fromarray(frames[k].astype(np.uint8)).save(os.path.join(tmp_path, name))
    nodes, el = _mesh()
    seq = gp.Sequence(image_dir=str(tmp_path), nodes=nodes, elements=el)
    seq.solve(gp.Mask("local", shape="circle", size=RADIUS), seed_coord=[128.0, 128.0], tolerance=0.75,
              seed_tolerance=0.75, guide=False)
    loaded = [O.load_convert(fr.astype(np.uint8)) for fr in frames]
    for s in seq.mesh_solutions:
        st, r = _oracle_pair(loaded[s.f_index], loaded[s.g_index], nodes, el, [128.0, 128.0], np.zeros(6),
                             tolerance=0.75, seed_tolerance=0.75)
        assert st == 0
        assert np.abs(s.p - r["p"]).max() <= 1e-4


class _PairCounter:
    """sequence.rs:1695-1712: records pair_start calls and node_done
    counts; cancels once the pair cancel_at has started."""

    def __init__(self, cancel_at=None):
        self.pairs, self.labels, self.nodes, self.cancel_at = [], [], 0, cancel_at

    def node_done(self):
        self.nodes += 1

    def pair_start(self, pair, n_pairs, label):
        self.pairs.append((pair, n_pairs))
        self.labels.append(label)

    def cancelled(self):
        return self.cancel_at is not None and len(self.pairs) > self.cancel_at


def test_sequence_reports_pairs_and_cancels():
    # sequence.rs:1714-1738: pairs (0, 2), (1, 2) reported with node counts;
    # cancelled as pair 1 starts: RuntimeError("solve cancelled"), not the
    # failed-pair fallback, and no solution stored
    frames = _mixtures([0.0, 0.1, 0.2])
    nodes, el = _mesh()
    mask = gp.Mask("local", shape="circle", size=RADIUS)
    seq = gp.Sequence(frames=frames, nodes=nodes, elements=el)
    p = _PairCounter()
    seq.solve(mask, seed_coord=[128.0, 128.0], tolerance=0.75, seed_tolerance=0.75, progress=p)
    assert p.pairs == [(0, 2), (1, 2)] and p.labels == ["0→1", "0→2"]
    assert p.nodes >= 2 * len(nodes) and len(seq.mesh_solutions) == 2
    seq = gp.Sequence(frames=frames, nodes=nodes, elements=el)
    p = _PairCounter(cancel_at=1)
    with pytest.raises(RuntimeError, match="solve cancelled"):
        seq.solve(mask, seed_coord=[128.0, 128.0], tolerance=0.75, seed_tolerance=0.75, progress=p)
    assert len(p.pairs) == 2 and not seq.solved

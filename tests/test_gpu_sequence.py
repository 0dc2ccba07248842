"""Incremental sequence on the GPU (sequence.rs:624-968 in batch form): the
frame loop replayed from one captured HIP graph equals the eager frame loop
bit for bit on every frame (several replays, with eager reference updates
and synchronisations between them), and both agree with the FP64 oracle
run through the same loop: converged displacements within 1e-4 px."""

import numpy as np
import pytest

import oracle as O
import geopyv_dev as gp
from geopyv_dev import _native, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def seq():
    frames, fields = workloads.sequence_frames(seed=3, shape=(320, 320), n_frames=7, n_speckles=2000, slip_max=1.2)
    coords = workloads.grid_coords((320, 320), 30, 40)
    mask = gp.Mask(mask_type="local", shape="circle", size=15)
    return frames, coords, mask


@pytest.mark.parametrize("prefilter", [_native.PREFILTER_FIR, _native.PREFILTER_FFT], ids=["fir", "fft"])
def test_graph_replay_equals_eager(seq, prefilter):
    frames, coords, mask = seq
    e = gp._engine()
    e.set_prefilter(prefilter)
    try:
        rg = gp.solve_sequence(frames, coords, mask, reference_every=3, use_graph=True)
        re = gp.solve_sequence(frames, coords, mask, reference_every=3, use_graph=False)
    finally:
        e.set_prefilter(_native.PREFILTER_FIR)
    assert len(rg) == len(re) == len(frames) - 1
    for k, (a, b) in enumerate(zip(rg, re)):
        assert np.array_equal(a["p"], b["p"]), f"frame {k + 1}"
        assert np.array_equal(a["iterations"], b["iterations"])
        assert np.array_equal(a["c_zncc"], b["c_zncc"])


def test_sequence_matches_oracle(seq):
    frames, coords, mask = seq
    res = gp.solve_sequence(frames, coords, mask, reference_every=3, use_graph=True)
    n = len(coords)
    xy = mask.coords
    ref = frames[0].astype(np.float64)
    Cf = O.get_c(ref, 20)
    p0 = np.zeros((n, 12))
    for k in range(1, len(frames)):
        Cg = O.get_c(frames[k].astype(np.float64), 20)
        orc = O.batch_solve(Cf, Cg, ref.shape, 20, coords, np.zeros(n, np.int32), xy, np.array([0, len(xy)]),
                            np.full(n, 1, np.int32), p0, max_norm=1e-5, max_iterations=50)
        r = res[k - 1]
        both = (r["converged"] == 1) & (orc["converged"] == 1)
        assert np.mean(r["converged"] == orc["converged"]) >= 0.95, f"frame {k}"
        du = np.abs(r["p"][both, :2] - orc["p"][both, :2]).max()
        print("frame", k, "max |du|", du, "converged", both.mean())
        assert du < 1e-4
        p0 = orc["p"].copy()
        if k % 3 == 0:
            Cf = Cg
            p0 = np.zeros((n, 12))


def test_graph_replay_equals_eager_uneven(seq):
    # two size classes (1 and 2 wavefronts per subset): the graph holds the
    # forked launches on the auxiliary streams
    frames, coords, _ = seq
    masks = [gp.Mask(mask_type="local", shape="circle", size=12 if i % 2 else 20) for i in range(len(coords))]
    rg = gp.solve_sequence(frames[:5], coords, masks, reference_every=2, use_graph=True)
    re = gp.solve_sequence(frames[:5], coords, masks, reference_every=2, use_graph=False)
    for k, (a, b) in enumerate(zip(rg, re)):
        assert np.array_equal(a["p"], b["p"]), f"frame {k + 1}"
        assert np.array_equal(a["iterations"], b["iterations"])

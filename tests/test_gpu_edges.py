"""GPU parity at the image's edges and corners, where the reference clamps
every evaluated position into the image (image.rs:130-132) and the batched
engine leaves its fast path: clamped cells, coefficient windows cut by the
image boundary, and gathers from HBM when the warped subset leaves its
window.  Orders 0, 1 and 2; subsets whose centres lie on, near and just
inside the border, from a zero warp and from a warp 1.3 px off.

Bars: the FP64 drop-in gives the oracle's status and convergence flag at
every subset and its displacement within 1e-4 px wherever the oracle
converges.  The FP32 engine gives the oracle's status wherever the oracle
converges, its displacement within 1e-4 px wherever both converge, and the
oracle's convergence flag at 97 % of the subsets.  Where the oracle does
not converge, the solve diverges (warps of 1e4 px and more, every position
clamped): the FP32 trajectory parts from the FP64 one and may reach a
singular warp composition or Hessian, which the reference would raise
(subset.rs:520, 595, 1143), so a singular status is accepted there.
"""

from geopyv_dev import _native

import numpy as np
import pytest

import oracle as O
import geopyv_dev as gp
from geopyv_dev import workloads

pytestmark = pytest.mark.gpu

RADIUS = 12
H, W = 220, 260


@pytest.fixture(scope="module")
def edge_pair():
    sp = workloads.Speckle((H, W), 900, 2.2, 17)
    field = workloads.Translation(0.37, -0.52)
    ref = sp.reference(2.0).astype(np.float64)
    tar = sp.target(field, 2.0).astype(np.float64)
    xs = [0.0, 0.4, 5.0, 11.5, 12.0, 30.0, W / 2.0, W - 31.0, W - 12.0, W - 6.3, W - 1.0]
    ys = [0.0, 3.7, 12.0, 45.0, H / 2.0, H - 12.5, H - 2.0, H - 1.0]
    coords = np.array([[x, y] for y in ys for x in xs], dtype=np.float64)
    return dict(ref=ref, tar=tar, coords=coords, Cf=O.get_c(ref, 20), Cg=O.get_c(tar, 20),
                f=gp.Image(image_gs=ref, border=20), g=gp.Image(image_gs=tar, border=20))


def _p0(n, offset):
    p0 = np.zeros((n, 12))
    if offset:
        p0[:, 0] = 1.3
        p0[:, 1] = -1.3
    return p0


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("offset", [False, True])
def test_edge_subsets_match_oracle(edge_pair, order, offset):
    c = edge_pair
    coords = c["coords"]
    n = len(coords)
    m = gp.Mask(mask_type="local", shape="circle", size=RADIUS)
    p0 = _p0(n, offset)
    print(f"edges: order {order}, offset {offset}, {n} subsets", flush=True)
    orc = O.batch_solve(c["Cf"], c["Cg"], (H, W), 20, coords, np.zeros(n, np.int32), m.coords,
                        np.array([0, len(m.coords)]), np.full(n, order, np.int32), p0, max_norm=1e-5,
                        max_iterations=50)
    f64 = gp.solve_subsets(c["f"], c["g"], coords, m, p0=p0, subset_order=order, precision="f64", max_norm=1e-5)
    f32 = gp.solve_subsets(c["f"], c["g"], coords, m, p0=p0, subset_order=order, precision="f32", max_norm=1e-5)
    np.testing.assert_array_equal(f64["status"], orc["status"])
    assert np.array_equal(f64["converged"], orc["converged"])
    ok = (orc["status"] == 0) & (orc["converged"] == 1)
    np.testing.assert_array_equal(f32["status"][ok], orc["status"][ok])
    singular = (_native.ERR_SINGULAR_HESSIAN, _native.ERR_SINGULAR_WARP)
    diverged = ~ok & (f32["status"] != orc["status"])
    assert all(st in singular for st in f32["status"][diverged])
    du64 = np.abs(f64["p"][ok, :2] - orc["p"][ok, :2]).max()
    both = ok & (f32["converged"] == 1)
    du32 = np.abs(f32["p"][both, :2] - orc["p"][both, :2]).max() if both.any() else 0.0
    same = float(np.mean(f32["converged"] == orc["converged"]))
    print(f"  f64 max |du| {du64:.2e}; f32 max |du| {du32:.2e} over {both.sum()} converged, "
          f"convergence flags equal {same:.3f}, oracle converged {orc['converged'].mean():.3f}, "
          f"f32 singular on {int(diverged.sum())} diverging subsets", flush=True)
    assert du64 < 1e-4
    assert du32 < 1e-4
    assert same >= 0.97

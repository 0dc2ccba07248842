"""The batched-engine kernel source run on the CPU through the SIMT
emulation (tests/simt), under AddressSanitizer and UBSan, against the FP64
oracle.  Catches out-of-bounds indexing, cross-lane operations reached by
part of a wavefront, and logic errors without a GPU."""

import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "simt")


@pytest.fixture(scope="module")
def driver():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return os.path.join(HERE, "simt_driver")


@pytest.mark.parametrize("order,radius,slack", [(0, 10, 3), (1, 10, 3), (2, 10, 3), (1, 12, 0), (1, 7, 40)])
def test_batch_kernel_emulated_matches_oracle(driver, order, radius, slack):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([driver, str(order), str(radius), "2", "50", str(slack)], capture_output=True, text=True,
                       timeout=300, env=env)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "mismatched 0" in r.stdout

"""The seeded input generator (theses_code_amd/rng.py) against an
independent C statement of splitmix64-seeded xoshiro256** lanes
(tests/cpp/xoshiro_check.c), and the distributions derived from it."""
import os
import subprocess

import numpy as np
import pytest

from theses_code_amd import rng as R

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cgen(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("xo") / "xoshiro_check")
    subprocess.run(["gcc", "-O2", os.path.join(REPO, "tests", "cpp", "xoshiro_check.c"), "-o", out],
                   check=True)
    return out


@pytest.mark.parametrize("seed", [0, 0x5EED0000, 0x5EED0003, 2 ** 64 - 1])
def test_stream_matches_c(cgen, seed):
    n = 3 * R.LANES + 17  # several rounds and a partial one
    r = subprocess.run([cgen, str(seed), str(n)], capture_output=True, text=True, check=True)
    want = np.array([int(v) for v in r.stdout.split()], dtype=np.uint64)
    g = R.Xoshiro(seed)
    # drawn in pieces of uneven size: the stream does not depend on them
    got = np.concatenate([g.next_u64(5), g.next_u64(R.LANES), g.next_u64(n - 5 - R.LANES)])
    np.testing.assert_array_equal(got, want)


def test_splitmix64_known_values():
    # splitmix64 from state 0: the first output is the mix of the golden
    # gamma, computed here with Python integers
    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) % (1 << 64)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) % (1 << 64)
        return z ^ (z >> 31)
    got = R.splitmix64(12345, 3)
    want = [mix((12345 + k * 0x9E3779B97F4A7C15) % (1 << 64)) for k in (1, 2, 3)]
    assert [int(v) for v in got] == want


def test_distributions():
    g = R.Xoshiro(0x5EED0001)
    u = g.uniform(0.1, 0.4, 200000)
    assert u.min() >= 0.1 and u.max() < 0.4
    assert abs(u.mean() - 0.25) < 1e-3
    z = g.normal(size=200001)
    assert abs(z.mean()) < 1e-2 and abs(z.std() - 1.0) < 1e-2
    k = g.integers(64, 1025, 100000)
    assert k.min() == 64 and k.max() == 1024
    c = g.choice(1000, 100)
    assert len(set(c.tolist())) == 100 and c.max() < 1000


def test_spawn_children_are_distinct_and_reproducible():
    a = [c.next_u64(4) for c in R.Xoshiro(7).spawn(3)]
    b = [c.next_u64(4) for c in R.Xoshiro(7).spawn(3)]
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert not np.array_equal(a[0], a[1])

"""Seeded synthetic inputs: splitmix64-seeded xoshiro256** (Blackman and
Vigna), vectorised over independent lanes, so that any C program can
reproduce the same numbers.

The stream of a generator with seed S:

* splitmix64 from state S gives outputs z_0, z_1, ... (x += 0x9E3779B97F4A7C15;
  z = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9; z = (z ^ (z >> 27)) *
  0x94D049BB133111EB; z ^ (z >> 31));
* lane i (i < LANES = 1024) is a xoshiro256** generator with state
  (z_{4i}, z_{4i+1}, z_{4i+2}, z_{4i+3});
* the stream's v-th 64-bit output is output v // LANES of lane v % LANES,
  so each round of LANES outputs takes one step of every lane.

Derived values consume the stream in order: a double in [0, 1) is
(u >> 11) * 2^-53 of one output; uniform(lo, hi) is lo + (hi - lo) * that;
an integer in [lo, hi) is lo + floor(double * (hi - lo)); normal deviates
come in Box-Muller pairs from two doubles d1, d2: r = sqrt(-2 ln(1 - d1)),
r cos(2 pi d2), r sin(2 pi d2) (an odd count drops the last sine).  Child
generators (spawn) take seeds S_q = z_q of splitmix64 from state S ^
0xD1B54A32D192ED03, q = 0, 1, ..."""
from __future__ import annotations

import numpy as np

LANES = 1024
_GOLD = 0x9E3779B97F4A7C15
_SPAWN_KEY = 0xD1B54A32D192ED03


def splitmix64(state: int, count: int) -> np.ndarray:
    """The first `count` outputs of splitmix64 from `state`, as uint64."""
    with np.errstate(over="ignore"):
        # uint64 arithmetic wraps modulo 2^64, as the C statement does
        z = np.uint64(int(state) % (1 << 64)) + np.uint64(_GOLD) * np.arange(1, count + 1,
                                                                          dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _rotl(x: np.ndarray, k: int) -> np.ndarray:
    return (x << np.uint64(k)) | (x >> np.uint64(64 - k))


class Xoshiro:
    """The generator of one seed (see the module text for the stream)."""

    def __init__(self, seed: int):
        self.seed = int(seed) % (1 << 64)
        z = splitmix64(self.seed, 4 * LANES).reshape(LANES, 4)
        self._s = [z[:, 0].copy(), z[:, 1].copy(), z[:, 2].copy(), z[:, 3].copy()]
        self._buf = np.empty(0, dtype=np.uint64)

    def _round(self) -> np.ndarray:
        s0, s1, s2, s3 = self._s
        with np.errstate(over="ignore"):
            out = _rotl(s1 * np.uint64(5), 7) * np.uint64(9)
        t = s1 << np.uint64(17)
        s2 ^= s0
        s3 ^= s1
        s1 ^= s2
        s0 ^= s3
        s2 ^= t
        self._s[3] = _rotl(s3, 45)
        return out

    def next_u64(self, n: int) -> np.ndarray:
        """The stream's next n 64-bit outputs."""
        n = int(n)
        parts, have = [self._buf], self._buf.size
        while have < n:
            r = self._round()
            parts.append(r)
            have += r.size
        allv = np.concatenate(parts) if len(parts) > 1 else self._buf
        out, self._buf = allv[:n], allv[n:]
        return out

    def random(self, size=None):
        n = int(np.prod(size)) if size is not None else 1
        d = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
        return d.reshape(size) if size is not None else float(d[0])

    def uniform(self, low=0.0, high=1.0, size=None):
        d = self.random(size)
        return low + (high - low) * d

    def integers(self, low: int, high: int | None = None, size=None):
        if high is None:
            low, high = 0, low
        d = self.random(size)
        v = low + np.floor(np.asarray(d) * (high - low)).astype(np.int64)
        return v if size is not None else int(v)

    def normal(self, loc=0.0, scale=1.0, size=None):
        n = int(np.prod(size)) if size is not None else 1
        m = (n + 1) // 2
        d = self.random(2 * m).reshape(m, 2)
        r = np.sqrt(-2.0 * np.log1p(-d[:, 0]))
        z = np.empty((m, 2))
        z[:, 0] = r * np.cos(2.0 * np.pi * d[:, 1])
        z[:, 1] = r * np.sin(2.0 * np.pi * d[:, 1])
        v = loc + scale * z.reshape(-1)[:n]
        return v.reshape(size) if size is not None else float(v[0])

    def choice(self, n: int, k: int) -> np.ndarray:
        """k distinct integers of [0, n) (a partial Fisher-Yates shuffle)."""
        perm = np.arange(n)
        for i in range(k):
            j = i + int(self.integers(0, n - i))
            perm[i], perm[j] = perm[j], perm[i]
        return perm[:k]

    def spawn(self, count: int) -> list["Xoshiro"]:
        seeds = splitmix64(self.seed ^ _SPAWN_KEY, count)
        return [Xoshiro(int(s)) for s in seeds]

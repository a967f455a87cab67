"""The per-position rule by which the GPU index build (ga_index_gpu.hip,
idxWindowKernel) decides the minimizers, restated position by position in
Python, against the oracle's sequential window walk (iterateMinimizersReal,
src/MinimizerSeeder.cpp:104-189) on random, low-complexity and N-broken strings:
the same positions, and in increasing order (the walk's discovery order)."""
import random

import pytest

import oracle_ffi as O
from test_oracle_invariants import _hash

CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


def gpu_rule(s, k, w):
    n = len(s)
    code = [CODE.get(c, 4) for c in s]
    valid, h, km = [False] * n, [0] * n, [0] * n
    for e in range(n):
        if e + 1 >= k and all(code[i] < 4 for i in range(e + 1 - k, e + 1)):
            v = 0
            for i in range(e + 1 - k, e + 1):
                v = (v << 2) | code[i]
            valid[e], km[e], h[e] = True, v, _hash(v)
    r = w - k + 1
    own, later = [False] * n, [False] * n
    for e in range(n):
        if not valid[e]:
            continue
        back = e - (w + 2) if e >= w + 2 else 0
        s0 = -1
        for i in range(e, back - 1, -1):
            if code[i] >= 4:
                s0 = i + 1
                break
        if s0 < 0 and e < w + 2:
            s0 = 0
        if s0 >= 0 and e <= s0 + w:
            end = s0 + w
            if end >= n or any(code[i] >= 4 for i in range(e + 1, end + 1)):
                continue
            own[e] = h[e] == min(h[s0 + k - 1: end + 1])
            continue
        m = min(h[e - r + 1: e + 1])
        own[e] = h[e] == m
        prev = min(h[s0 + k - 1: s0 + w + 1]) if (s0 >= 0 and e == s0 + w + 1) else min(h[e - r: e])
        if m > prev:
            for j in range(e - 1, e - r, -1):
                if h[j] == m:
                    later[j] = True
                    break
    return [(e, km[e]) for e in range(n) if own[e] or later[e]]


@pytest.mark.parametrize("k,w", [(15, 20), (19, 30), (5, 9), (3, 3), (4, 4), (3, 12)])
def test_gpu_rule_equals_window_walk(k, w):
    rng = random.Random(7 * k + w)
    strings = ["".join(rng.choice("ACGT") for _ in range(rng.randrange(1, 300))) for _ in range(120)]
    strings += ["".join(rng.choice("AC") for _ in range(rng.randrange(w, 250))) for _ in range(60)]
    strings += ["".join(rng.choice("ACGTN" if rng.random() < 0.5 else "ACGT") for _ in range(rng.randrange(k, 200))) for _ in range(80)]
    strings += ["A" * n for n in (k - 1, k, w, w + 1, w + 2, 3 * w)] + ["N" * 5 + "A" * (w + 1), "A" * (w + 1) + "N" + "C" * (w + 1)]
    for s in strings:
        want = [tuple(x) for x in O.minimizers(s, k, w).tolist()]
        assert gpu_rule(s, k, w) == want, (s, k, w)

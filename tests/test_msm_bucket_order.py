"""The MSM's bucket order (csrc/msm.hip, k_msm_entries and the radix sort in
MsmEngine::enqueue), checked on the CPU.

The entries are written window-major with the global bucket id w * B + b
(B = 2^(c-1); the top window's unsigned digits use 2B ids) in the high word,
and sorted stably on the c low bits of that id only.  The accumulation needs
every bucket's entries to be contiguous after the sort, which holds when the
c low bits of the ids of one window are all distinct.  This test builds
entries the same way for c = 2..20 (zero digits as skip entries in bucket
term mod B of their window), sorts them stably on the c low bits and checks
that no bucket id reappears after its run ends and that no entry is lost.
"""
import random

import pytest


def _entries(c, n2, zero_frac, rnd):
    W = (128 + c - 1) // c
    B = 1 << (c - 1)
    out = []
    for w in range(W):
        for i in range(n2):
            top = w == W - 1
            if rnd.random() < zero_frac:
                b, skip = i & (B - 1), True
            else:
                b, skip = rnd.randrange(2 * B if top else B), False
            out.append((w * B + b, i, skip))
    return out


@pytest.mark.parametrize("c", [2, 3, 5, 8, 10, 13, 15, 16, 17, 20])
def test_low_c_bits_keep_buckets_contiguous(c):
    rnd = random.Random(1000 + c)
    ents = _entries(c, 200 if c < 14 else 2500, 0.05, rnd)
    mask = (1 << c) - 1
    srt = sorted(ents, key=lambda e: e[0] & mask)          # Python's sort is stable
    seen, prev = set(), None
    for k, _, _ in srt:
        if k != prev:
            assert k not in seen, f"bucket {k} split at c = {c}"
            seen.add(k)
            prev = k
    assert sorted(ents) == sorted(srt)


def test_window_ids_distinct_in_low_bits():
    # exhaustive over every (window, digit) id for the widths the engine allows
    for c in range(2, 21):
        W = (128 + c - 1) // c
        B = 1 << (c - 1)
        for w in range(W):
            n = 2 * B if w == W - 1 else B
            low = {(w * B + b) & ((1 << c) - 1) for b in range(n)}
            assert len(low) == n, (c, w)

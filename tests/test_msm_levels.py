"""CPU replay of the MSM bucket-sum stages (csrc/msm.hip: k_seg0, the piece
step k_run_heads + k_piece_reduce over the compacted partials, k_level and
the host loop that runs them).

Integers stand in for points: an entry's value is added to its bucket's sum.
The first stage cuts the bucket-sorted entries into chunks of K0 and writes a
(head, tail) slot pair per chunk.  The piece step compacts the partials,
splits each bucket's run of partials into pieces of at most PL, completes the
buckets whose run is one piece and leaves one partial per piece of the longer
runs, at the piece's index.  Every later level reads the previous slots in
place, empty slots where they are, and relies on the rule that two
consecutive partials of one bucket are at most one empty slot apart.  Every
bucket must be written exactly once and hold the sum of its entries, for
bucket layouts from uniform to one bucket holding everything, and for chunk
sizes down to 2 (entries) and 4 (slots: a level writes two slots per chunk,
so only chunks of four or more slots shrink every slot count above one chunk;
with three, four slots stay four).
"""
import random
import zlib

import pytest

EMPTY = None


class Sums:
    def __init__(self):
        self.v = {}

    def write(self, k, x):
        assert k not in self.v, f"bucket {k} written twice"
        self.v[k] = x


def seg0(keys, vals, skip, K, bsum):
    """k_seg0: chunks of K sorted entries; skip entries add nothing."""
    E = len(keys)
    chunks = (E + K - 1) // K
    okeys, opart = [EMPTY] * (2 * chunks), [0] * (2 * chunks)
    for j in range(chunks):
        beg, end = j * K, min(j * K + K, E)
        before = beg > 0 and keys[beg - 1] == keys[beg]
        after = end < E and keys[end] == keys[end - 1]
        cur, first, acc = keys[beg], True, 0
        hk = tk = EMPTY
        hv = tv = 0
        for e in range(beg, end):
            if keys[e] != cur:
                if first and before:
                    hk, hv = cur, acc
                else:
                    bsum.write(cur, acc)
                first, cur, acc = False, keys[e], 0
            if not skip[e]:
                acc += vals[e]
        if first and before:
            hk, hv = cur, acc
        elif after:
            tk, tv = cur, acc
        else:
            bsum.write(cur, acc)
        okeys[2 * j], opart[2 * j], okeys[2 * j + 1], opart[2 * j + 1] = hk, hv, tk, tv
    return okeys, opart, chunks


def level(ikeys, ipart, n, K, bsum):
    """k_level: chunks of K slots read in place."""
    chunks = (n + K - 1) // K
    okeys, opart = [EMPTY] * (2 * chunks), [0] * (2 * chunks)
    for j in range(chunks):
        beg, end = j * K, min(j * K + K, n)
        nz = [e for e in range(beg, end) if ikeys[e] is not EMPTY]
        hk = tk = EMPTY
        hv = tv = 0
        if nz:
            f, l = nz[0], nz[-1]
            kf, kl = ikeys[f], ikeys[l]
            before = False
            if f >= 1:
                q = f - 1
                if ikeys[q] is EMPTY and q >= 1:
                    q -= 1
                before = ikeys[q] == kf
            after = False
            if l + 1 < n:
                q = l + 1
                if ikeys[q] is EMPTY and q + 1 < n:
                    q += 1
                after = ikeys[q] == kl
            cur, first, acc = kf, True, 0
            for e in range(f, l + 1):
                k = ikeys[e]
                if k is EMPTY:
                    continue
                if k != cur:
                    if first and before:
                        hk, hv = cur, acc
                    else:
                        bsum.write(cur, acc)
                    first, cur, acc = False, k, 0
                acc += ipart[e]
            if first and before:
                hk, hv = cur, acc
            elif after:
                tk, tv = cur, acc
            else:
                bsum.write(cur, acc)
        okeys[2 * j], opart[2 * j], okeys[2 * j + 1], opart[2 * j + 1] = hk, hv, tk, tv
    return okeys, opart, chunks


def pieces(ikeys, ipart, S, PL, bsum):
    """Compaction, run starts and k_piece_reduce: piece p's partial (or an
    empty slot) at index p of S slots."""
    comp = [e for e in range(S) if ikeys[e] is not EMPTY]
    E = len(comp)
    rstart = [0] * E
    for i in range(E):
        rstart[i] = i if i == 0 or ikeys[comp[i]] != ikeys[comp[i - 1]] else rstart[i - 1]
    starts = [i for i in range(E) if (i - rstart[i]) % PL == 0]
    okeys, opart = [EMPTY] * S, [0] * S
    for p, beg in enumerate(starts):
        end = starts[p + 1] if p + 1 < len(starts) else E
        key = ikeys[comp[beg]]
        acc = sum(ipart[comp[j]] for j in range(beg, end))
        whole = rstart[beg] == beg and (end == E or ikeys[comp[end]] != key)
        if whole:
            bsum.write(key, acc)
        else:
            okeys[p], opart[p] = key, acc
    return okeys, opart


def bucket_sums(keys, vals, skip, K0, PL, K):
    bsum = Sums()
    slot_keys, slot_part, chunks0 = seg0(keys, vals, skip, K0, bsum)
    slots = 2 * chunks0
    slot_keys, slot_part = pieces(slot_keys, slot_part, slots, PL, bsum)
    levels = 0
    while True:
        chunks = (slots + K - 1) // K
        slot_keys, slot_part, _ = level(slot_keys, slot_part, slots, K, bsum)
        levels += 1
        if chunks == 1:
            break
        slots = 2 * chunks
    assert all(s is EMPTY for s in slot_keys)      # nothing left over
    return bsum.v, levels


def layout(kind, rng, E, nb):
    if kind == "uniform":
        keys = sorted(rng.randrange(nb) for _ in range(E))
    elif kind == "one_bucket":
        keys = [7] * E
    elif kind == "few_huge":          # a narrow top window: a few buckets hold most entries
        keys = sorted(rng.choice([3, 5, nb - 1]) if rng.random() < 0.8 else rng.randrange(nb) for _ in range(E))
    elif kind == "runs":              # run lengths from 1 to several hundred
        keys, b = [], 0
        while len(keys) < E:
            keys += [b] * rng.choice([1, 1, 2, 3, 17, 31, 32, 33, 64, 257, 900])
            b += rng.randrange(1, 4)
        keys = keys[:E]
    else:
        raise ValueError(kind)
    vals = [rng.randrange(1, 1 << 20) for _ in keys]
    skip = [rng.random() < 0.05 for _ in keys]
    return keys, vals, skip


@pytest.mark.parametrize("kind", ["uniform", "one_bucket", "few_huge", "runs"])
@pytest.mark.parametrize("K0,PL,K", [(32, 16, 16), (2, 2, 4), (3, 4, 5), (5, 3, 4), (32, 1, 16), (4, 8, 4)])
def test_every_bucket_once_with_its_sum(kind, K0, PL, K):
    rng = random.Random(zlib.crc32(f"{kind}/{K0}/{PL}/{K}".encode()))
    for E, nb in ((1, 4), (2, 3), (37, 5), (600, 64), (5000, 300)):
        keys, vals, skip = layout(kind, rng, E, nb)
        got, _ = bucket_sums(keys, vals, skip, K0, PL, K)
        want = {}
        for k, v, s in zip(keys, vals, skip):
            want[k] = want.get(k, 0) + (0 if s else v)
        assert got == want, (kind, E, nb)


def test_level_count_at_the_2p20_shape():
    """2^20 terms at c = 16: 2^24 entries, 2^19 first-stage chunks, 2^20
    slots after the pieces; the levels go 2^20 -> 2^17 -> 2^14 -> 2048 -> 256
    -> 32 -> 4 slots and one chunk, i.e. seven launches of the level kernel."""
    slots, levels = 2 * ((1 << 24) // 32), 0
    while True:
        chunks = (slots + 15) // 16
        levels += 1
        if chunks == 1:
            break
        slots = 2 * chunks
    assert levels == 7


def test_uniform_buckets_finish_in_the_piece_step():
    """About 64 entries per bucket, as at 2^20 terms and c = 16: every bucket
    completes in the piece step, so the levels find only empty slots."""
    rng = random.Random(5)
    E, nb = 60000, 940
    keys = sorted(rng.randrange(nb) for _ in range(E))
    bsum = Sums()
    sk, sp, ch0 = seg0(keys, [1] * E, [False] * E, 32, bsum)
    ok, _ = pieces(sk, sp, 2 * ch0, 16, bsum)
    assert all(k is EMPTY for k in ok)
    assert len(bsum.v) == len(set(keys))

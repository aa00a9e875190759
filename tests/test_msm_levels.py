"""CPU replay of the MSM bucket-sum stages (csrc/msm.hip: k_seg0, k_join,
k_level and the host loop that runs them).

Integers stand in for points: an entry's value is added to its bucket's sum.
The first stage cuts the bucket-sorted entries into chunks of K0 and writes a
(head, tail) slot pair per chunk; a chunk that lies inside one bucket's run
marks its tail slot CONT.  The join step lets the lane of a bucket's first
chunk add the bucket's tail partial and the heads of the chunks it reaches
when they are at most SPAN chunks away, write the bucket and empty those
slots.  Every later level reads the previous slots in place, empty (and CONT)
slots where they are, and relies on the rule that two consecutive partials of
one bucket are at most one empty slot apart.  Every bucket must be written
exactly once and hold the sum of its entries, for bucket layouts from uniform
to one bucket holding everything, and for chunk sizes down to 2 (entries) and
4 (slots: a level writes two slots per chunk, so only chunks of four or more
slots shrink every slot count above one chunk; with three, four slots stay
four).
"""
import random
import zlib

import pytest

EMPTY = None
CONT = "cont"      # the kernel's kCont: an empty tail slot of a chunk inside one run


def is_empty(k):
    return k is EMPTY or k == CONT


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
            if after:
                tk = CONT
        elif after:
            tk, tv = cur, acc
        else:
            bsum.write(cur, acc)
        okeys[2 * j], opart[2 * j], okeys[2 * j + 1], opart[2 * j + 1] = hk, hv, tk, tv
    return okeys, opart, chunks


def join(keys, part, chunks, span, bsum):
    """k_join, in place.  The lanes run in any order on the device; each reads
    only its own chain's slots (and the tail key of the chain's last chunk,
    which that chunk's lane may empty: both values read as 'not CONT'), so a
    random lane order here stands for any interleaving."""
    order = list(range(chunks))
    random.Random(chunks).shuffle(order)
    snapshot_tail = [keys[2 * j + 1] for j in range(chunks)]
    for j in order:
        tk = keys[2 * j + 1]
        if is_empty(tk):
            continue
        m, ok = 1, True
        while m <= span and j + m < chunks:
            if keys[2 * (j + m)] != tk:
                ok = False
                break
            # the last chunk's tail may already be emptied by its own lane
            t = keys[2 * (j + m) + 1]
            assert (t == CONT) == (snapshot_tail[j + m] == CONT)
            if t != CONT:
                break
            m += 1
        if not ok or m > span or j + m >= chunks:
            continue
        bsum.write(tk, part[2 * j + 1] + sum(part[2 * (j + i)] for i in range(1, m + 1)))
        keys[2 * j + 1] = EMPTY
        for i in range(1, m + 1):
            assert keys[2 * (j + i)] == tk      # each head joins one chain only
            keys[2 * (j + i)] = EMPTY
    return keys, part


def level(ikeys, ipart, n, K, bsum):
    """k_level: chunks of K slots read in place."""
    chunks = (n + K - 1) // K
    okeys, opart = [EMPTY] * (2 * chunks), [0] * (2 * chunks)
    for j in range(chunks):
        beg, end = j * K, min(j * K + K, n)
        nz = [e for e in range(beg, end) if not is_empty(ikeys[e])]
        hk = tk = EMPTY
        hv = tv = 0
        if nz:
            f, l = nz[0], nz[-1]
            kf, kl = ikeys[f], ikeys[l]
            before = False
            if f >= 1:
                q = f - 1
                if is_empty(ikeys[q]) and q >= 1:
                    q -= 1
                before = ikeys[q] == kf
            after = False
            if l + 1 < n:
                q = l + 1
                if is_empty(ikeys[q]) and q + 1 < n:
                    q += 1
                after = ikeys[q] == kl
            cur, first, acc = kf, True, 0
            for e in range(f, l + 1):
                k = ikeys[e]
                if is_empty(k):
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


def bucket_sums(keys, vals, skip, K0, SPAN, K):
    bsum = Sums()
    slot_keys, slot_part, chunks0 = seg0(keys, vals, skip, K0, bsum)
    slots = 2 * chunks0
    # the two slot buffers as the host sizes them: the first stage's slots,
    # and two per chunk of the first level; levels alternate between them
    cap = [slots, 2 * ((slots + K - 1) // K)]
    slot_keys, slot_part = join(slot_keys, slot_part, chunks0, SPAN, bsum)
    levels = 0
    while True:
        chunks = (slots + K - 1) // K
        slot_keys, slot_part, _ = level(slot_keys, slot_part, slots, K, bsum)
        assert 2 * chunks <= cap[(levels + 1) % 2], "a level's output overflows its buffer"
        levels += 1
        if chunks == 1:
            break
        slots = 2 * chunks
    assert all(is_empty(s) for s in slot_keys)      # nothing left over
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
@pytest.mark.parametrize("K0,SPAN,K", [(32, 8, 16), (2, 2, 4), (3, 4, 5), (5, 3, 4), (32, 1, 16), (4, 8, 4),
                                       (32, 0, 16)])
def test_every_bucket_once_with_its_sum(kind, K0, SPAN, K):
    rng = random.Random(zlib.crc32(f"{kind}/{K0}/{SPAN}/{K}".encode()))
    for E, nb in ((1, 4), (2, 3), (37, 5), (600, 64), (5000, 300)):
        keys, vals, skip = layout(kind, rng, E, nb)
        got, _ = bucket_sums(keys, vals, skip, K0, SPAN, K)
        want = {}
        for k, v, s in zip(keys, vals, skip):
            want[k] = want.get(k, 0) + (0 if s else v)
        assert got == want, (kind, E, nb)


def test_level_count_at_the_2p20_shape():
    """2^20 terms at c = 16, one window group of 8 windows: 2^24 entries, 2^19
    first-stage chunks, 2^20 slots after the join; the levels go 2^20 -> 2^17
    -> 2^14 -> 2048 -> 256 -> 32 -> 4 slots and one chunk, i.e. seven launches
    of the level kernel."""
    slots, levels = 2 * ((1 << 24) // 32), 0
    while True:
        chunks = (slots + 15) // 16
        levels += 1
        if chunks == 1:
            break
        slots = 2 * chunks
    assert levels == 7


def test_uniform_buckets_finish_in_the_join():
    """About 64 entries per bucket, as at 2^20 terms and c = 16: every bucket
    completes in the first stage or the join, so the levels find only empty
    slots."""
    rng = random.Random(5)
    E, nb = 60000, 940
    keys = sorted(rng.randrange(nb) for _ in range(E))
    bsum = Sums()
    sk, sp, ch0 = seg0(keys, [1] * E, [False] * E, 32, bsum)
    ok, _ = join(sk, sp, ch0, 8, bsum)
    assert all(is_empty(k) for k in ok)
    assert len(bsum.v) == len(set(keys))


def test_cont_marks_only_chunks_inside_one_run():
    """A CONT tail appears exactly on chunks whose entries all belong to one
    bucket that also continues on both sides."""
    rng = random.Random(11)
    for E, nb in ((5000, 40), (3000, 7), (64, 1)):
        keys = sorted(rng.randrange(nb) for _ in range(E))
        sk, _, ch0 = seg0(keys, [1] * E, [False] * E, 32, Sums())
        for j in range(ch0):
            beg, end = 32 * j, min(32 * j + 32, E)
            inside = (beg > 0 and end < E and keys[beg - 1] == keys[beg] == keys[end - 1] == keys[end])
            assert (sk[2 * j + 1] == CONT) == inside

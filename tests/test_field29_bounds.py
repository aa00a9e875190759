"""Bounds and fold constants of the 9 x 29-bit field layer (secp_field29.h),
checked on the CPU.

The layer serves exponentiation chains only (the square root), so every
multiplication input is a previous output or an exact conversion.  This test
replays the column / fold / carry sequence of fe29_columns and the two
fold-and-carry rounds of fe29_to_fe on Python integers:

* with every input limb at its documented maximum (every intermediate is
  monotone in the inputs, so this bounds all inputs): 64-bit column sums, the
  32-bit operands of v_mad_u64_u32, the carry t and the output limb bounds,
  which must again be valid inputs (the chain is closed);
* on random and extreme inputs: the replayed result is congruent to a * b
  mod p, so a wrong fold constant, column index or shift fails here, and
  fe29_to_fe returns a value below 2^256.
"""
import random

P = 2**256 - 2**32 - 977
M29, M24, M32, M64 = (1 << 29) - 1, (1 << 24) - 1, (1 << 32) - 1, (1 << 64) - 1
Q0, Q1 = 0x7A20, 0x100
IN_MAX = [M29 + (1 << 12)] * 8 + [M24]      # documented input / output bound


def test_constants():
    assert (2**261 - (Q1 * 2**29 + Q0)) % P == 0
    assert Q1 * Q0 <= M32 and Q1 * Q1 <= M32


def _mad(acc, a, b):
    assert 0 <= a <= M32 and 0 <= b <= M32, "v_mad_u64_u32 operand above 32 bits"
    r = acc + a * b
    assert r <= M64, "64-bit column overflow"
    return r


def columns(a, b):
    """fe29_columns on limb lists a, b; returns the 9 output limbs and t."""
    h = [0] * 9
    acc = 0
    for k in range(9, 17):
        for i in range(k - 8, 9):
            acc = _mad(acc, a[i], b[k - i])
        h[k - 9] = acc & M29
        acc >>= 29
    h[8] = acc
    l = [0] * 9
    acc = 0
    for k in range(9):
        for i in range(k + 1):
            acc = _mad(acc, a[i], b[k - i])
        acc = _mad(acc, h[k], Q0)
        if k >= 1:
            acc = _mad(acc, h[k - 1], Q1)
        if k == 0:
            acc = _mad(acc, h[8], Q1 * Q0)
        if k == 1:
            acc = _mad(acc, h[8], Q1 * Q1)
        if k < 8:
            l[k] = acc & M29
            acc >>= 29
        else:
            l[8] = acc & M24
            acc >>= 24
    t = acc
    assert t * 977 + M29 <= M64 and (t << 3) + M64 // 2**29 + M29 <= M64
    x0 = t * 977 + l[0]
    x1 = (t << 3) + (x0 >> 29) + l[1]
    r = [x0 & M29, x1 & M29, l[2] + (x1 >> 29)] + l[3:]
    return r, t


def value(l):
    return sum(x << (29 * i) for i, x in enumerate(l))


def to_fe(l):
    l = list(l)
    for _ in range(2):
        x = l[8] >> 24
        l[8] &= M24
        l[0] += x * 977
        l[1] += x << 3
        for i in range(8):
            l[i + 1] += l[i] >> 29
            l[i] &= M29
    return value(l)


def test_worst_case_is_closed():
    r, t = columns(IN_MAX, IN_MAX)
    assert t < 2**38
    assert all(x <= m for x, m in zip(r, IN_MAX)), "outputs exceed the documented input bound"
    # squaring doubles one operand (d = 2a): the cross products are the same
    # column sums, so the multiply's bound covers the square


def test_values_against_big_integers():
    rng = random.Random(29)
    cases = [(IN_MAX, IN_MAX), ([0] * 9, IN_MAX), ([M29] * 8 + [M24], [M29] * 8 + [M24])]
    for _ in range(400):
        a = [rng.randint(0, m) for m in IN_MAX]
        b = [rng.randint(0, m) for m in IN_MAX]
        cases.append((a, b))
    for a, b in cases:
        r, _ = columns(a, b)
        assert value(r) % P == value(a) * value(b) % P
        v = to_fe(r)
        assert v < 2**256 and v % P == value(r) % P


def test_to_fe_extremes():
    for l in (IN_MAX, [M29] * 8 + [M24], [0] * 9, [M29] * 9):
        v = to_fe(l)
        assert v < 2**256 and v % P == value(l) % P

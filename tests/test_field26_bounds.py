"""Overflow bounds of the 10 x 26-bit field layer (secp_field26.h,
secp_point26.h), checked on the CPU.

The layer reduces lazily: additions and subtractions are limb-wise 32-bit adds
and the multiply accumulates 64-bit column sums without carry-outs.  Its
correctness therefore rests on limb bounds, which this test derives:

* the multiply's worst case (every input limb at its maximum) for the largest
  input magnitude the layer allows, replaying the exact column / fold / carry
  sequence of fe26_columns with 64-bit and 32-bit overflow checks (every
  intermediate is monotone in the inputs, so the all-maximum input bounds
  every other input);
* the same column / fold / carry sequence on actual values, against
  a * b mod p for random and all-maximum inputs up to that magnitude;
* limb-bound propagation through the doubling, the mixed addition and the
  general addition exactly as secp_point26.h writes them, iterated from the
  loop's entry values to a fixed point: every multiply input stays within the
  allowed magnitude, every fe26_norm input below 2^31, every limb below 2^32, and each fe26_sub<K> has
  K * p limb-wise at least its subtrahend's bound.  The K constants written
  in secp_point26.h are asserted to be sufficient.  The general addition takes
  both operands from the same bound set (loop outputs or exact conversions).
"""
P = 2**256 - 2**32 - 977
M26, M22, M32, M64 = (1 << 26) - 1, (1 << 22) - 1, (1 << 32) - 1, (1 << 64) - 1
LIM = [M26] * 9 + [M22]
PL = [(P >> (26 * i)) & lim for i, lim in enumerate(LIM)]
R0, R1 = 0x3D10, 0x400
MUL_MAX_MAG = 20          # largest input magnitude the multiply accepts
LOOP_MUL_MAG = 16         # largest magnitude the loops feed it

# fe26_sub<K> / fe26_neg<K> constants as written in secp_point26.h
DBL_K = {"-A": 2, "-C": 2, "X3": 3, "D-X3": 5, "Y3": 9}
MADD_K = {"h": 7, "s2-y": 11, "-j": 2, "X3": 3, "v-X3": 7, "Y3": 3, "-z1z1": 2, "Z3": 2, "neg y": 2}
ADD_K = {"h": 2, "s2-s1": 2, "-j": 2, "X3": 3, "v-X3": 7, "Y3": 3}


def test_constants():
    assert sum(l << (26 * i) for i, l in enumerate(PL)) == P
    assert PL[:2] == [0x3FFFC2F, 0x3FFFFBF] and PL[2:9] == [M26] * 7 and PL[9] == M22
    assert (2**260 - (R1 * 2**26 + R0)) % P == 0


def _mad(a, b, c):
    assert 0 <= a <= M32 and 0 <= b <= M32
    r = a * b + c
    assert r <= M64, "64-bit column overflow"
    return r


def mul_worst(mag):
    """fe26_columns on all-maximum inputs of magnitude `mag`; returns the
    output limbs and the intermediate maxima (t, h[9])."""
    a = [mag * lim for lim in LIM]
    b = list(a)
    h = [0] * 10
    acc = 0
    for k in range(10, 19):
        for i in range(k - 9, 10):
            acc = _mad(a[i], b[k - i], acc)
        h[k - 10] = acc & M26
        acc >>= 26
    h[9] = acc
    assert h[9] <= M32
    l = [0] * 10
    acc = 0
    for k in range(10):
        for i in range(k + 1):
            acc = _mad(a[i], b[k - i], acc)
        acc = _mad(h[k], R0, acc)
        if k >= 1:
            acc = _mad(h[k - 1], R1, acc)
        if k == 0:
            acc = _mad(h[9], R1 * R0, acc)
        if k == 1:
            acc = _mad(h[9], R1 * R1, acc)
        if k < 9:
            l[k] = acc & M26
            acc >>= 26
        else:
            l[9] = acc & M22
            acc >>= 22
    t = acc
    assert t < 2**42
    x0 = t * 977 + M26                      # l[0] at most M26
    x1 = (t << 6) + (x0 >> 26) + M26
    assert x0 <= M64 and x1 <= M64
    return {"t": t, "h9": h[9], "l2_excess": x1 >> 26}


def test_multiply_worst_case():
    w = mul_worst(MUL_MAX_MAG)
    assert w["l2_excess"] <= 1 << 22          # documented: l[2] <= 2^26 - 1 + 2^22
    try:
        mul_worst(24)
        raised = False
    except AssertionError:
        raised = True
    assert raised, "the multiply's limit moved: recheck MUL_MAX_MAG"


def mul_replay(a, b):
    """fe26_columns on actual limb values (the same columns, folds and
    carries as the device code), returning the output limbs."""
    h = [0] * 10
    acc = 0
    for k in range(10, 19):
        for i in range(k - 9, 10):
            acc += a[i] * b[k - i]
        h[k - 10] = acc & M26
        acc >>= 26
    h[9] = acc
    l = [0] * 10
    acc = 0
    for k in range(10):
        for i in range(k + 1):
            acc += a[i] * b[k - i]
        acc += h[k] * R0
        if k >= 1:
            acc += h[k - 1] * R1
        if k == 0:
            acc += h[9] * R1 * R0
        if k == 1:
            acc += h[9] * R1 * R1
        assert acc <= M64
        if k < 9:
            l[k] = acc & M26
            acc >>= 26
        else:
            l[9] = acc & M22
            acc >>= 22
    t = acc
    x0 = t * 977 + l[0]
    x1 = (t << 6) + (x0 >> 26) + l[1]
    return [x0 & M26, x1 & M26, l[2] + (x1 >> 26)] + l[3:]


def test_multiply_replay_values():
    """The fold algebra gives a * b mod p for inputs up to the multiply's
    magnitude limit, all-maximum limbs included."""
    import random
    rnd = random.Random(2636)
    val = lambda v: sum(x << (26 * i) for i, x in enumerate(v))
    cases = [([m * lim for lim in LIM], [m * lim for lim in LIM]) for m in (1, 16, MUL_MAX_MAG)]
    for _ in range(3000):
        m1, m2 = rnd.choice((1, 2, 8, 16, MUL_MAX_MAG)), rnd.choice((1, 4, 16, MUL_MAX_MAG))
        cases.append(([rnd.randint(0, m1 * lim) for lim in LIM], [rnd.randint(0, m2 * lim) for lim in LIM]))
    for a, b in cases:
        r = mul_replay(a, b)
        assert val(r) % P == val(a) * val(b) % P
        assert all(x <= bound for x, bound in zip(r, [M26, M26, M26 + (1 << 22)] + [M26] * 6 + [M22]))


# --------------------------------------------------------------------------
# bound propagation through the formulas
# --------------------------------------------------------------------------
OUT = [M26, M26, M26 + (1 << 22)] + [M26] * 6 + [M22]     # multiply / square output
NORM_OUT = [M26] * 9 + [(1 << 22) + 64]


def _mag(v):
    return max(x / l for x, l in zip(v, LIM))


def _chk(v):
    assert all(x <= M32 for x in v)
    return v


def mul(a, b):
    assert _mag(a) <= LOOP_MUL_MAG and _mag(b) <= LOOP_MUL_MAG
    return list(OUT)


def add(a, b):
    return _chk([x + y for x, y in zip(a, b)])


def shl(a, s):
    return _chk([x << s for x in a])


def sub(a, b, k):
    assert all(k * pl >= x for pl, x in zip(PL, b)), "fe26_sub K too small"
    return _chk([x + k * pl for x, pl in zip(a, PL)])


def norm(a):
    assert all(x < 2**31 for x in a)
    return list(NORM_OUT)


def dbl(X, Y, Z):
    A = mul(X, X)
    B = mul(Y, Y)
    C = mul(B, B)
    t = mul(add(X, B), add(X, B))
    t = sub(t, A, DBL_K["-A"])
    t = sub(t, C, DBL_K["-C"])
    D = norm(add(t, t))
    E = _chk([3 * x for x in A])
    F = mul(E, E)
    Z3 = mul(Y, Z)
    Z3 = add(Z3, Z3)
    X3 = sub(F, add(D, D), DBL_K["X3"])
    t = mul(E, sub(D, X3, DBL_K["D-X3"]))
    Y3 = sub(t, shl(C, 3), DBL_K["Y3"])
    return X3, Y3, Z3


def madd(X, Y, Z, bx, by):
    z1z1 = mul(Z, Z)
    u2 = mul(bx, z1z1)
    s2 = mul(by, mul(Z, z1z1))
    h = norm(sub(u2, X, MADD_K["h"]))
    t = sub(s2, Y, MADD_K["s2-y"])
    rr = norm(add(t, t))
    hh = mul(h, h)
    i4 = shl(hh, 2)
    j = mul(h, i4)
    v = mul(X, i4)
    t = sub(mul(rr, rr), j, MADD_K["-j"])
    X3 = sub(t, add(v, v), MADD_K["X3"])
    t = mul(rr, sub(v, X3, MADD_K["v-X3"]))
    j = mul(Y, j)
    Y3 = sub(t, add(j, j), MADD_K["Y3"])
    t = mul(add(Z, h), add(Z, h))
    t = sub(t, z1z1, MADD_K["-z1z1"])
    Z3 = sub(t, hh, MADD_K["Z3"])
    return X3, Y3, Z3


def gadd(X1, Y1, Z1, X2, Y2, Z2):
    """jac26_add (add-2007-bl with Z3 = 2 Z1 Z2 H)."""
    z1z1 = mul(Z1, Z1)
    z2z2 = mul(Z2, Z2)
    u1 = mul(X1, z2z2)
    u2 = mul(X2, z1z1)
    s1 = mul(Y1, mul(Z2, z2z2))
    s2 = mul(Y2, mul(Z1, z1z1))
    h = norm(sub(u2, u1, ADD_K["h"]))
    t = sub(s2, s1, ADD_K["s2-s1"])
    rr = norm(add(t, t))
    i4 = shl(mul(h, h), 2)
    j = mul(h, i4)
    v = mul(u1, i4)
    t = sub(mul(rr, rr), j, ADD_K["-j"])
    X3 = sub(t, add(v, v), ADD_K["X3"])
    t = mul(rr, sub(v, X3, ADD_K["v-X3"]))
    j = mul(s1, j)
    Y3 = sub(t, add(j, j), ADD_K["Y3"])
    z = mul(mul(Z1, Z2), h)
    Z3 = add(z, z)
    return X3, Y3, Z3


def _vmax(*vs):
    return [max(c) for c in zip(*vs)]


def test_formulas_reach_a_fixed_point():
    exact = list(LIM)                                   # fe26_from_fe of a stored coordinate
    neg_y = sub([0] * 10, exact, MADD_K["neg y"])       # fe26_neg<2>
    bx = _vmax(exact, OUT)                              # x or beta * x
    by = _vmax(exact, neg_y)
    one = [1] + [0] * 9
    # entry values: an affine point (jac26_from_aff), infinity (1, 1, 0)
    X, Y, Z = exact, exact, _vmax(one, OUT)
    for _ in range(10):
        d = dbl(X, Y, Z)
        m = madd(X, Y, Z, bx, by)
        g = gadd(X, Y, Z, X, Y, Z)      # both operands anywhere in the state
        # also the madd's own special outputs: b itself (a = inf) and a itself
        Xn = _vmax(X, d[0], m[0], g[0], bx)
        Yn = _vmax(Y, d[1], m[1], g[1], by)
        Zn = _vmax(Z, d[2], m[2], g[2], OUT)
        if (Xn, Yn, Zn) == (X, Y, Z):
            break
        X, Y, Z = Xn, Yn, Zn
    else:
        raise AssertionError("no fixed point")
    assert _mag(X) < 6.2 and _mag(Y) < 10.2 and _mag(Z) < 5.2


# --------------------------------------------------------------------------
# the wave-cooperative form (secp_wide26.h): lane t of a 21-lane group holds
# limb / column t; replayed lane by lane
# --------------------------------------------------------------------------
WIDE_OUT = [(1 << 26) + (1 << 10)] * 9 + [1 << 22]          # wmul output bound
WIDE_NORM_OUT = [1 << 26] * 9 + [(1 << 22) + 33]           # wnorm output bound


def _shr1(v):
    """DPP wave_shr:1 within one group (lane 0 receives 0: the previous
    group's lane 20 is zero in every operand, and the carry into it is
    masked)."""
    return [0] + v[:-1]


def wmul_replay(a, b, check=True):
    """wmul on limb lists a, b (10 limbs): columns, two carry rounds, the
    2^260 fold with lanes 19 and 20 folded twice, two carry rounds with limb 9
    split at 22 bits.  With `check`, every 64-bit intermediate is bounded."""
    lim64 = (lambda x: x <= M64) if check else (lambda x: True)
    A = list(a) + [0] * 11
    acc = [0] * 21
    sh = list(A)
    for j in range(10):
        for t in range(21):
            assert sh[t] <= M32 and b[j] <= M32
            acc[t] += sh[t] * b[j]
            assert lim64(acc[t]), "column overflow"
        if j < 9:
            sh = _shr1(sh)
    for _ in range(2):
        hi = [x >> 26 for x in acc]
        inc = _shr1(hi)
        acc = [(x & M26) + i for x, i in zip(acc, inc)]
    L = acc
    assert all(x <= M32 for x in L)
    V = R1 * L[19] + R0 * L[20]
    V11 = R1 * L[20]
    m = [0] * 21
    for t in range(10):
        m[t] = L[t] + R0 * L[t + 10] + (R1 * L[t + 9] if t >= 1 else 0)
        if t == 0:
            m[t] += R0 * V
        if t == 1:
            m[t] += R1 * V + R0 * V11
        if t == 2:
            m[t] += R1 * V11
        assert lim64(m[t])
    for _ in range(2):
        hi = [(x >> (22 if t == 9 else 26)) for t, x in enumerate(m)]
        inc = _shr1(hi)
        h9 = hi[9]
        new = [0] * 21
        for t in range(10):
            keep = m[t] & ((1 << (22 if t == 9 else 26)) - 1)
            new[t] = keep + (inc[t] if t >= 1 else 0)
            if t == 0:
                new[t] += h9 * 977
            if t == 1:
                new[t] += h9 << 6
            assert lim64(new[t])
        m = new
    assert all(x == 0 for x in m[10:])
    return m[:10]


def wnorm_replay(a):
    assert all(x < 2**31 for x in a)
    x = a[9] >> 22
    v = list(a)
    v[9] &= M22
    v[0] += x * 977
    v[1] += x << 6
    for _ in range(2):
        hi = [0 if t == 9 else v[t] >> 26 for t in range(10)]
        inc = [0] + hi[:9]
        v = [(v[t] if t == 9 else v[t] & M26) + inc[t] for t in range(10)]
    return v


def test_wide_multiply_worst_case_and_values():
    """The wide product: no 64-bit overflow at the largest input magnitude,
    outputs within WIDE_OUT, and a * b mod p on random and all-maximum
    inputs; the wide normalisation within WIDE_NORM_OUT."""
    import random
    rnd = random.Random(2626)
    val = lambda v: sum(x << (26 * i) for i, x in enumerate(v))
    cases = [([m * lim for lim in LIM], [m * lim for lim in LIM]) for m in (1, 16, MUL_MAX_MAG)]
    for _ in range(3000):
        m1, m2 = rnd.choice((1, 2, 8, 16, MUL_MAX_MAG)), rnd.choice((1, 4, 16, MUL_MAX_MAG))
        cases.append(([rnd.randint(0, m1 * lim) for lim in LIM], [rnd.randint(0, m2 * lim) for lim in LIM]))
    for a, b in cases:
        r = wmul_replay(a, b)
        assert val(r) % P == val(a) * val(b) % P
        assert all(x <= bound for x, bound in zip(r, WIDE_OUT)), r
    for _ in range(3000):
        a = [rnd.randint(0, 2**31 - 1) for _ in range(10)]
        r = wnorm_replay(a)
        assert val(r) % P == val(a) % P
        assert all(x <= bound for x, bound in zip(r, WIDE_NORM_OUT)), r
    r = wnorm_replay([2**31 - 1] * 10)
    assert all(x <= bound for x, bound in zip(r, WIDE_NORM_OUT))


def test_wide_doubling_and_final_additions_reach_a_fixed_point():
    """The MSM final: wide doublings (wdbl, same K constants as the one-lane
    doubling, wide multiply / normalisation output bounds) alternate with the
    one-lane general addition of an exactly converted window sum.  From an
    exactly converted entry point the state reaches a fixed point with every
    multiply input within LOOP_MUL_MAG and every K sufficient."""
    global OUT, NORM_OUT
    saved = OUT, NORM_OUT
    exact = list(LIM)
    one = [1] + [0] * 9

    def wdbl(X, Y, Z):
        global OUT, NORM_OUT
        OUT, NORM_OUT = WIDE_OUT, WIDE_NORM_OUT
        try:
            return dbl(X, Y, Z)
        finally:
            OUT, NORM_OUT = saved

    X, Y, Z = exact, exact, _vmax(one, exact)
    try:
        for _ in range(12):
            d = wdbl(X, Y, Z)
            g = gadd(X, Y, Z, exact, exact, exact)
            g2 = gadd(X, Y, Z, X, Y, Z)
            Xn, Yn, Zn = _vmax(X, d[0], g[0], g2[0]), _vmax(Y, d[1], g[1], g2[1]), _vmax(Z, d[2], g[2], g2[2])
            if (Xn, Yn, Zn) == (X, Y, Z):
                break
            X, Y, Z = Xn, Yn, Zn
        else:
            raise AssertionError("no fixed point")
    finally:
        OUT, NORM_OUT = saved
    assert _mag(X) < 6.2 and _mag(Y) < 10.2 and _mag(Z) < 5.2

"""Device field / scalar / GLV / point primitives against textbook integer math.

Each op runs in the HIP self-test kernel (``ufsecp_gpu_ext_selftest``) on
random values plus the edge values a limb implementation gets wrong
(0, 1, p-1, p, 2^256-1, values just above p, carries across every limb).
"""
import random

import numpy as np
import pytest

from ec_math import BETA, G, LAMBDA, N, P, add, from_limbs, lift_x, mul, to_limbs

pytestmark = pytest.mark.gpu

OP_FE_MUL, OP_FE_SQR, OP_FE_ADD, OP_FE_SUB, OP_FE_INV, OP_FE_SQRT = 0, 1, 2, 3, 4, 5
OP_FE_INV_FERMAT = 8
OP_SC_MUL, OP_SC_INV, OP_GLV, OP_SC_INV_FERMAT = 10, 11, 12, 13
OP_GEN_MUL, OP_VAR_MUL = 20, 21

M256 = (1 << 256) - 1
EDGE = [0, 1, 2, P - 1, P, P + 1, M256, M256 - 1, (1 << 255), (1 << 32) - 1, 977, 0x1000003D1,
        (1 << 224) - 1, P - 0x1000003D1, N - 1, N, N + 1] + [((1 << 32) - 1) << (32 * i) for i in range(8)]


def _vals(n, seed, edge=EDGE):
    rnd = random.Random(seed)
    vals = list(edge) + [rnd.getrandbits(256) for _ in range(n)]
    return vals


def _pack(vals):
    a = np.zeros((len(vals), 16), np.uint32)
    for i, v in enumerate(vals):
        a[i, :8] = to_limbs(v)
    return a


def test_field_mul_sqr_add_sub(gpu):
    xs = _vals(300, 1)
    ys = list(reversed(_vals(300, 2)))
    a, b = _pack(xs), _pack(ys)
    for op, f in ((OP_FE_MUL, lambda x, y: x * y % P), (OP_FE_SQR, lambda x, y: x * x % P),
                  (OP_FE_ADD, lambda x, y: (x + y) % P), (OP_FE_SUB, lambda x, y: (x - y) % P)):
        out = gpu.selftest(op, a, b)
        for i, (x, y) in enumerate(zip(xs, ys)):
            assert from_limbs(out[i, :8]) == f(x, y), (op, hex(x), hex(y))


def test_field_inv_sqrt(gpu):
    xs = [v for v in _vals(100, 3) if v % P != 0]
    a = _pack(xs)
    out = gpu.selftest(OP_FE_INV, a, a)
    for i, x in enumerate(xs):
        assert from_limbs(out[i, :8]) == pow(x, -1, P), hex(x)
    out = gpu.selftest(OP_FE_SQRT, a, a)
    for i, x in enumerate(xs):
        r = pow(x % P, (P + 1) // 4, P)
        ok = r * r % P == x % P
        assert int(out[i, 8]) == int(ok)
        assert from_limbs(out[i, :8]) == r


def test_scalar_mul_inv(gpu):
    xs = [v % N for v in _vals(200, 4)]
    ys = [v % N for v in _vals(200, 5)][::-1]
    out = gpu.selftest(OP_SC_MUL, _pack(xs), _pack(ys))
    for i, (x, y) in enumerate(zip(xs, ys)):
        assert from_limbs(out[i, :8]) == x * y % N
    nz = [x for x in xs if x]
    out = gpu.selftest(OP_SC_INV, _pack(nz), _pack(nz))
    for i, x in enumerate(nz):
        assert from_limbs(out[i, :8]) == pow(x, -1, N)


def _inv_inputs(m, n, seed):
    """Random residues plus inputs that stress the divsteps batches: 0, 1,
    m-1, powers of two (long runs of even steps), 2^k - 1 (long odd runs),
    (m+1)/2 and values that differ from m in the low word only."""
    rnd = random.Random(seed)
    xs = [0, 1, 2, 3, m - 1, m - 2, (m + 1) // 2, (m - 1) // 2, m - (1 << 32), (1 << 256) - 1 - m]
    xs += [1 << k for k in range(0, 256, 7)] + [(1 << k) - 1 for k in range(2, 256, 9)]
    xs += [m - (1 << k) for k in range(1, 250, 13)]
    xs += [rnd.randrange(m) for _ in range(n)]
    return [x % m for x in xs]


@pytest.mark.parametrize("mod", ["p", "n"])
def test_divsteps_inverse(gpu, mod):
    """Divsteps inversion == Python pow(x, -1, m) == the a^(m-2) chain, 0 -> 0."""
    m = P if mod == "p" else N
    op, op_fermat = (OP_FE_INV, OP_FE_INV_FERMAT) if mod == "p" else (OP_SC_INV, OP_SC_INV_FERMAT)
    xs = _inv_inputs(m, 4000, 11 if mod == "p" else 12)
    a = _pack(xs)
    out = gpu.selftest(op, a, a)
    ref = gpu.selftest(op_fermat, a, a)
    for i, x in enumerate(xs):
        exp = pow(x, -1, m) if x else 0
        assert from_limbs(out[i, :8]) == exp, hex(x)
        assert from_limbs(ref[i, :8]) == exp, hex(x)
    if mod == "p":   # weakly reduced inputs in [p, 2^256)
        ys = [P, P + 1, P + 977, (1 << 256) - 1]
        out = gpu.selftest(OP_FE_INV, _pack(ys), _pack(ys))
        for i, y in enumerate(ys):
            exp = pow(y % P, -1, P) if y % P else 0
            assert from_limbs(out[i, :8]) == exp, hex(y)


def test_glv_split(gpu):
    ks = [v % N for v in _vals(300, 6)] + [LAMBDA, N - LAMBDA, (N - 1) // 2, (N + 1) // 2]
    out = gpu.selftest(OP_GLV, _pack(ks), _pack(ks))
    for i, k in enumerate(ks):
        k1, k2 = from_limbs(out[i, :8]), from_limbs(out[i, 8:16])
        n1, n2 = int(out[i, 16]), int(out[i, 17])
        assert k1 < 2 ** 128 and k2 < 2 ** 128
        s1 = -k1 if n1 else k1
        s2 = -k2 if n2 else k2
        assert (s1 + s2 * LAMBDA - k) % N == 0


def test_generator_mul(gpu):
    rnd = random.Random(7)
    ks = [1, 2, 3, N - 1, N - 2, (N - 1) // 2, 0, 1 << 255, (1 << 128) + 1] + [rnd.randrange(N) for _ in range(60)]
    out = gpu.selftest(OP_GEN_MUL, _pack(ks), _pack(ks))
    for i, k in enumerate(ks):
        exp = mul(k)
        if exp is None:
            assert int(out[i, 16]) == 1
        else:
            assert int(out[i, 16]) == 0
            assert (from_limbs(out[i, :8]), from_limbs(out[i, 8:16])) == exp, k


def test_variable_mul_glv(gpu):
    rnd = random.Random(8)
    xs, ks, pts = [], [], []
    while len(xs) < 40:
        x = rnd.getrandbits(256) % P
        pt = lift_x(x)
        if pt is None:
            continue
        xs.append(x)
        pts.append(pt)
        ks.append(rnd.randrange(N))
    ks[0], ks[1], ks[2], ks[3] = 0, 1, N - 1, LAMBDA
    out = gpu.selftest(OP_VAR_MUL, _pack(xs), _pack(ks))
    for i, (pt, k) in enumerate(zip(pts, ks)):
        assert int(out[i, 17]) == 1
        exp = mul(k, pt)
        if exp is None:
            assert int(out[i, 16]) == 1
        else:
            assert (from_limbs(out[i, :8]), from_limbs(out[i, 8:16])) == exp, i
    # lambda * P == (beta * x, y)
    assert mul(LAMBDA, pts[3]) == (BETA * pts[3][0] % P, pts[3][1])
    _ = (G, add)


def test_field_mul_sqr_chains(gpu):
    """4096 dependent multiplications / squarings per lane (the hot carry chains
    back to back) agree with x*y^4096 and x^(2^4096)."""
    xs = _vals(60, 9)
    ys = list(reversed(_vals(60, 10)))
    a, b = _pack(xs), _pack(ys)
    out = gpu.selftest(6, a, b)
    for i, (x, y) in enumerate(zip(xs, ys)):
        assert from_limbs(out[i, :8]) == x * pow(y, 4096, P) % P
    out = gpu.selftest(7, a, b)
    for i, x in enumerate(xs):
        assert from_limbs(out[i, :8]) == pow(x, 1 << 4096, P)


# ---------------------------------------------------------------------------
# 10 x 26-bit limb layer (secp_field26.h) used by the scalar-multiplication
# loops: products at every input magnitude the multiply allows (up to 20 x
# the limb maximum), the documented output bounds, weak normalisation and the
# 8 x 32 round trip.
# ---------------------------------------------------------------------------
OP_FE26_MUL, OP_FE26_SQR, OP_FE26_NORM, OP_FE26_FROMTO = 30, 31, 32, 33
M26, M22 = (1 << 26) - 1, (1 << 22) - 1
LIM26 = [M26] * 9 + [M22]


def _v26(limbs):
    return sum(int(x) << (26 * i) for i, x in enumerate(limbs))


def _limbs26(rnd, mag):
    return [rnd.randint(0, mag * lim) for lim in LIM26]


def _pack26(rows):
    a = np.zeros((len(rows), 16), np.uint32)
    for i, r in enumerate(rows):
        a[i, :10] = r
    return a


def test_fe26_mul_sqr(gpu):
    rnd = random.Random(26)
    xs, ys = [], []
    for mag in (1, 2, 4, 8, 12, 16, 20):
        xs.append([mag * lim for lim in LIM26])        # every limb at its maximum
        ys.append([mag * lim for lim in LIM26])
        for _ in range(40):
            xs.append(_limbs26(rnd, mag))
            ys.append(_limbs26(rnd, rnd.choice((1, 2, mag))))
    xs += [[0] * 10, [1] + [0] * 9, [((P >> (26 * i)) & lim) for i, lim in enumerate(LIM26)]]
    ys += [[M26] * 9 + [M22], [0] * 10, [2] + [0] * 9]
    a, b = _pack26(xs), _pack26(ys)
    for op, f in ((OP_FE26_MUL, lambda x, y: x * y), (OP_FE26_SQR, lambda x, y: x * x)):
        out = gpu.selftest(op, a, b)
        for i, (x, y) in enumerate(zip(xs, ys)):
            want = f(_v26(x), _v26(y)) % P
            assert from_limbs(out[i, :8]) % P == want, (op, i)
            r26 = out[i, 8:18]
            assert all(int(r26[l]) <= M26 for l in (0, 1, 3, 4, 5, 6, 7, 8)), (op, i)
            assert int(r26[2]) <= M26 + (1 << 22) and int(r26[9]) <= M22, (op, i)
            assert _v26(r26) % P == want


def test_fe26_norm_and_round_trip(gpu):
    rnd = random.Random(27)
    rows = [[(1 << 31) - 1] * 10, [0] * 10, [((P >> (26 * i)) & lim) for i, lim in enumerate(LIM26)]]
    rows += [[rnd.randrange(1 << 31) for _ in range(10)] for _ in range(300)]
    rows += [_limbs26(rnd, m) for m in (1, 2, 24) for _ in range(50)]
    out = gpu.selftest(OP_FE26_NORM, _pack26(rows), _pack26(rows))
    for i, r in enumerate(rows):
        n = [int(v) for v in out[i, :10]]
        assert all(v <= M26 for v in n[:9]) and n[9] <= (1 << 22) + 64, i
        assert _v26(n) < 2 * P and _v26(n) % P == _v26(r) % P, i
    xs = _vals(200, 28)
    out = gpu.selftest(OP_FE26_FROMTO, _pack(xs), _pack(xs))
    for i, x in enumerate(xs):
        assert _v26(out[i, 8:18]) == x, i                 # exact limbs
        assert from_limbs(out[i, :8]) % P == x % P and from_limbs(out[i, :8]) < 2**256, i


OP_JAC26_MADD, OP_JAC26_DBL, OP_JAC26_ADD = 34, 35, 36


def _aff_pack(points):
    a = np.zeros((len(points), 16), np.uint32)
    for i, (x, y) in enumerate(points):
        a[i, :8] = to_limbs(x)
        a[i, 8:] = to_limbs(y)
    return a


def test_jac26_madd_and_dbl_exceptional_cases(gpu):
    """The 10 x 26 mixed addition against the group law, including its
    exceptional paths: B' = A (doubling requested, accumulator unchanged),
    B' = -A (infinity), A at infinity (result B'), with negated and
    beta-mapped addends; and the doubling, infinity included.  Item i uses
    Z = i + 3, flags i & 3, A = infinity when i % 11 == 10."""
    rnd = random.Random(34)
    As, Bs = [], []
    n = 220
    for i in range(n):
        A = mul(rnd.randrange(1, N))
        flags = i & 3
        kind = i % 5
        if kind == 0:     # B' = A: choose B so that the flag transform gives A
            x, y = A
            if flags & 2:
                x = x * pow(BETA, 2, P) % P      # beta^3 = 1, so beta^2 x maps to x
            if flags & 1:
                y = (P - y) % P
            B = (x, y)
        elif kind == 1:   # B' = -A
            x, y = A
            if flags & 2:
                x = x * pow(BETA, 2, P) % P
            if not flags & 1:
                y = (P - y) % P
            B = (x, y)
        else:
            B = mul(rnd.randrange(1, N))
        As.append(A)
        Bs.append(B)
    a, b = _aff_pack(As), _aff_pack(Bs)
    out = gpu.selftest(OP_JAC26_MADD, a, b)
    for i in range(n):
        flags = i & 3
        bx, by = Bs[i]
        if flags & 2:
            bx = bx * BETA % P
        if flags & 1:
            by = (P - by) % P
        Bp = (bx, by)
        a_inf = i % 11 == 10
        nd = int(out[i, 17])
        if a_inf:
            exp, want_nd = Bp, 0
        elif Bp == As[i]:
            exp, want_nd = As[i], 1                  # accumulator unchanged, caller doubles
        else:
            exp, want_nd = add(As[i], Bp), 0
        assert nd == want_nd, i
        if exp is None:
            assert int(out[i, 16]) == 1, i
        else:
            assert int(out[i, 16]) == 0, i
            assert (from_limbs(out[i, :8]), from_limbs(out[i, 8:16])) == exp, i
    assert sum(int(out[i, 17]) for i in range(n)) >= 30            # the a == b path ran
    assert sum(int(out[i, 16]) for i in range(n) if i % 11 != 10) >= 30   # the a == -b path ran
    out = gpu.selftest(OP_JAC26_DBL, a, b)
    for i in range(n):
        if i % 11 == 10:
            assert int(out[i, 16]) == 1, i
        else:
            assert (from_limbs(out[i, :8]), from_limbs(out[i, 8:16])) == add(As[i], As[i]), i


@pytest.mark.gpu
def test_jac26_add_all_cases(gpu):
    """The 10 x 26 general Jacobian addition (MSM window sums, tree reduce,
    final) against the affine group law: B = A (doubling), B = -A (infinity),
    either operand at infinity, and generic sums.  Item i uses Z = i + 3 for A
    and Z = i + 5 for B; A = infinity when i % 11 == 10, B = infinity when
    i % 13 == 12."""
    rnd = random.Random(36)
    As, Bs = [], []
    n = 260
    for i in range(n):
        A = mul(rnd.randrange(1, N))
        kind = i % 5
        if kind == 0:
            B = A
        elif kind == 1:
            B = (A[0], (P - A[1]) % P)
        else:
            B = mul(rnd.randrange(1, N))
        As.append(A)
        Bs.append(B)
    out = gpu.selftest(OP_JAC26_ADD, _aff_pack(As), _aff_pack(Bs))
    doubled = at_inf = 0
    for i in range(n):
        a = None if i % 11 == 10 else As[i]
        b = None if i % 13 == 12 else Bs[i]
        exp = add(a, b)
        if a is not None and a == b:
            doubled += 1
        if exp is None:
            at_inf += 1
            assert int(out[i, 16]) == 1, i
        else:
            assert int(out[i, 16]) == 0, i
            assert (from_limbs(out[i, :8]), from_limbs(out[i, 8:16])) == exp, i
    assert doubled >= 40 and at_inf >= 40


OP_JAC26_DBL_WIDE = 37


@pytest.mark.gpu
def test_jac26_dbl_wide_against_group_law(gpu):
    """The wave-cooperative doubling of the MSM final (field products spread
    over 21 lanes, three at a time): item i is A with Z = i + 3 doubled
    1 + i % 7 times (infinity when i % 11 == 10), against 2^k A by the affine
    group law; 300 items (several waves, every repeat count), G, -G and 2G
    among them."""
    rnd = random.Random(37)
    n = 300
    As = [mul(rnd.randrange(1, N)) for _ in range(n)]
    As[0], As[1], As[2] = mul(1), mul(N - 1), mul(2)
    out = gpu.selftest(OP_JAC26_DBL_WIDE, _aff_pack(As), _aff_pack(As))
    for i in range(n):
        if i % 11 == 10:
            assert int(out[i, 16]) == 1, i
            continue
        exp = As[i]
        for _ in range(1 + i % 7):
            exp = add(exp, exp)
        assert int(out[i, 16]) == 0, i
        assert (from_limbs(out[i, :8]), from_limbs(out[i, 8:16])) == exp, i


OP_JAC26_ADD_LANES = 38


@pytest.mark.gpu
def test_jac26_add_lanes_all_cases(gpu):
    """The MSM final's lane-spread general addition (one product per lane, five
    levels) against the affine group law, with the cases of
    test_jac26_add_all_cases: B = A (doubling), B = -A (infinity), either
    operand at infinity, generic sums; one whole wave per item."""
    rnd = random.Random(38)
    As, Bs = [], []
    n = 260
    for i in range(n):
        A = mul(rnd.randrange(1, N))
        kind = i % 5
        if kind == 0:
            B = A
        elif kind == 1:
            B = (A[0], (P - A[1]) % P)
        else:
            B = mul(rnd.randrange(1, N))
        As.append(A)
        Bs.append(B)
    out = gpu.selftest(OP_JAC26_ADD_LANES, _aff_pack(As), _aff_pack(Bs))
    doubled = at_inf = 0
    for i in range(n):
        a = None if i % 11 == 10 else As[i]
        b = None if i % 13 == 12 else Bs[i]
        exp = add(a, b)
        if a is not None and a == b:
            doubled += 1
        if exp is None:
            at_inf += 1
            assert int(out[i, 16]) == 1, i
        else:
            assert int(out[i, 16]) == 0, i
            assert (from_limbs(out[i, :8]), from_limbs(out[i, 8:16])) == exp, i
    assert doubled >= 40 and at_inf >= 40

"""CPU replay of the variable-base table build (ultrafastsecp256k1_amd/csrc/
secp_ecmult.h, build_ptable): the doubling that leaves P and 2P at one Z, the
seven co-Z additions (XYCZ-ADD), the rescale of entries 0..6 to entry 7's Z
and the stored Z scale zs.

The replay follows the device code step by step on Python integers mod p
(same operations, same order, same Z-ratio bookkeeping), and the textbook
affine group law of ec_math checks the result: every stored entry j, read
with Z = zs, must be (2j+1) P.  A dropped factor in a ratio, a swapped
operand in the co-Z addition or a wrong power in the rescale fails here for
every P.  The GPU tests check the same kernels end to end (bit-exact
verdicts against the reference library, and the arithmetic self-tests'
b*P against this group law)."""
import random

from ec_math import G, P, add, lift_x, mul

K_ENTRIES = 8


def build_ptable(x, y):
    """The device build_ptable on affine P = (x, y): returns the stored
    entries (x_j, y_j) and zs."""
    # the doubling from Z = 1: B = x^2, E = y^2, L = E^2, S = 2((x+E)^2 - B - L),
    # M = 3B; 2P = (M^2 - 2S, M(S - X) - 8L), P = (S, 8L), Z = 2y
    b = x * x % P
    e = y * y % P
    l = e * e % P
    s = 2 * (((x + e) ** 2 - b - l) % P) % P
    m = 3 * b % P
    qx = (m * m - 2 * s) % P
    qy = (m * (s - qx) - 8 * l) % P
    cx, cy = s, 8 * l % P
    entries = [(cx, cy)]
    zr = [None]
    for _ in range(1, K_ENTRIES):
        dx = (cx - qx) % P
        a = dx * dx % P
        bq = qx * a % P
        bc = cx * a % P
        dy = (cy - qy) % P
        ee = qy * (bc - bq) % P
        nx = (dy * dy - bq - bc) % P
        ny = (dy * (bq - nx) - ee) % P
        qx, qy = bq, ee
        cx, cy = nx, ny
        entries.append((cx, cy))
        zr.append(dx)
    # rescale entries 0..6 by s = zr[j+1] * ... * zr[7]
    sc = zr[K_ENTRIES - 1]
    for j in range(K_ENTRIES - 2, -1, -1):
        ex, ey = entries[j]
        s2 = sc * sc % P
        s3 = s2 * sc % P
        entries[j] = (ex * s2 % P, ey * s3 % P)
        if j > 0:
            sc = sc * zr[j] % P
    zs = sc * (2 * y) % P
    return entries, zs


def affine(X, Y, Z):
    zi = pow(Z, -1, P)
    return X * zi * zi % P, Y * zi * zi * zi % P


def test_entries_are_the_odd_multiples():
    rnd = random.Random(340)
    points = [G, mul(2), mul(7), mul(0xFFFF), mul(3 ** 100)]
    while len(points) < 40:
        pt = lift_x(rnd.getrandbits(256) % P, rnd.random() < 0.5)
        if pt is not None:
            points.append(pt)
    for pt in points:
        entries, zs = build_ptable(*pt)
        want = pt
        two = mul(2, pt)
        for j, (ex, ey) in enumerate(entries):
            assert affine(ex, ey, zs) == want, (pt, j)
            want = add(want, two)


def test_common_z_is_the_product_of_the_ratios():
    """zs is 2y times every step's ratio: entry 7 is stored unscaled, so its
    own Jacobian Z must be zs (the loop multiplies the accumulator's Z by zs
    before the fixed-base additions)."""
    pt = mul(12345)
    entries, zs = build_ptable(*pt)
    x7, y7 = entries[-1]
    assert affine(x7, y7, zs) == mul(15, pt)

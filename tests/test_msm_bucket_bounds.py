"""CPU replay of the MSM's term preparation and bucket-entry arithmetic
(ultrafastsecp256k1_amd/csrc/msm.hip: the GLV split, masking and digits of
k_msm_entries and msm_digit) for adversarial scalars and every window width
c = 2..20.

The accumulation kernel indexes bsum[bucket id] and pts[term index] with the
entries' fields, so these must stay below the bucket count (W + 1) * 2^(c-1)
and 2N for any bytes a caller passes: scalars >= n, 2^256 - 1, zero, halves at
the 2^128 edge.  Invalid points do not change the entries (their terms keep
their buckets and the call returns the error), so only the scalars matter
here.  The replay follows the device code step by
step (same constants, one conditional subtraction in sc_add, round-half-up of
k * g / 2^384), and the closed forms check it: k1 + lambda k2 == k (mod n) and
sum_w d_w 2^(c w) == the half, so a replay that drifted from the kernel would
fail here before it could hide a bound violation."""
import pathlib
import random
import re

import pytest

from ec_math import LAMBDA, N

M256 = (1 << 256) - 1
_HDR = (pathlib.Path(__file__).resolve().parents[1] / "ultrafastsecp256k1_amd" / "csrc" / "secp_field.h").read_text()


def _limbs(name):
    """A constant of the device code, from its 8 x 32-bit little-endian limbs."""
    ws = re.search(r"#define UFS_%s_LIMBS\s+(.*)" % name, _HDR).group(1).replace("U", "").split(",")
    return sum(int(w.strip(), 16) << (32 * i) for i, w in enumerate(ws))


G1, G2, MB1, MB2, MLAMBDA = (_limbs(x) for x in ("G1", "G2", "MB1", "MB2", "MLAMBDA"))
B1, B2 = -0xE4437ED6010E88286F547FA90ABFE4C3, 0x3086D221A7D46BCDE86C90E49284EB15   # published GLV basis


def test_device_split_constants_closed_form():
    """-lambda, -b1, -b2 mod n and g1 = round(2^384 b2 / n), g2 = round(2^384 (-b1) / n)."""
    assert MLAMBDA == N - LAMBDA
    assert MB1 == -B1 % N and MB2 == -B2 % N
    assert G1 == (B2 * (1 << 384) + N // 2) // N
    assert G2 == (-B1 * (1 << 384) + N // 2) // N
    assert (B2 + B1 * LAMBDA) % N == 0      # (a1, b1) = (b2, b1) is a lattice vector: a1 + b1 lambda = 0


def sc_add(a, b):
    """secp_field.h sc_add: a + b, minus n once if it carried or is >= n."""
    s = a + b
    return (s - N) & M256 if s >= N else s


def glv_split_dev(k):
    """secp_ecmult.h glv_split for any 256-bit k: (|k1|, |k2|, neg1, neg2)."""
    c1 = ((k * G1) >> 384) + (((k * G1) >> 383) & 1)
    c2 = ((k * G2) >> 384) + (((k * G2) >> 383) & 1)
    c1 = (c1 & ((1 << 128) - 1)) * MB1 % N
    c2 = (c2 & ((1 << 128) - 1)) * MB2 % N
    k2 = sc_add(c1, c2)
    k1 = sc_add(k2 * MLAMBDA % N, k)
    neg1, neg2 = k1 > N // 2, k2 > N // 2
    k1 = (N - k1) & M256 if neg1 and k1 else k1
    k2 = (N - k2) & M256 if neg2 and k2 else k2
    return k1, k2, neg1, neg2


def prepare(k):
    """k_msm_entries: halves cut to 128 bits; a scalar >= n gets zero halves,
    so it places only skip entries."""
    k1, k2, n1, n2 = glv_split_dev(k)
    if not k < N:
        return 0, 0
    return k1 & ((1 << 128) - 1), k2 & ((1 << 128) - 1)


def plan(c):
    W = (128 + c - 1) // c
    return W, 1 << (c - 1)


def digits(k, c):
    """msm_digit over the W windows of a value (any width: bits are read as
    the kernel reads the 8-limb array)."""
    W, _ = plan(c)
    out, carry = [], 0
    for w in range(W):
        raw = ((k >> (w * c)) & ((1 << c) - 1)) + carry
        if w == W - 1:
            out.append(raw)
            carry = 0
        elif raw > (1 << (c - 1)):
            out.append(raw - (1 << c))
            carry = 1
        else:
            out.append(raw)
            carry = 0
    return out


def entries(k, c, i):
    """k_msm_entries for half-term i: (bucket id, skip) per window."""
    W, B = plan(c)
    res = []
    for w, d in enumerate(digits(k, c)):
        b = abs(d) - 1 if d != 0 else i & (B - 1)
        res.append((w * B + b, d == 0))
    return res


def adversarial_scalars():
    rnd = random.Random(0xB0C4E7)
    xs = [0, 1, 2, N - 1, N - 2, N, N + 1, N + 3, M256, M256 - 1, 1 << 255, (1 << 255) - 1, 1 << 128,
          (1 << 128) - 1, N // 2, N // 2 + 1, LAMBDA, N - LAMBDA]
    # k = a + lambda b for halves at the 2^128 edge (the split reduces them
    # against its own lattice basis, so these stress the output bound)
    for a in (0, 1, (1 << 128) - 1, 1 << 127, (1 << 128) - 12345):
        for b in (0, 1, (1 << 128) - 1, 1 << 127, N - ((1 << 128) - 1)):
            xs.append((a + LAMBDA * b) % N)
    xs += [rnd.getrandbits(256) for _ in range(400)]
    xs += [N + rnd.getrandbits(128) for _ in range(50)]
    return xs


def test_split_reconstructs_and_halves_below_2_128():
    rnd = random.Random(7)
    for k in adversarial_scalars() + [rnd.randrange(N) for _ in range(3000)]:
        if k >= N:
            continue
        k1, k2, n1, n2 = glv_split_dev(k)
        a = N - k1 if n1 else k1
        b = N - k2 if n2 else k2
        assert (a + LAMBDA * b) % N == k
        assert k1 < (1 << 128) and k2 < (1 << 128), hex(k)


@pytest.mark.parametrize("c", range(2, 21))
def test_digits_recompose_every_window_width(c):
    rnd = random.Random(c)
    vals = [0, 1, (1 << 128) - 1, 1 << 127, int("AA" * 16, 16), int("55" * 16, 16)]
    vals += [rnd.getrandbits(128) for _ in range(200)]
    W, B = plan(c)
    for v in vals:
        ds = digits(v, c)
        assert sum(d << (c * w) for w, d in enumerate(ds)) == v
        assert all(-B < d <= B for d in ds[:-1]) and 0 <= ds[-1] <= 2 * B


@pytest.mark.parametrize("c", range(2, 21))
def test_bucket_ids_in_range_for_any_input(c):
    """Split terms of adversarial scalars, and (without the masking) raw
    256-bit values: every id < (W + 1) B, every skip id inside
    its own window, every term index < 2N."""
    W, B = plan(c)
    nbk = (W + 1) * B
    scalars = adversarial_scalars()
    n2 = 2 * len(scalars)
    for t, k in enumerate(scalars):
        for h, half in enumerate(prepare(k)):
            i = 2 * t + h
            assert i < n2
            for w, (bid, skip) in enumerate(entries(half, c, i)):
                assert bid < nbk, (hex(k), c, w)
                assert w * B <= bid < (w + 2 if w == W - 1 else w + 1) * B
    rnd = random.Random(100 + c)
    for v in [M256, 1 << 255, (1 << 256) - (1 << 128)] + [rnd.getrandbits(256) for _ in range(64)]:
        for bid, _ in entries(v, c, 5):
            assert bid < nbk

"""Pippenger MSM on the GPU vs the reference CPU ``ufsecp_multi_scalar_mul``.

Small cases and error precedence come from tests/golden/edge_cases.json (the
reference CPU ABI's outputs and codes).  Larger MSMs use bases k_i*G, whose
result has the closed form (sum s_i k_i mod n) * G.
"""
from pathlib import Path

import numpy as np
import pytest

from ec_math import N, compress, mul
from ultrafastsecp256k1_amd import UFSECP_ERR_ARITH, UFSECP_OK
from ultrafastsecp256k1_amd import workload as wl

pytestmark = pytest.mark.gpu


def test_msm_golden_cases(gpu, golden):
    for c in golden["edge_cases"]["msm"]:
        rc, out = gpu.msm_rc(bytes.fromhex(c["scalars"]), bytes.fromhex(c["points"]), c["n"])
        assert rc == c["rc"], c["name"]
        if rc == UFSECP_OK:
            assert out.hex() == c["result"], c["name"]
        else:
            assert out == b"\x00" * 33, c["name"]


def test_msm_workload_4096_matches_reference(gpu, golden):
    f = golden["workload_heads"]["msm"]
    sc, pts, _ = wl.generate_msm(f["n"], f["seed"])
    assert gpu.msm(sc, pts).hex() == f["result"]


@pytest.mark.parametrize("n", [1 << 12, 1 << 16])
def test_msm_closed_form(gpu, n):
    sc, pts, ks = wl.generate_msm(n, seed=wl.SEED_MSM + 77)
    e = wl.msm_closed_form_scalar(sc, ks)
    assert gpu.msm(sc, pts) == compress(mul(e))


def test_msm_infinity_is_arith_error(gpu):
    sc, pts, ks = wl.generate_msm(64, seed=5)
    sc2 = sc.copy()
    # s_63 chosen so that the total is 0 mod n
    e = wl.msm_closed_form_scalar(sc[:63], ks[:63])
    k63 = int.from_bytes(bytes(ks[63]), "big")
    s63 = (-e) * pow(k63, -1, N) % N
    sc2[63] = np.frombuffer(s63.to_bytes(32, "big"), np.uint8)
    rc, out = gpu.msm_rc(sc2, pts)
    assert rc == UFSECP_ERR_ARITH and out == b"\x00" * 33


@pytest.mark.parametrize("c", [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20])
def test_msm_every_window_width(gpu, c):
    """Explicit window widths: every c gives the closed-form point, including
    the top window's unsigned digits (128 - c(W-1) bits plus the carry) and
    the small bucket counts of c < 4."""
    n = 777
    sc, pts, ks = wl.generate_msm(n, seed=wl.SEED_MSM + c)
    e = wl.msm_closed_form_scalar(sc, ks)
    rc, out = gpu.msm_rc(sc, pts, n, c)
    assert rc == UFSECP_OK
    assert out == compress(mul(e))


@pytest.mark.parametrize("c", [2, 5, 8, 11, 13, 16, 17, 20])
def test_msm_top_window_extremes(gpu, c):
    """Scalars whose signed recoding carries into the top window up to its
    largest unsigned digit (n - 1, n - 2, 2^255 + 2^254, all-ones low bits),
    beside zero and one: the top window's second bucket group is used."""
    N = wl.N_ORDER
    specials = [N - 1, N - 2, (1 << 255) + (1 << 254), (1 << 256) - (1 << 200) - 1 - ((1 << 256) - N),
                (1 << 128) - 1, 1, 0, N // 2, N // 2 + 1]
    specials = [v % N for v in specials]
    ks = [3 + 7 * i for i in range(len(specials))]
    sc = b"".join(v.to_bytes(32, "big") for v in specials)
    pts = b"".join(compress(mul(k)) for k in ks)
    e = sum(v * k for v, k in zip(specials, ks)) % N
    rc, out = gpu.msm_rc(sc, pts, len(ks), c)
    assert rc == UFSECP_OK
    assert out == compress(mul(e))


@pytest.mark.parametrize("c", [10, 12, 14, 16])
def test_msm_window_widths_at_2p18(gpu, c):
    """The configuration sweep's widths on a size where the top window's
    buckets hold hundreds of thousands of terms."""
    n = 1 << 18
    sc, pts, ks = wl.generate_msm(n, seed=wl.SEED_MSM + 100 + c)
    e = wl.msm_closed_form_scalar(sc, ks)
    rc, out = gpu.msm_rc(sc, pts, n, c)
    assert rc == UFSECP_OK
    assert out == compress(mul(e))


@pytest.mark.parametrize("c", [8, 16])
def test_msm_repeated_bases(gpu, c):
    """Every term on the same base point: equal scalars put every term of a
    window in one bucket, so the bucket sum adds a point to itself (the
    mixed addition asks for a doubling), and scalars s and n - s cancel
    (the sum passes through infinity and continues)."""
    N = wl.N_ORDER
    k = 0x1D3F5A7C9E0B2D4F6A8C0E1F3A5C7E9B1D3F5A7C9E0B2D4F6A8C0E1F3A5C7E9B % N
    base = compress(mul(k))
    s = 0x3C5E7091B3D5F7193B5D7F91B3D5F7193B5D7F91B3D5F7193B5D7F91B3D5F719 % N
    cases = {
        "equal": [s] * 1000,
        "cancel": [s if i % 2 == 0 else N - s for i in range(1001)],
        "mixed": [s, s, N - s, 2 * s % N, s, N - s, N - s] * 97,
    }
    for name, scal in cases.items():
        sc = b"".join(v.to_bytes(32, "big") for v in scal)
        pts = base * len(scal)
        e = sum(scal) * k % N
        rc, out = gpu.msm_rc(sc, pts, len(scal), c)
        assert rc == UFSECP_OK, name
        assert out == compress(mul(e)), name


@pytest.mark.parametrize("c", [11, 13, 15])
def test_msm_odd_window_widths_at_2p16(gpu, c):
    """The sweep's odd widths (a partial top window: 128 - c(W - 1) < c bits)
    at 2^16 terms, against the closed form."""
    n = 1 << 16
    sc, pts, ks = wl.generate_msm(n, seed=wl.SEED_MSM + 200 + c)
    rc, out = gpu.msm_rc(sc, pts, n, c)
    assert rc == UFSECP_OK
    assert out == compress(mul(wl.msm_closed_form_scalar(sc, ks)))


def _msm_affine(gpu, sc, p64, c=0):
    """(status, 33-byte output) of ufsecp_gpu_ext_msm_affine_device."""
    n = len(sc)
    d_s, d_p, d_o = gpu.to_device(sc), gpu.to_device(p64), gpu.empty(64)
    d_o.upload(np.full(64, 0xAB, np.uint8))
    gpu.msm_affine_device(d_s.ptr, d_p.ptr, n, d_o.ptr, d_o.at(48), c)
    gpu.synchronize()
    raw = d_o.download()
    return int(np.frombuffer(raw[48:52].tobytes(), np.int32)[0]), raw[:33].tobytes()


@pytest.mark.parametrize("n,c", [(5000, 0), (1 << 16, 12), (1 << 16, 16)])
def test_msm_affine_bases_match_compressed(gpu, n, c):
    """Pre-decompressed bases (64-byte x || y): the same point as the
    compressed path and the closed form."""
    sc, pts, ks, p64 = wl.generate_msm(n, seed=wl.SEED_MSM + 300 + c, affine=True)
    want = compress(mul(wl.msm_closed_form_scalar(sc, ks)))
    assert _msm_affine(gpu, sc, p64, c) == (0, want)
    assert gpu.msm_rc(sc, pts, n, c) == (UFSECP_OK, want)


def test_msm_affine_bases_invalid(gpu):
    """x >= p, y >= p, a point off the curve -> BAD_PUBKEY (3); a scalar >= n
    -> BAD_INPUT (5); the first failing term decides; output zeroed."""
    from ec_math import P
    sc, _, _, p64 = wl.generate_msm(300, seed=11, affine=True)
    for mut in ("x_ge_p", "y_ge_p", "off_curve", "y_neg_ok"):
        q = p64.copy()
        x = int.from_bytes(bytes(q[77, :32]), "big")
        y = int.from_bytes(bytes(q[77, 32:]), "big")
        if mut == "x_ge_p":
            q[77, :32] = np.frombuffer((x + P).to_bytes(32, "big") if x + P < 1 << 256 else P.to_bytes(32, "big"),
                                       np.uint8)
        elif mut == "y_ge_p":
            q[77, 32:] = np.frombuffer((y + P).to_bytes(32, "big") if y + P < 1 << 256 else P.to_bytes(32, "big"),
                                       np.uint8)
        elif mut == "off_curve":
            q[77, 32:] = np.frombuffer(((y + 1) % P).to_bytes(32, "big"), np.uint8)
        else:                       # -P is on the curve: valid, result changes
            q[77, 32:] = np.frombuffer((P - y).to_bytes(32, "big"), np.uint8)
        st, out = _msm_affine(gpu, sc, q)
        if mut == "y_neg_ok":
            assert st == 0
        else:
            assert (st, out) == (3, bytes(33)), mut
    s2 = sc.copy()
    s2[200] = 0xFF          # scalar >= n at term 200, bad point at term 77: the earlier term decides
    q = p64.copy()
    q[77, 63] ^= 1
    assert _msm_affine(gpu, s2, q)[0] == 3
    assert _msm_affine(gpu, s2, p64)[0] == 5


@pytest.mark.parametrize("n,c", [(1 << 16, 16), (1 << 16, 8), (200_000, 13)])
def test_msm_one_scalar_many_bases(gpu, n, c):
    """Every term with the same scalar (and a second run with a quarter of the
    terms sharing one scalar): whole windows fall into one or two buckets, so
    the bucket partials span far more chunks than the join takes and go
    through several level rounds.  Against the closed form s * sum k_i (+ the
    rest)."""
    N = wl.N_ORDER
    sc, pts, ks = wl.generate_msm(n, seed=wl.SEED_MSM + 400 + c, threads=16)
    s = int.from_bytes(bytes(sc[0]), "big")
    same = np.broadcast_to(sc[0], sc.shape).copy()
    assert gpu.msm_rc(same, pts, n, c) == (UFSECP_OK, compress(mul(s * wl.msm_closed_form_scalar(
        np.broadcast_to(np.frombuffer((1).to_bytes(32, "big"), np.uint8), sc.shape), ks) % N)))
    mixed = sc.copy()
    mixed[::4] = sc[0]
    assert gpu.msm_rc(mixed, pts, n, c) == (UFSECP_OK, compress(mul(wl.msm_closed_form_scalar(mixed, ks))))


def test_msm_two_stream_form(tmp_path, monkeypatch, golden):
    """The split form behind the tuning key msm_split_from (off by default):
    the high windows' sums and their Horner share run on the engine's second
    stream beside the low windows' bucket phase, and the last kernel adds the
    share.  Enabled here from 2^10 terms, against the closed form at several
    window widths (odd window counts included), the golden cases, the
    infinity result and back-to-back calls on one context."""
    from ultrafastsecp256k1_amd import UfsecpGpu
    tuned = Path(__file__).resolve().parents[1] / "ultrafastsecp256k1_amd" / "ufsecp_tuning.txt"
    f = tmp_path / "split_tuning.txt"
    f.write_text(tuned.read_text() + "\nmsm_split_from 10\n")
    monkeypatch.setenv("UFSECP_TUNING_FILE", str(f))
    with UfsecpGpu(0) as g:
        for n, c in ((1 << 10, 0), (5000, 13), (3000, 7), (1 << 16, 16), (1 << 16, 15), (1 << 12, 18)):
            sc, pts, ks = wl.generate_msm(n, seed=wl.SEED_MSM + 900 + n + c)
            assert g.msm_rc(sc, pts, n, c) == (UFSECP_OK, compress(mul(wl.msm_closed_form_scalar(sc, ks)))), (n, c)
        for case in golden["edge_cases"]["msm"]:
            rc, out = g.msm_rc(bytes.fromhex(case["scalars"]), bytes.fromhex(case["points"]), case["n"])
            assert rc == case["rc"], case["name"]
        sc, pts, ks = wl.generate_msm(2048, seed=11)
        e = wl.msm_closed_form_scalar(sc[:2047], ks[:2047])
        k_last = int.from_bytes(bytes(ks[2047]), "big")
        sc[2047] = np.frombuffer(((-e) * pow(k_last, -1, N) % N).to_bytes(32, "big"), np.uint8)
        assert g.msm_rc(sc, pts, 2048, 11) == (UFSECP_ERR_ARITH, b"\x00" * 33)


@pytest.mark.parametrize("c", [0, 16])
def test_msm_invalid_points_in_a_large_batch(gpu, c):
    """An invalid base among 2^16 terms: its term's bucket entries still go
    through the accumulation (the bases are checked beside the scalar side,
    not before it), and the call returns BAD_PUBKEY (3) with a zeroed output;
    a scalar >= n at an earlier term returns BAD_INPUT (5) instead.  A valid
    MSM on the same context afterwards is unaffected."""
    from ec_math import P
    n = 1 << 16
    sc, pts, ks = wl.generate_msm(n, seed=wl.SEED_MSM + 700 + c, threads=16)
    x_ge_p = np.frombuffer(b"\x02" + (P + 5).to_bytes(32, "big"), np.uint8)
    off_curve = None
    for x in range(1, 100):                   # x^3 + 7 not a square: no point with this x
        if pow(x ** 3 + 7, (P - 1) // 2, P) != 1:
            off_curve = np.frombuffer(b"\x03" + x.to_bytes(32, "big"), np.uint8)
            break
    bad_prefix = pts[n // 3].copy()
    bad_prefix[0] = 0x04
    for pos, bad in ((n // 2, x_ge_p), (n - 1, off_curve), (n // 3, bad_prefix), (0, x_ge_p)):
        q = pts.copy()
        q[pos] = bad
        assert gpu.msm_rc(sc, q, n, c) == (3, b"\x00" * 33), pos
        s2 = sc.copy()
        if pos + 1 < n:
            s2[pos + 1] = 0xFF                          # scalar >= n after the bad point
            assert gpu.msm_rc(s2, q, n, c)[0] == 3, pos
        s2 = sc.copy()
        s2[pos] = 0xFF                                  # at the same term: the scalar is checked first
        assert gpu.msm_rc(s2, q, n, c)[0] == 5, pos
        if pos > 0:
            s3 = sc.copy()
            s3[pos - 1] = 0xFF                          # before it
            assert gpu.msm_rc(s3, q, n, c)[0] == 5, pos
    assert gpu.msm_rc(sc, pts, n, c) == (UFSECP_OK, compress(mul(wl.msm_closed_form_scalar(sc, ks))))


def test_msm_random_shapes(gpu):
    """Seeded random MSMs: sizes log-uniform in [1, 2^15], explicit and tuned
    window widths, and scalar patterns that stress the bucket stages
    (duplicated scalars, half zero, small scalars, one scalar for every term,
    repeated bases), against the closed form (sum s_i k_i) G."""
    import random
    rnd = random.Random(20261204)
    base_sc, base_pts, base_ks = wl.generate_msm(1 << 15, seed=4242, threads=16)
    for it in range(40):
        n = max(1, min(int(round(2 ** rnd.uniform(0, 15))), 1 << 15))
        c = rnd.choice([0, 0, 0, 2, 3, 5, 8, 11, 13, 16, 18, 20])
        off = rnd.randrange(0, (1 << 15) - n + 1)
        sc, pts, ks = base_sc[off:off + n].copy(), base_pts[off:off + n].copy(), base_ks[off:off + n].copy()
        pat = rnd.choice(["uniform", "dup", "zeros", "small", "onebucket", "rep_points"])
        if pat == "dup":
            sc = sc[[rnd.randrange(n) for _ in range(n)]]
        elif pat == "zeros":
            sc[np.array([rnd.random() < 0.5 for _ in range(n)])] = 0
        elif pat == "small":
            for i in range(n):
                sc[i] = np.frombuffer(rnd.randrange(0, 1 << rnd.choice([1, 8, 20, 40])).to_bytes(32, "big"), np.uint8)
        elif pat == "onebucket":
            sc[:] = sc[0]
        elif pat == "rep_points":
            j = [rnd.randrange(max(1, n // 8)) for _ in range(n)]
            pts, ks = pts[j], ks[j]
        e = wl.msm_closed_form_scalar(sc, ks)
        want = (UFSECP_ERR_ARITH, b"\x00" * 33) if e == 0 else (UFSECP_OK, compress(mul(e)))
        assert gpu.msm_rc(sc, pts, n, c) == want, (it, n, c, pat, off)

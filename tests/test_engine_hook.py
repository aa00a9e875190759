"""The reference CPU engine's GPU column hook served by this library.

The reference engine (secp256k1::ecdsa_batch_verify_opaque_columns /
schnorr_batch_verify_bip340_columns, src/cpu/src/batch_verify.cpp:851-960)
asks an installed hook first and falls back to its CPU path when the hook
declines (batch_verify.hpp:149-168).  Each test runs a child process that
loads the reference engine (oracle/_ref, built from the reference sources)
with global symbols, then this library, installs the hook and calls the
engine's column entry points on the opaque-signature fixture.

Without a device the hook declines and the engine's own CPU path answers.
That path reduces r, s >= n once modulo n (Scalar::from_limbs,
src/cpu/src/scalar.cpp:118-125) where the C ABI and the GPU reject them, so
its verdicts differ from the fixture exactly on the s + n twins of valid
signatures; with the device the hook answers and the verdicts are the
fixture's.
"""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = ROOT / "oracle" / "_ref" / "libufsecp_ref.so"

CHILD = r"""
import ctypes, json, sys
sys.path.insert(0, sys.argv[1])
eng = ctypes.CDLL(sys.argv[2], mode=ctypes.RTLD_GLOBAL)
import ultrafastsecp256k1_amd as ufs
lib = ufs.load_library()
installed = lib.ufsecp_gpu_ext_engine_hook_install()
items = json.loads(open(sys.argv[3]).read())["items"]
cat = lambda k: b"".join(bytes.fromhex(c[k]) for c in items)
n = len(items)
out = ctypes.create_string_buffer(n)
f = eng._ZN9secp256k133ecdsa_batch_verify_opaque_columnsEPKhS1_S1_mPhm
f.restype = ctypes.c_bool
f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
all_ok = f(cat("msg"), cat("pub"), cat("sig"), n, ctypes.cast(out, ctypes.c_void_p), 1)
ecdsa = list(out.raw)
bip = json.loads(open(sys.argv[4]).read())
cb = lambda k: b"".join(bytes.fromhex(c[k]) for c in bip)
g = eng._ZN9secp256k135schnorr_batch_verify_bip340_columnsEPKhS1_S1_mPhm
g.restype = ctypes.c_bool
g.argtypes = f.argtypes
out2 = ctypes.create_string_buffer(len(bip))
g(cb("msg"), cb("pubkey"), cb("sig"), len(bip), ctypes.cast(out2, ctypes.c_void_p), 1)
print(json.dumps({"installed": installed, "all_ok": bool(all_ok), "ecdsa": ecdsa, "schnorr": list(out2.raw),
                  "stats": ufs.engine_hook_stats(lib)}))
"""


def _run_child():
    r = subprocess.run([sys.executable, "-c", CHILD, str(ROOT), str(REF),
                        str(ROOT / "tests" / "golden" / "opaque_vectors.json"),
                        str(ROOT / "tests" / "golden" / "bip340_vectors.json")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def _fixture():
    items = json.loads((ROOT / "tests" / "golden" / "opaque_vectors.json").read_text())["items"]
    bip = json.loads((ROOT / "tests" / "golden" / "bip340_vectors.json").read_text())
    return items, bip


@pytest.mark.skipif(not REF.is_file(), reason="reference CPU library (oracle/_ref) not built")
def test_engine_hook_declines_without_device():
    import ultrafastsecp256k1_amd as ufs
    if ufs.load_library().ufsecp_gpu_device_count(4) > 0:
        pytest.skip("a device is present: the hook answers instead of declining")
    res = _run_child()
    items, bip = _fixture()
    assert res["installed"] == 1
    assert res["stats"]["calls"] == 2 and res["stats"]["declined"] == 2 and res["stats"]["handled"] == 0
    # the s + n twins of valid small-s signatures (r + n overflows 256 bits for
    # the fixture's r, so its r_plus_n item is not a twin)
    reduced = {c["name"] for c in items if c["name"].startswith("small_s_") and c["name"].endswith("_plus_n")}
    assert len(reduced) == 3
    for c, v in zip(items, res["ecdsa"]):
        assert v == (1 if c["name"] in reduced else c["verdict"]), c["name"]
    assert res["schnorr"] == [c["verdict"] for c in bip]


@pytest.mark.gpu
@pytest.mark.skipif(not REF.is_file(), reason="reference CPU library (oracle/_ref) not built")
def test_engine_hook_on_device():
    res = _run_child()
    items, bip = _fixture()
    assert res["installed"] == 1
    assert res["stats"]["handled"] == 2 and res["stats"]["declined"] == 0
    assert res["stats"]["items"] == len(items) + len(bip)
    assert res["ecdsa"] == [c["verdict"] for c in items]
    assert res["schnorr"] == [c["verdict"] for c in bip]
    assert res["all_ok"] is False

"""Build the in-tree native libraries.

* ``libufsecp_gpu.so``      -- the HIP/gfx950 batch-verify + MSM library
                               behind the ufsecp GPU C ABI (hipcc).
* ``libufsecp_workload.so`` -- CPU synthetic-input generator (keys, RFC6979
                               ECDSA and BIP-340 signatures, MSM terms) used by
                               the benchmark and the serving driver (g++).

Run ``python -m ultrafastsecp256k1_amd.build`` (or ``__graft_entry__.build()``).
Object files go to ``ultrafastsecp256k1_amd/_build/``; the shared libraries
are written next to this file so they travel with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
GPU_LIB = PKG / "libufsecp_gpu.so"
WORKLOAD_LIB = PKG / "libufsecp_workload.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("UFSECP_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
GPU_SOURCES = ["kernels_verify.hip", "kernels_pubdata.hip", "kernels_recover.hip",
               "kernels_rlc.hip", "msm.hip", "backend.cpp", "server.cpp", "tuning.cpp", "ufsecp_gpu_abi.cpp",
               "engine_hook.cpp", "batch_rlc.cpp", "ufsecp_gpu_unsupported.cpp"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "feature flag" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{out}")


def _compile_hip(src: str) -> Path:
    obj = BUILD / (src + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC, *HIP_FLAGS, *lang, "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def build_gpu(jobs: int = 4) -> Path:
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, GPU_SOURCES))
    tmp = GPU_LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


VALU_PEAK = PKG / "valu_peak"


def build_valu_peak() -> Path:
    """Integer-VALU throughput probe (measured peak for roofline fractions)."""
    tmp = VALU_PEAK.with_suffix(".tmp")
    _run([HIPCC, *HIP_FLAGS, str(CSRC / "valu_peak.hip"), "-o", str(tmp)])
    os.replace(tmp, VALU_PEAK)
    return VALU_PEAK


def build_workload() -> Path:
    BUILD.mkdir(exist_ok=True)
    gxx = shutil.which("g++") or "g++"
    tmp = WORKLOAD_LIB.with_suffix(".so.tmp")
    _run([gxx, "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-fvisibility=hidden",
          str(CSRC / "workload.cpp"), "-o", str(tmp)])
    os.replace(tmp, WORKLOAD_LIB)
    return WORKLOAD_LIB


SERVE_BENCH = PKG / "serve_bench"


def build_serve_bench() -> Path:
    """Serving benchmark executable over the two libraries (built first)."""
    gxx = shutil.which("g++") or "g++"
    tmp = SERVE_BENCH.with_suffix(".tmp")
    _run([gxx, "-O2", "-std=c++17", "-pthread", f"-I{ROOT / 'include'}", str(CSRC / "serve_bench.cpp"),
          f"-L{PKG}", "-lufsecp_gpu", "-lufsecp_workload", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(tmp)])
    os.replace(tmp, SERVE_BENCH)
    return SERVE_BENCH


def build_all() -> None:
    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_gpu), ex.submit(build_workload), ex.submit(build_valu_peak)]
        for f in futs:
            f.result()
    build_serve_bench()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"gpu": build_gpu, "workload": build_workload, "valu": build_valu_peak, "serve": build_serve_bench,
     "all": build_all}[what]()
    print("built", what)

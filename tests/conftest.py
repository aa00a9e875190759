"""Shared test configuration.

Tests marked ``gpu`` need an MI355X (they call the HIP library through the C
ABI); everything else runs on the CPU.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (run with -m gpu on the MI355X box)")


@pytest.fixture(scope="session")
def gpu():
    from ultrafastsecp256k1_amd import UfsecpGpu
    ctx = UfsecpGpu()
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def golden():
    import json
    return {name: json.loads((GOLDEN / f"{name}.json").read_text())
            for name in ("workload_heads", "bip340_vectors", "rfc6979_vectors", "edge_cases", "opaque_vectors",
                         "pubdata_vectors", "recover_vectors", "full_size_digests")}

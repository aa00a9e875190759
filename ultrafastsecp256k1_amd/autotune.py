"""Autotuner for the library's launch parameters on the GPU it runs on.

Measures, on synthetic inputs of the benchmark workloads, and writes the
winners to a tuning file that every context reads at creation
(``ufsecp_tuning.txt`` beside the library, or ``UFSECP_TUNING_FILE``):

* ``gcomb_w``: width of the fixed-base comb table for G (more width = fewer
  additions per signature, a larger table in HBM/MALL), by BIP-340 verify
  throughput;
* ``verify_blocks_per_cu``: resident blocks per CU of the persistent verify
  grid, by throughput of the uneven multi-batch mix (10,000-batch plan,
  first batches) and of one 2^20 batch;
* ``msm_window <log2 n> <c>``: the Pippenger window for each size band;
* ``server_groups``: launch groups in flight in the serving loop: among the
  settings within 2% of the best saturated throughput of log-uniform
  requests, the one with the lowest p99 latency at 80% of that best rate
  (more groups in flight make each launch group smaller at a given load).

    python -m ultrafastsecp256k1_amd.autotune [--out PATH] [--quick]
                                             [--only comb,grid,msm,server]

With ``--only`` the other parameters are kept from the existing file.
"""
from __future__ import annotations

import argparse
import ast
import math
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
DEFAULT_OUT = PKG / "ufsecp_tuning.txt"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _ctx_with(values: dict):
    """A fresh context reading a tuning file that holds exactly `values`."""
    from . import UfsecpGpu
    fd, path = tempfile.mkstemp(suffix=".txt", prefix="ufsecp_tune_")
    with os.fdopen(fd, "w") as f:
        for k, v in values.items():
            if isinstance(v, dict):
                for a, b in v.items():
                    f.write(f"{k} {a} {b}\n")
            else:
                f.write(f"{k} {v}\n")
    old = os.environ.get("UFSECP_TUNING_FILE")
    os.environ["UFSECP_TUNING_FILE"] = path
    try:
        return UfsecpGpu()
    finally:
        if old is None:
            os.environ.pop("UFSECP_TUNING_FILE", None)
        else:
            os.environ["UFSECP_TUNING_FILE"] = old
        os.unlink(path)


def _device_batch(gpu, b):
    d = [gpu.to_device(b.msgs), gpu.to_device(b.pubkeys), gpu.to_device(b.sigs), gpu.empty(b.n)]
    return d


def _time_verify(gpu, d, n, kind, reps=3) -> float:
    """Best-of-reps kernel time (ms) of one device-resident batch."""
    fn = gpu.schnorr_verify_device if kind == "schnorr" else gpu.ecdsa_verify_device
    fn(d[0].ptr, d[1].ptr, d[2].ptr, n, d[3].ptr)
    gpu.synchronize()
    best = float("inf")
    for _ in range(reps):
        e0, e1 = gpu.event(), gpu.event()
        e0.record()
        fn(d[0].ptr, d[1].ptr, d[2].ptr, n, d[3].ptr)
        e1.record()
        best = min(best, e0.elapsed_ms(e1))
    return best


def _smallest_within(res: dict, tol: float):
    """The smallest setting whose throughput is within `tol` of the best: a
    wider comb table must win by more than measurement noise to be chosen."""
    best = max(res.values())
    return min(k for k, v in res.items() if v >= best * (1 - tol))


def tune_comb(tuned: dict, n: int, widths, passes: int = 2) -> None:
    """Each width is measured in `passes` interleaved passes (best kept), so a
    slow drift of the clock does not favour the widths measured first; a wider
    table must beat the best by more than 0.5% to be chosen."""
    from . import workload as wl
    b = wl.generate("schnorr", n)
    res = {}
    for _ in range(passes):
        for w in widths:
            gpu = _ctx_with({**tuned, "gcomb_w": w})
            d = _device_batch(gpu, b)
            ms = _time_verify(gpu, d, n, "schnorr", reps=5)
            res[w] = max(res.get(w, 0.0), n / (ms * 1e-3))
            log(f"[tune] gcomb_w {w}: {n / (ms * 1e-3) / 1e6:.2f} M sigs/s")
            for x in d:
                x.free()
            gpu.close()
    tuned["gcomb_w"] = _smallest_within(res, 0.005)
    tuned.setdefault("_measured", {})["gcomb_w"] = {str(k): round(v) for k, v in res.items()}


def tune_grid(tuned: dict, n: int, nbatches: int) -> None:
    from . import KIND_ECDSA, workload as wl
    b = wl.generate("schnorr", n)
    plan = wl.mix_plan(nbatches, 4096, wl.SEED_MIX)
    pools, layout = wl.generate_mix(plan)
    res = {}
    for bpc in (1, 2, 3, 0):
        gpu = _ctx_with({**tuned, "verify_blocks_per_cu": bpc})
        d = _device_batch(gpu, b)
        ms1 = _time_verify(gpu, d, n, "schnorr")
        dev = [(gpu.to_device(p.msgs), gpu.to_device(p.pubkeys), gpu.to_device(p.sigs), gpu.empty(p.n))
               for p in pools]
        rows = []
        for kind, first, size in layout:
            m, pk, s, o = dev[kind]
            stride = 33 if kind == KIND_ECDSA else 32
            rows.append((kind, size, m.at(32 * first), pk.at(stride * first), s.at(64 * first), o.at(first)))
        table = gpu.batch_table(rows)
        gpu.verify_multi_device(table, len(rows), use_graph=True)
        gpu.synchronize()
        best = float("inf")
        for _ in range(3):
            e0, e1 = gpu.event(), gpu.event()
            e0.record()
            gpu.verify_multi_device(table, len(rows), use_graph=True)
            e1.record()
            best = min(best, e0.elapsed_ms(e1))
        items = sum(size for _, _, size in layout)
        res[bpc] = (n / (ms1 * 1e-3), items / (best * 1e-3))
        log(f"[tune] verify_blocks_per_cu {bpc}: batch {res[bpc][0] / 1e6:.2f} M/s, "
            f"mix {res[bpc][1] / 1e6:.2f} M/s")
        for x in d:
            x.free()
        for t in dev:
            for x in t:
                x.free()
        gpu.close()
    score = {k: v[0] * v[1] for k, v in res.items()}
    tuned["verify_blocks_per_cu"] = max(score, key=score.get)
    tuned.setdefault("_measured", {})["verify_blocks_per_cu"] = {
        str(k): [round(v[0]), round(v[1])] for k, v in res.items()}


def tune_msm(tuned: dict, sizes) -> None:
    """Device-resident MSM (inputs already in HBM), best of three HIP-event
    timings per window width."""
    from . import workload as wl
    gpu = _ctx_with(tuned)
    big = max(sizes)
    sc, pts, _ = wl.generate_msm(1 << big)
    d_sc, d_pt = gpu.to_device(sc), gpu.to_device(pts)
    d_out = gpu.empty(64)
    table = {}
    meas = {}
    for lg in sizes:
        n = 1 << lg
        cands = range(max(2, lg - 9), min(18, lg - 1) + 1)
        best_c, best_ms = None, float("inf")
        for c in cands:
            gpu.msm_device(d_sc.ptr, d_pt.ptr, n, d_out.ptr, d_out.at(48), c)   # warm-up (buffers)
            gpu.synchronize()
            if int(d_out.download("int32")[12]) != 0:
                raise RuntimeError(f"msm failed n={n} c={c}")
            t = float("inf")
            for _ in range(3):
                e0, e1 = gpu.event(), gpu.event()
                e0.record()
                gpu.msm_device(d_sc.ptr, d_pt.ptr, n, d_out.ptr, d_out.at(48), c)
                e1.record()
                t = min(t, e0.elapsed_ms(e1))
            meas.setdefault(str(lg), {})[str(c)] = round(t, 3)
            if t < best_ms:
                best_c, best_ms = c, t
        table[lg] = best_c
        log(f"[tune] msm 2^{lg}: c = {best_c} ({best_ms:.2f} ms)")
    # fill the bands in between with the nearest measured band
    full = {}
    meas_lgs = sorted(table)
    for lg in range(0, 32):
        near = min(meas_lgs, key=lambda m: (abs(m - lg), m))
        full[lg] = table[near]
    tuned["msm_window"] = full
    tuned.setdefault("_measured", {})["msm_ms"] = meas
    for x in (d_sc, d_pt, d_out):
        x.free()
    gpu.close()


def _serve_bench(tuned: dict, groups: int, ecdsa_items: int, schnorr_items: int, extra) -> dict:
    import json
    import subprocess
    fd, path = tempfile.mkstemp(suffix=".txt", prefix="ufsecp_tune_")
    os.close(fd)
    write({k: v for k, v in tuned.items() if k != "_measured"}, Path(path), "candidate")
    env = dict(os.environ, UFSECP_TUNING_FILE=path)
    try:
        r = subprocess.run([str(PKG / "serve_bench"), "--groups", str(groups), "--ecdsa-items", str(ecdsa_items),
                            "--schnorr-items", str(schnorr_items), *extra],
                           capture_output=True, text=True, timeout=600, env=env)
    finally:
        os.unlink(path)
    if r.returncode != 0:
        raise RuntimeError(f"serve_bench failed: {r.stderr[-1000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])["batched"]


def tune_server(tuned: dict, ecdsa_items: int, schnorr_items: int) -> None:
    """Serving loop (serve_bench: log-uniform requests, zero-copy submission
    from a C++ client): saturated throughput for 1..4 groups in flight, then
    p99 latency at 80% of the best of those rates for the settings within 2%
    of it; the lowest p99 wins."""
    peak = {}
    for groups in (1, 2, 3, 4):
        peak[groups] = _serve_bench(tuned, groups, ecdsa_items, schnorr_items, ["--peak-only"])["peak_sigs_per_s"]
        log(f"[tune] server_groups {groups}: {peak[groups] / 1e6:.2f} M sigs/s saturated")
    best = max(peak.values())
    p99 = {}
    for groups in sorted(g for g, v in peak.items() if v >= best * 0.98):
        run = _serve_bench(tuned, groups, ecdsa_items, schnorr_items,
                           ["--no-plain", "--loads", "0.8", "--seconds", "2", "--load-rate", str(best)])
        p99[groups] = run["loads"][0]["p99_ms"]
        log(f"[tune] server_groups {groups}: p99 {p99[groups]:.2f} ms at {0.8 * best / 1e6:.1f} M sigs/s")
    tuned["server_groups"] = min(p99, key=lambda g: (p99[g], g))
    tuned.setdefault("_measured", {})["server_groups"] = {str(k): round(v) for k, v in peak.items()}
    tuned["_measured"]["server_p99_ms_at_80pct"] = {str(k): v for k, v in p99.items()}


def read(path: Path) -> dict:
    """The parameters (and the measured record) of an existing tuning file."""
    tuned: dict = {}
    for line in path.read_text().splitlines():
        if line.startswith("# measured: "):
            tuned["_measured"] = ast.literal_eval(line[len("# measured: "):])
            continue
        f = line.split()
        if not f or f[0].startswith("#"):
            continue
        if f[0] == "msm_window":
            tuned.setdefault("msm_window", {})[int(f[1])] = int(f[2])
        else:
            tuned[f[0]] = int(f[1])
    return tuned


def write(tuned: dict, path: Path, device: str) -> None:
    meas = tuned.get("_measured", {})
    lines = [f"# ufsecp MI355X launch parameters, written by ultrafastsecp256k1_amd/autotune.py",
             f"# device: {device}",
             f"# measured: {meas}"]
    for k in ("gcomb_w", "verify_blocks_per_cu", "server_groups"):
        if k in tuned:
            lines.append(f"{k} {tuned[k]}")
    for lg, c in sorted(tuned.get("msm_window", {}).items()):
        lines.append(f"msm_window {lg} {c}")
    path.write_text("\n".join(lines) + "\n")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(DEFAULT_OUT))
    ap.add_argument("--quick", action="store_true", help="smaller measurement sizes")
    ap.add_argument("--only", default="comb,grid,msm,server",
                    help="comma list of comb, grid, msm, server; the rest is kept from --out")
    args = ap.parse_args()
    only = set(args.only.split(","))
    if not only <= {"comb", "grid", "msm", "server"}:
        ap.error(f"unknown --only item in {args.only}")
    from . import UfsecpGpu
    g = UfsecpGpu()
    dev = g.lib.ufsecp_gpu_backend_name(4).decode()
    g.close()
    out = Path(args.out)
    tuned: dict = read(out) if only != {"comb", "grid", "msm", "server"} and out.exists() else {}
    n = 1 << (17 if args.quick else 20)
    t0 = time.perf_counter()
    if "comb" in only:
        tune_comb(tuned, n, (14, 16, 18, 19, 20))
    if "grid" in only:
        tune_grid(tuned, n, 500 if args.quick else 2000)
    if "msm" in only:
        tune_msm(tuned, (8, 10, 12, 14, 16) if args.quick else (8, 10, 12, 14, 16, 18, 20))
    if "server" in only:
        tune_server(tuned, *((1 << 19, 1 << 18) if args.quick else (1 << 21, 1 << 20)))
    write(tuned, Path(args.out), dev)
    log(f"[tune] wrote {args.out} in {time.perf_counter() - t0:.1f} s")
    print(Path(args.out).read_text())
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""STRIDE decode benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one decode of the whole workload (one kernel launch for a single
mesh).  The default workload is the 1900 x 1900 heightfield (7,212,402
triangles) with 12-bit bbox quantisation, encoded by this package's
STRIDEEncoder; inputs are resident in HBM before timing starts.  Prints one
JSON line (rank 0) with the decode throughput in Gtris/s, the dominant
kernel's HBM roofline figures and a CPU baseline (the C oracle, one thread,
timed on this host).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "decode Gtris/s + HBM GB/s (batched vs per-mesh); learned-encode speedup @≤1% bpv"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def bbox_error(bits: int) -> float:
    """bbox_frac precision giving exactly `bits` bits on the longest axis."""
    return 1.0 / ((1 << bits) - 1.5)


def encode_mesh(V, T, bits):
    from meshpress_amd.encoder import STRIDEEncoder
    from meshpress_amd.utils.types import Model
    enc = STRIDEEncoder(max_verts=256, max_tris=256, precision_error=bbox_error(bits),
                        precision_mode="bbox_frac")
    t0 = time.perf_counter()
    prep = enc.prepare(Model.from_arrays(V, T))
    t1 = time.perf_counter()
    data = enc.encode_prepared(prep)
    t2 = time.perf_counter()
    return data, prep, {"prep_s": t1 - t0, "search_write_s": t2 - t1, **enc.last_stats}


def time_decoder(dec, steps, warmup):
    import torch
    for _ in range(warmup):
        dec.decode()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(steps):
        dec.decode()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall / steps, ev0.elapsed_time(ev1) / 1e3 / steps


def cpu_baseline(data, n_t, budget_s=10.0):
    from tests.oracle_c import load
    co = load()
    reps, t_total = 0, 0.0
    while t_total < budget_s and reps < 20:
        t0 = time.perf_counter()
        co.decode(data, n_threads=1)
        t_total += time.perf_counter() - t0
        reps += 1
    return n_t / (t_total / reps) / 1e9, reps, t_total


def check_parity(dec, data):
    from tests.oracle_c import load
    V, T = dec.decode_to_host()
    Vo, To = load().decode(data, n_threads=min(16, os.cpu_count() or 1))
    ok = np.array_equal(T.astype(np.int64), To) and np.array_equal(V.view(np.uint32), Vo.view(np.uint32))
    if not ok:
        raise SystemExit("bench workload decode does not match the oracle")
    return ok


def read_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary."""
    p = ROOT / "profiles" / f"{name}_traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="heightfield_1900_q12")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration)")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        n = a.copy_calibration << 20
        x = torch.ones(n // 4, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "read": n, "written": n}), flush=True)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(decode shards by meshlet ranges with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        sys.exit("bench.py needs a ROCm GPU")
    torch.cuda.set_device(0)
    from meshpress_amd import workloads
    from meshpress_amd.utils.paradelta_v5_cuda import ParaDeltaV5GpuDecoder

    name = a.workload
    if name.startswith("heightfield_1900"):
        bits = 16 if name.endswith("q16") else 12
        log("generating 1900x1900 heightfield")
        V, T = workloads.heightfield_1900()
    elif name.startswith("icosphere_l8"):
        bits = 12
        log("generating icosphere level 8")
        V, T = workloads.icosphere_l8()
    else:
        sys.exit(f"unknown workload {name}")
    log(f"encoding {len(T):,} tris at {bits}-bit")
    data, prep, enc_stats = encode_mesh(V, T, bits)
    log(f"encoded: {len(data):,} bytes, {prep['n_meshlets']:,} meshlets, g_bits {list(prep['g_bits'])}, "
        f"prep {enc_stats['prep_s']:.1f}s search+write {enc_stats['search_write_s']:.1f}s")
    dec = ParaDeltaV5GpuDecoder(data)
    check_parity(dec, data)
    log("parity vs oracle: bit-exact")
    s_wall, s_gpu = time_decoder(dec, a.steps, a.warmup)
    n_t = dec.n_t_total
    bytes_alg = dec.bytes_read + dec.bytes_written
    gtris = n_t / s_wall / 1e9
    achieved = bytes_alg / s_gpu / 1e9
    log(f"decode: {s_gpu * 1e3:.3f} ms/launch (events), {gtris:.3f} Gtris/s, {achieved:.1f} GB/s")
    cpu = None
    if not a.no_cpu:
        v, reps, tt = cpu_baseline(data, n_t)
        cpu = {"value": v, "unit": "Gtris/s", "cores": 1, "kind": "port",
               "sample": f"C oracle decode_dup restatement, full {name} stream x{reps} ({tt:.1f} s)"}
        log(f"cpu baseline: {v:.4f} Gtris/s")
    out = {
        "metric": METRIC, "value": gtris, "unit": "Gtris/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": s_wall * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32+f64", "data": "synthetic",
        "config": {"workload": name, "n_tris": n_t, "n_verts_out": dec.n_v_total,
                   "meshlets": dec.n_meshlets, "g_bits": [int(x) for x in prep["g_bits"]],
                   "stream_bytes": dec.stream_bytes, "side_bytes": dec.side_bytes,
                   "tune": dec.get_tune()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(name),
                     "algorithmic_bytes": bytes_alg, "kernel_ms": s_gpu * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""STRIDE decode benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--batch N] [--batch-cache NPZ] [--quick] [--no-cpu]

A step is one decode of the whole workload: the token pass and then the axis
pass of one mesh, or for the batch one hipGraph replay of both passes over
every mesh.  The headline workload is the 1900 x 1900 heightfield (7,212,402
triangles) with 12-bit bbox quantisation, encoded by this package's
STRIDEEncoder; inputs are resident in HBM before timing starts.  The same
run also decodes

* the 16-bit variant of the heightfield and the level-8 icosphere
  (1,310,720 triangles), reported per variant;
* the batched workload: N meshes (default 2,048) with triangle counts
  log-uniform in [1K, 200K] drawn from {displaced UV sphere, displaced torus,
  noise heightfield}, decoded by one hipGraph replay of the batched kernel
  versus a loop of one launch per mesh (equal output, checked);
* the learned encoder proposal against the exhaustive search on the
  headline mesh and on held-out meshes: search-time speedup, bit-rate
  regret, and decode equality of the two streams;
* the headline mesh as a PDLT v5 stream (global boundary table, LIN3
  predictor), decoded by the same class with the float64 chain of
  decode_paradelta_v5 and with the encoder's float32 chain;

Prints one JSON line with the headline throughput in Gtris/s, the decode's
HBM roofline figures (with the VALU issue bound of the committed counter
profile: roofline.issue_model), a CPU baseline (the C oracle on one thread,
and under cpu_baseline.all_threads on the host's thread share, both timed on
this host) and the other workloads under "details".
"""
from __future__ import annotations

import argparse
import functools
import json
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


def time_calls(fn, steps, warmup):
    """(wall seconds per call, event seconds per call) on torch's current stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(steps):
        fn()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall / steps, ev0.elapsed_time(ev1) / 1e3 / steps


def autotune(dec, mode_kw=None):
    from meshpress_amd.autotune import BENCH_CANDIDATES, tune
    best, res = tune(dec, steps=10, cands=BENCH_CANDIDATES, mode_kw=mode_kw)
    return best, [[c["mode"], c["group_size"], c["lds_bytes"], c.get("tok_group", 0),
                   c.get("tok_chains", 0), c["grid"], c["use_queue"], round(ms, 4)] for c, ms in res]


def cpu_baseline(data, n_t, gpu_out, budget_s=10.0, n_threads=1):
    """The C oracle decode on n_threads host threads (meshlets split across
    them) over the full stream; its output also checks the GPU decode that
    was timed."""
    from tests.oracle_c import load
    co = load()
    reps, t_total, out = 0, 0.0, None
    while t_total < budget_s and reps < 20:
        t0 = time.perf_counter()
        out = co.decode(data, n_threads=n_threads)
        t_total += time.perf_counter() - t0
        reps += 1
    V, T = gpu_out
    Vo, To = out
    same = bool(np.array_equal(T.astype(np.int64), To)
                and np.array_equal(V.view(np.uint32), Vo.view(np.uint32)))
    return n_t / (t_total / reps) / 1e9, reps, t_total, same


def python_c1_baseline(budget_s=3.0):
    """The oracle's pure-Python decode_dup restatement (the speed class of the
    reference's pure-Python decoder) on the 65,536-triangle UV sphere (lat-long
    128 x 256, displaced, 12 bits in world mode: STRIDEEncoder(max_verts=256,
    precision_error=0.0005)), checked against the GPU decode of the same
    stream."""
    from meshpress_amd import workloads
    from meshpress_amd.encoder import STRIDEEncoder
    from meshpress_amd.utils.paradelta_v5_cuda import ParaDeltaV5GpuDecoder
    from meshpress_amd.utils.types import Model
    from oracle import stride_oracle as O
    V, T = workloads.uv_sphere_c1()
    data = STRIDEEncoder(max_verts=256, precision_error=0.0005).encode(Model(V, T)).data
    Vg, Tg = ParaDeltaV5GpuDecoder(data).decode_to_host()
    reps, t_total, out = 0, 0.0, None
    while t_total < budget_s and reps < 10:
        t0 = time.perf_counter()
        out = O.decode_dup(data)
        t_total += time.perf_counter() - t0
        reps += 1
    same = bool(np.array_equal(Tg.astype(np.int64), out[1])
                and np.array_equal(Vg.view(np.uint32), out[0].view(np.uint32)))
    return {"value": len(T) / (t_total / reps) / 1e9, "unit": "Gtris/s", "cores": 1,
            "kind": "port", "sample": f"UV sphere 128 x 256 (65,536 tris), x{reps} ({t_total:.1f} s)",
            "gpu_output_matches": same}


def host_threads():
    """The host threads a run may use: OMP_NUM_THREADS when set (the GPU
    machine's share of cores), else the visible cores."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return max(1, os.cpu_count() or 1)


def read_profile(name):
    """The committed rocprofv3 --pmc summary of the workload, or {}."""
    p = ROOT / "profiles" / f"{name}_traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text())
        except Exception:
            return {}
    return {}


def read_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary."""
    return read_profile(name).get("hbm_bytes_per_launch")


MI355X_SIMDS = 256 * 4          # 256 CUs x 4 SIMD units
MI355X_CLOCK_GHZ = 2.4          # peak engine clock


def issue_model(name, kernel_ms):
    """The decode's VALU issue bound from the committed counter profile: every
    wave64 VALU instruction takes a 16-lane SIMD 4 cycles, so
    sum(VALU) * 4 / (SIMDs * clock) is the time at full issue; frac = that
    time / the measured kernel time."""
    valu = read_profile(name).get("valu_insts_per_launch")
    if not valu:
        return None
    n = sum(valu.values())
    bound_ms = n * 4 / MI355X_SIMDS / (MI355X_CLOCK_GHZ * 1e9) * 1e3
    return {"valu_insts": n, "simds": MI355X_SIMDS, "clock_ghz": MI355X_CLOCK_GHZ,
            "bound_ms": bound_ms, "frac": bound_ms / kernel_ms}


def single_mesh(name, V, T, bits, steps, warmup, tune=True):
    from meshpress_amd.utils.paradelta_v5_cuda import ParaDeltaV5GpuDecoder
    log(f"{name}: encoding {len(T):,} tris at {bits}-bit")
    data, prep, enc_stats = encode_mesh(V, T, bits)
    log(f"{name}: {len(data):,} bytes, {prep['n_meshlets']:,} meshlets, g_bits {list(prep['g_bits'])}, "
        f"prep {enc_stats['prep_s']:.1f}s search+write {enc_stats['search_write_s']:.1f}s")
    dec = ParaDeltaV5GpuDecoder(data)
    sweep = None
    if isinstance(tune, dict):          # a given launch shape (profiling runs)
        dec.set_tune(**tune)
        log(f"{name}: tune {tune}")
    elif tune:
        best, sweep = autotune(dec)
        log(f"{name}: tuned {best}")
    s_wall, s_gpu = time_calls(dec.decode, steps, warmup)
    bytes_alg = dec.bytes_read + dec.bytes_written
    r = {"workload": name, "n_tris": dec.n_t_total, "n_verts_out": dec.n_v_total,
         "meshlets": dec.n_meshlets, "g_bits": [int(x) for x in prep["g_bits"]],
         "stream_bytes": dec.stream_bytes, "side_bytes": dec.side_bytes,
         "bpv": 8 * len(data) / prep["n_v"], "tune": dec.get_tune(),
         "ms_wall": s_wall * 1e3, "ms_kernel": s_gpu * 1e3,
         "gtris_s": dec.n_t_total / s_wall / 1e9,
         "hbm_gb_s": bytes_alg / s_gpu / 1e9, "algorithmic_bytes": bytes_alg,
         "pred_n": enc_stats["pred_n"], "pred_K": enc_stats["pred_K"],
         "prep_s": enc_stats["prep_s"], "search_write_s": enc_stats["search_write_s"]}
    r["hbm_frac"] = r["hbm_gb_s"] / HBM_PEAK_GBS
    if sweep is not None:
        r["tune_sweep_ms"] = sweep
    log(f"{name}: {s_gpu * 1e3:.3f} ms/launch, {r['gtris_s']:.3f} Gtris/s, {r['hbm_gb_s']:.1f} GB/s")
    return dec, data, r, prep


# PDLT v5 launch shapes: mode 4 (token pass, then the chain pass over
# groups of group_size meshlets in lds_bytes) with token lanes (tok_chains 0;
# stage_words = the token pass's LDS words, tok_group 64) or token chains
# (about tok_chains a meshlet, tok_group meshlets a wave), or mode 3 (one
# kernel)
PDLT_LAYOUTS = ([dict(mode=4, group_size=gs, lds_bytes=lds, stage_words=sw, tok_chains=0,
                      tok_group=64)
                 for gs, lds, sw in ((16, 24576, 4096), (8, 24576, 4096), (16, 16384, 4096),
                                     (16, 49152, 4096), (16, 24576, 8192), (16, 16384, 8192),
                                     (8, 12288, 8192))]
                + [dict(mode=4, group_size=gs, lds_bytes=lds, tok_chains=kc, tok_group=tg)
                   for kc, tg in ((16, 4), (8, 8))
                   for gs, lds in ((16, 20480), (12, 20480), (16, 24576), (16, 16384), (8, 12288),
                                   (8, 16384))]
                + [dict(mode=3, group_size=16, lds_bytes=lds, stage_words=4096, tok_chains=0,
                        tok_group=64) for lds in (49152, 24576)])


def tune_layout(dec):
    """The fastest PDLT_LAYOUTS entry for a v5 decoder (its own chain: the
    float32 chain's smaller LDS blocks favour other group sizes)."""
    sweep = []
    for t in PDLT_LAYOUTS:
        dec.set_tune(**t)
        _, g = time_calls(dec.decode_world, 10, 3)
        sweep.append(dict(t, ms=round(g * 1e3, 4)))
    best = dict(min(sweep, key=lambda x: x["ms"]))
    best.pop("ms")
    dec.set_tune(**best)
    return best, sweep


def pdlt_v5(prep, name, steps, warmup):
    """The headline mesh as a PDLT v5 stream (encode_from_prepared_v5 over the
    same plans), decoded by the same class: float64 chain (bit for bit
    decode_paradelta_v5) and the encoder's float32 chain."""
    import torch
    from meshpress_amd.encoder.paradelta_v5 import encode_from_prepared_v5
    from meshpress_amd.utils.paradelta_v5_cuda import ParaDeltaV5GpuDecoder
    t0 = time.perf_counter()
    data, info = encode_from_prepared_v5(prep, return_info=True)
    enc_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    dec = ParaDeltaV5GpuDecoder(data)
    create_s = time.perf_counter() - t0
    log(f"{name} v5: {len(data):,} bytes ({8 * len(data) / prep['n_v']:.2f} bpv), "
        f"{dec.n_boundary:,} boundary vertices, encode {enc_s:.1f}s, create {create_s:.2f}s")
    best, sweep = tune_layout(dec)
    s_wall, s_gpu = time_calls(dec.decode_world, steps, warmup)
    v64, t64 = dec.decode_world()
    v64 = v64.clone()
    dec32 = ParaDeltaV5GpuDecoder(data, chain="f32")
    best32, _ = tune_layout(dec32)
    w32, g32 = time_calls(dec32.decode_world, steps, warmup)
    v32, t32 = dec32.decode_world()
    torch.cuda.synchronize()
    step = float(prep["per_coord_err"]) * float(prep["scale"])
    drift = float((v64.double() - v32.double()).abs().max()) / step
    bytes_alg = dec.bytes_read + dec.bytes_written
    r = {"workload": f"{name}_pdlt_v5", "n_tris": dec.n_t_total, "n_verts_out": dec.n_v_total,
         "meshlets": dec.n_meshlets, "n_boundary": dec.n_boundary, "stream_bytes": dec.stream_bytes,
         "side_bytes": dec.side_bytes, "bpv": 8 * len(data) / prep["n_v"],
         "lin3": [float(x) for x in info["lin3"]], "encode_s": enc_s, "create_s": create_s,
         "tune": best, "tune_sweep_ms": sweep,
         "ms_wall": s_wall * 1e3, "ms_kernels": s_gpu * 1e3, "gtris_s": dec.n_t_total / s_wall / 1e9,
         "hbm_gb_s": bytes_alg / s_gpu / 1e9, "algorithmic_bytes": bytes_alg,
         "f32_chain": {"ms_kernels": g32 * 1e3, "gtris_s": dec.n_t_total / w32 / 1e9, "tune": best32,
                       "tris_equal": bool(torch.equal(t64, t32)),
                       "max_f64_f32_drift_steps": drift}}
    log(f"{name} v5: {s_gpu * 1e3:.3f} ms, {r['gtris_s']:.3f} Gtris/s (f32 chain {g32 * 1e3:.3f} ms), "
        f"drift {drift:.4f} steps")
    return r


def batch_blobs(n_meshes, predictor, cache=None):
    """The batch's DPR4 streams: encoded here, or read from (and then kept
    in) the .npz file `cache`, so that a profiled run need not encode."""
    from meshpress_amd import workloads
    from meshpress_amd.encoder.many import encode_many
    key = f"{n_meshes}_{predictor}"
    if cache and os.path.exists(cache):
        with np.load(cache) as z:
            if str(z["key"]) == key and "n_v" in z.files:
                n = z["sizes"]
                flat = z["blobs"]
                ends = np.cumsum(n)
                log(f"batch: {n_meshes} streams read from {cache}")
                return ([flat[e - k:e].tobytes() for e, k in zip(ends, n)],
                        [int(x) for x in z["n_v"]], 0.0)
    specs = workloads.batch_specs(n_meshes, 1000, 200000, 0)
    sources = [functools.partial(workloads.mesh_with_tris, f, n, s) for f, n, s in specs]
    t0 = time.perf_counter()

    def progress(i, n):
        if i % 256 == 0 or i == n:
            log(f"batch: encoded {i}/{n} meshes ({time.perf_counter() - t0:.0f} s)")

    blobs, st = encode_many(sources, progress=progress, precision_error=bbox_error(12),
                            precision_mode="bbox_frac", predictor=predictor)
    t_enc = time.perf_counter() - t0
    n_v = [int(s["n_v"]) for s in st]
    if cache:
        np.savez(cache, key=np.array(key), sizes=np.array([len(b) for b in blobs], np.int64),
                 n_v=np.array(n_v, np.int64), blobs=np.frombuffer(b"".join(blobs), np.uint8))
    return blobs, n_v, t_enc


def batched(n_meshes, steps, warmup, predictor, tune=True, modes=("graph", "launch", "per_mesh"),
            cache=None):
    import torch
    from meshpress_amd.batch import BatchDecoder
    from meshpress_amd.utils.paradelta_v5_cuda import ParaDeltaV5GpuDecoder
    blobs, n_v, t_enc = batch_blobs(n_meshes, predictor, cache)
    bd = BatchDecoder(blobs)
    n_t = bd.n_t_total
    log(f"batch: {n_meshes} meshes, {n_t:,} tris, {bd.n_meshlets:,} meshlets, "
        f"{bd.stream_bytes:,} stream bytes, encode {t_enc:.0f} s")
    sweep = None
    if isinstance(tune, dict):          # a given launch shape (profiling runs)
        bd.set_tune(**tune)
        log(f"batch: tune {tune}")
    elif tune:
        best, sweep = autotune(bd, mode_kw={"mode": "graph"})
        log(f"batch: tuned {best}")
    r = {"workload": f"batch{n_meshes}_log_uniform_1k_200k_q12", "n_meshes": n_meshes,
         "n_tris": n_t, "n_verts_out": bd.n_v_total, "meshlets": bd.n_meshlets,
         "stream_bytes": bd.stream_bytes, "side_bytes": bd.side_bytes, "predictor": predictor,
         "encode_s": t_enc, "tune": bd.get_tune()}
    if sweep is not None:
        r["tune_sweep_ms"] = sweep
    bytes_alg = bd.bytes_read + bd.bytes_written
    r["algorithmic_bytes"] = bytes_alg
    r["bpv_mean"] = float(np.mean([8 * len(b) / nv for b, nv in zip(blobs, n_v)]))
    for mode in modes:
        # the per-mesh loop is ~100x slower: fewer repetitions keep the run short
        k = steps if mode != "per_mesh" else max(3, steps // 10)
        w, g = time_calls(functools.partial(bd.decode, mode=mode), k, min(warmup, k))
        r[mode] = {"ms_wall": w * 1e3, "ms_gpu": g * 1e3, "gtris_s": n_t / w / 1e9,
                   "hbm_gb_s": bytes_alg / g / 1e9}
        log(f"batch {mode}: {w * 1e3:.3f} ms wall, {n_t / w / 1e9:.3f} Gtris/s")
    if not {"graph", "per_mesh"} <= set(modes):
        return r
    # equal output: graph replay vs the per-mesh loop, bit for bit
    bd.decode(mode="graph")
    v1, t1 = bd.d_verts.clone(), bd.d_tris.clone()
    bd.d_verts.fill_(float("nan"))
    bd.d_tris.fill_(-1)
    bd.decode(mode="per_mesh")
    torch.cuda.synchronize()
    r["equal_output"] = bool(torch.equal(v1.view(torch.int32), bd.d_verts.view(torch.int32))
                             and torch.equal(t1, bd.d_tris))
    # the per-mesh loop a caller of the single-mesh surface writes
    decs = [ParaDeltaV5GpuDecoder(b) for b in blobs]
    for d in decs:
        d.set_tune(**{k: bd.get_tune()[k] for k in ("mode", "group_size", "lds_bytes", "tok_group",
                                                   "tok_chains")})

    def loop():
        for d in decs:
            d.decode()

    w, g = time_calls(loop, max(3, steps // 10), 1)
    r["per_mesh_python"] = {"ms_wall": w * 1e3, "ms_gpu": g * 1e3, "gtris_s": n_t / w / 1e9}
    log(f"batch per-mesh python loop: {w * 1e3:.3f} ms wall")
    del decs
    r["speedup_vs_per_mesh"] = r["per_mesh"]["ms_wall"] / r["graph"]["ms_wall"]
    r["speedup_vs_per_mesh_python"] = r["per_mesh_python"]["ms_wall"] / r["graph"]["ms_wall"]
    return r


def cpu_exhaustive(prep, n_threads):
    """The exhaustive predictor and coder search on n_threads host threads:
    the oracle's restatement of fit_predictor's control flow
    (encoder/_irlp_fit.py:364-394) with the C restatement of
    _cost_seeds_njit (:120-183) as the candidate cost, then _pick_best_k
    (encoder/paradelta_v5_dup.py:159-171) per stream in C.  The residual pass
    between the two is timed apart, as the encoder's own search_s does."""
    from oracle import stride_oracle as O
    from tests.oracle_c import load
    co = load()
    nat = prep["native"]
    A, B, C, V, off = (np.ascontiguousarray(x, dtype=np.int64) for x in nat.para_samples())

    def factory(a, b, c, v, o):
        return lambda cands, K: co.irlp_cost_mt(a, b, c, v, o, cands, K, n_threads)

    def full(a, b, c, v, o, n3, K):
        return int(co.irlp_cost_mt(a, b, c, v, o, [n3], K, n_threads)[0])

    t0 = time.perf_counter()
    pn, pk = O.fit_predictor(A, B, C, V, off, cost_fn_factory=factory, full_cost_fn=full)
    t1 = time.perf_counter()
    vals, soff, _ = nat.residuals(pn, pk)
    t2 = time.perf_counter()
    tagk, _ = co.pick_streams(vals, soff, n_threads)
    t3 = time.perf_counter()
    return {"predictor_s": t1 - t0, "coder_s": t3 - t2, "search_s": (t1 - t0) + (t3 - t2),
            "pred_n": np.asarray(pn).tolist(), "pred_K": np.asarray(pk).tolist(), "tagk": tagk,
            "vals": vals, "soff": soff}


def learned_encode(head_prep, n_heldout):
    """Search time and bytes of the learned proposal vs the exhaustive search."""
    from meshpress_amd import workloads
    from meshpress_amd.encoder import STRIDEEncoder
    from meshpress_amd.encoder.prep import prepare
    from meshpress_amd.utils.paradelta_v5_cuda import ParaDeltaV5GpuDecoder
    kw = dict(precision_error=bbox_error(12), precision_mode="bbox_frac")
    ex = STRIDEEncoder(**kw)
    le = STRIDEEncoder(selector="learned", **kw)
    lf = STRIDEEncoder(selector="learned+fallback", **kw)
    specs = workloads.learned_specs(50, 1)[:n_heldout]
    n_cpu = host_threads()
    preps = []
    for fam, n, seed in specs:
        V, T = workloads.mesh_with_tris(fam, n, seed)
        preps.append(prepare(V, T, **kw))
    for e in (ex, le, lf):                               # warm-up
        e.encode_prepared(preps[0])
    rows = []
    for name, prep in [("heightfield_1900_q12", head_prep)] + [(f"{f}_{n}_{sd}", p) for (f, n, sd), p
                                                                in zip(specs, preps)]:
        r = {"mesh": name, "n_tris": int(prep["n_t"])}
        outs = []
        for tag, e in (("exhaustive", ex), ("learned", le), ("learned_fallback", lf)):
            data = e.encode_prepared(prep)
            st = e.last_stats
            r[tag] = {"bytes": len(data), "search_s": st["search_s"],
                      "predictor_s": st["predictor_s"], "coder_s": st["coder_s"],
                      "pred_n": [int(x) for x in np.asarray(st["pred_n"]).reshape(-1)],
                      "pred_K": [int(x) for x in np.asarray(st["pred_K"]).reshape(-1)]}
            V, T = ParaDeltaV5GpuDecoder(data).decode_to_host()
            outs.append((V.view(np.uint32).copy(), T))
        r["decode_equal"] = all(bool(np.array_equal(outs[0][0], o[0]) and np.array_equal(outs[0][1], o[1]))
                                for o in outs[1:])
        # the same exhaustive search on the host's threads (encode-side CPU
        # baseline); its choices equal the GPU search's
        cpu = cpu_exhaustive(prep, n_cpu)
        from meshpress_amd.encoder import search
        gpu_tagk, _ = search.select_codes(cpu.pop("vals").astype(np.int32), cpu.pop("soff"))
        cpu_n, cpu_k, cpu_tagk = (cpu.pop(k) for k in ("pred_n", "pred_K", "tagk"))
        flat = lambda x: [int(v) for v in np.asarray(x).reshape(-1)]  # noqa: E731
        cpu["same_choice"] = bool(flat(cpu_n) == r["exhaustive"]["pred_n"]
                                  and flat(cpu_k) == r["exhaustive"]["pred_K"]
                                  and np.array_equal(cpu_tagk, gpu_tagk))
        r["exhaustive_cpu"] = cpu
        rows.append(r)
        ex_, le_ = r["exhaustive"], r["learned"]
        log(f"learned {name}: search {ex_['search_s'] * 1e3:.1f} ms (predictor "
            f"{ex_['predictor_s'] * 1e3:.1f} + coders {ex_['coder_s'] * 1e3:.1f}) -> "
            f"{le_['search_s'] * 1e3:.1f} ms ({le_['predictor_s'] * 1e3:.1f} + "
            f"{le_['coder_s'] * 1e3:.1f}), bytes {ex_['bytes']} -> {le_['bytes']} "
            f"(with fallback {r['learned_fallback']['search_s'] * 1e3:.1f} ms, "
            f"{r['learned_fallback']['bytes']}); "
            f"CPU exhaustive {r['exhaustive_cpu']['search_s'] * 1e3:.1f} ms on {n_cpu} threads "
            f"(same choice: {r['exhaustive_cpu']['same_choice']})")

    def agg(rs):
        be = sum(r["exhaustive"]["bytes"] for r in rs)
        bl = sum(r["learned"]["bytes"] for r in rs)
        bf = sum(r["learned_fallback"]["bytes"] for r in rs)
        te = sum(r["exhaustive"]["search_s"] for r in rs)
        tl = sum(r["learned"]["search_s"] for r in rs)
        tf = sum(r["learned_fallback"]["search_s"] for r in rs)
        tc = sum(r["exhaustive_cpu"]["search_s"] for r in rs)
        return {"meshes": len(rs), "exhaustive_search_s": te, "learned_search_s": tl,
                "speedup": te / tl, "regret": (bl - be) / be,
                "fallback_search_s": tf, "fallback_speedup": te / tf,
                "fallback_regret": (bf - be) / be,
                "decode_equal": all(r["decode_equal"] for r in rs),
                "cpu_exhaustive_search_s": tc, "cpu_threads": n_cpu,
                "speedup_vs_cpu_exhaustive": tc / tl,
                "cpu_same_choice": all(r["exhaustive_cpu"]["same_choice"] for r in rs)}

    return {"headline_mesh": agg(rows[:1]), "heldout": agg(rows[1:]), "all": agg(rows),
            "per_mesh": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="heightfield_1900_q12",
                    help="heightfield_1900_q12 | heightfield_1900_q16 | icosphere_l8_q12 | "
                         "batch (the batched workload alone, as the headline)")
    ap.add_argument("--batch", type=int, default=2048, help="meshes in the batched workload (0: skip)")
    ap.add_argument("--batch-cache", default=None, metavar="NPZ",
                    help="read the batch's streams from this file, or encode and write them there")
    ap.add_argument("--batch-modes", default="graph,launch,per_mesh",
                    help="batched decode modes to time (graph, launch, per_mesh)")
    ap.add_argument("--batch-predictor", default="generalized", choices=("generalized", "canonical"))
    ap.add_argument("--learned-heldout", type=int, default=50,
                    help="held-out meshes in the learned-encode comparison (0: skip)")
    ap.add_argument("--quick", action="store_true", help="headline workload only")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the decoders' default launch shape (profiling runs)")
    ap.add_argument("--tune", default=None, metavar="JSON",
                    help="launch shape of the headline (or --workload batch) decoder instead "
                         "of autotuning it, "
                         "e.g. '{\"tok_chains\": 8, \"tok_group\": 8}' (profiling runs)")
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
              "(decode shards by meshlet ranges with no exchange step)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        sys.exit("bench.py needs a ROCm GPU")
    torch.cuda.set_device(0)
    from meshpress_amd import workloads

    name = a.workload
    if name == "batch":
        r = batched(a.batch, a.steps, a.warmup, a.batch_predictor,
                    tune=json.loads(a.tune) if a.tune else not a.no_autotune,
                    modes=tuple(a.batch_modes.split(",")), cache=a.batch_cache)
        mode = "graph" if "graph" in r else tuple(k for k in ("launch", "per_mesh") if k in r)[0]
        print(json.dumps({
            "metric": METRIC, "value": r[mode]["gtris_s"], "unit": "Gtris/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": r[mode]["ms_wall"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32+f64", "data": "synthetic",
            "config": {"workload": r["workload"], "mode": mode, "n_tris": r["n_tris"],
                       "meshlets": r["meshlets"], "tune": r["tune"]},
            "details": {"batched": r}}), flush=True)
        return
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
    head_tune = json.loads(a.tune) if a.tune else not a.no_autotune
    dec, data, head, head_prep = single_mesh(name, V, T, bits, a.steps, a.warmup,
                                             tune=head_tune)
    details = {}
    cpu = None
    if not a.no_cpu:
        gpu_out = dec.decode_to_host()
        v, reps, tt, same = cpu_baseline(data, head["n_tris"], gpu_out)
        cpu = {"value": v, "unit": "Gtris/s", "cores": 1, "kind": "port",
               "sample": f"C oracle decode_dup restatement, full {name} stream x{reps} ({tt:.1f} s)",
               "gpu_output_matches": same}
        log(f"cpu baseline: {v:.4f} Gtris/s, GPU output {'matches' if same else 'DIFFERS'}")
        if not same:
            sys.exit("GPU decode differs from the oracle on the bench workload")
        nt = host_threads()
        vm, reps_m, tt_m, same_m = cpu_baseline(data, head["n_tris"], gpu_out, budget_s=5.0,
                                                n_threads=nt)
        cpu["all_threads"] = {"value": vm, "unit": "Gtris/s", "cores": nt,
                              "sample": f"same stream x{reps_m} ({tt_m:.1f} s), meshlets split "
                                        f"over {nt} threads", "gpu_output_matches": same_m}
        log(f"cpu baseline, {nt} threads: {vm:.4f} Gtris/s")
        if not same_m:
            sys.exit("GPU decode differs from the threaded oracle on the bench workload")
        cpu["python_c1"] = python_c1_baseline()
        log(f"cpu baseline, pure Python on the UV sphere: {cpu['python_c1']['value']:.6f} Gtris/s")
    del dec
    if not a.quick:
        details[f"{name}_pdlt_v5"] = pdlt_v5(head_prep, name, a.steps, a.warmup)
        torch.cuda.empty_cache()
        if name.startswith("heightfield_1900"):
            other = "heightfield_1900_q16" if bits == 12 else "heightfield_1900_q12"
            _, _, details[other], _ = single_mesh(other, V, T, 28 - bits, a.steps, a.warmup)
        del V, T
        torch.cuda.empty_cache()
        if not name.startswith("icosphere_l8"):
            log("generating icosphere level 8")
            Vi, Ti = workloads.icosphere_l8()
            _, _, details["icosphere_l8_q12"], _ = single_mesh("icosphere_l8_q12", Vi, Ti, 12,
                                                               a.steps, a.warmup)
            del Vi, Ti
        torch.cuda.empty_cache()
        if a.learned_heldout > 0 and bits == 12:
            details["learned_encode"] = learned_encode(head_prep, a.learned_heldout)
        del head_prep
        if a.batch > 0:
            details["batched"] = batched(a.batch, a.steps, a.warmup, a.batch_predictor,
                                         modes=tuple(a.batch_modes.split(",")),
                                         cache=a.batch_cache)
    out = {
        "metric": METRIC, "value": head["gtris_s"], "unit": "Gtris/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": head["ms_wall"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32+f64", "data": "synthetic",
        "config": {k: head[k] for k in ("workload", "n_tris", "n_verts_out", "meshlets", "g_bits",
                                        "stream_bytes", "side_bytes", "bpv", "tune")},
        # the HBM roofline is secondary: the decode is bound by instruction
        # issue along the format's serial chains (issue_model)
        "roofline": {"bound": "hbm", "achieved": head["hbm_gb_s"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["hbm_frac"], "traffic": read_traffic(name),
                     "algorithmic_bytes": head["algorithmic_bytes"], "kernel_ms": head["ms_kernel"],
                     "issue_model": issue_model(name, head["ms_kernel"]),
                     "limit": "instruction issue (issue_model); the HBM fraction is secondary"},
        "cpu_baseline": cpu,
        "details": details,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

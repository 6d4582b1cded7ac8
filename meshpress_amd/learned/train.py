"""Label, train and evaluate the proposal network on one GPU.

    python -m meshpress_amd.learned.train --out DIR [--train 500] [--eval 50]
        [--epochs 40] [--batch 4096] [--lr 3e-3] [--workers 16]

1. Labels: `--train` seeded meshes (displaced UV sphere, displaced torus,
   noise heightfield; 20K-500K triangles, 12-bit bbox quantisation) are
   prepared on a pool of host threads and labelled by the exhaustive GPU
   search (dataset.label_prep); the `--eval` held-out meshes (another seed)
   are labelled the same way for validation.
2. Training: Adam on the grouped cross-entropy, entirely in the
   stride_mlp_* HIP kernels (forward, backward, Adam); torch only gathers
   batches.  Weights go to DIR/proposal.npz.
3. Evaluation on the held-out meshes: each is encoded with the exhaustive
   search, the learned proposal, the learned proposal with fallback and the
   network-free baseline; reported are total-byte regret against the
   exhaustive encoding, search time and speedup, and whether every stream
   decodes (on the GPU) to exactly what the exhaustive stream decodes to.
   Results go to DIR/eval.json.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np


def log(*a):
    print("[learned]", *a, flush=True)


def bbox_error(bits: int) -> float:
    return 1.0 / ((1 << bits) - 1.5)


def _prep_one(spec, bits=12, n_threads=1):
    from .. import workloads
    from ..encoder.prep import prepare
    fam, n, seed = spec
    V, T = workloads.mesh_with_tris(fam, n, seed)
    return prepare(V, T, precision_error=bbox_error(bits), precision_mode="bbox_frac",
                   n_threads=n_threads)


def _label_one(spec):
    from .dataset import label_prep
    prep = _prep_one(spec)
    st = {}
    F, Y = label_prep(prep, st)
    st.update(family=spec[0], n_tris=int(prep["n_t"]))
    return F, Y, st


def label_set(specs, workers, what):
    from threadpoolctl import threadpool_limits
    t0 = time.perf_counter()
    out = [None] * len(specs)
    with threadpool_limits(limits=1), ThreadPoolExecutor(max_workers=workers) as pool:
        futs = [pool.submit(_label_one, s) for s in specs]
        for i, f in enumerate(futs):
            out[i] = f.result()
            if (i + 1) % 50 == 0 or i + 1 == len(specs):
                log(f"{what}: labelled {i + 1}/{len(specs)} meshes ({time.perf_counter() - t0:.0f} s)")
    X = np.concatenate([o[0] for o in out]).astype(np.float32)
    Y = np.concatenate([o[1] for o in out]).astype(np.uint8)
    return X, Y, [o[2] for o in out], time.perf_counter() - t0


def evaluate_heads(net, X, Y, chunk=65536):
    """Validation loss and per-head accuracy (ignored heads excluded)."""
    import torch
    from .mlp import PRED_CLASSES
    tot_loss, n = 0.0, len(X)
    hits = np.zeros(9)
    cnt = np.zeros(9)
    for s in range(0, n, chunk):
        x = net.features_to_device(X[s:s + chunk])
        y = Y[s:s + chunk]
        _, _, z = net.forward(x)
        _, loss = net.xent(z, torch.from_numpy(y).to("cuda"), 1.0)
        tot_loss += float(loss.sum())
        pp, cls, _ = net.heads(z)
        pp = pp.cpu().numpy()
        pred = np.stack([pp[:, PRED_CLASSES * d:PRED_CLASSES * (d + 1)].argmax(1) for d in range(3)], 1)
        guess = np.concatenate([pred, cls.cpu().numpy()], axis=1)
        ok = y != 255
        hits += ((guess == y) & ok).sum(0)
        cnt += ok.sum(0)
    return tot_loss / max(n, 1), (hits / np.maximum(cnt, 1)).tolist()


def train(X, Y, Xv, Yv, *, epochs, batch, lr, seed=0, progress=True):
    import torch
    from .mlp import ProposalMLP
    net = ProposalMLP(seed=seed)
    dX = net.features_to_device(X)
    dY = torch.from_numpy(Y).to("cuda")
    n = len(X)
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    warm = min(100, total // 10)
    hist = []
    step = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = torch.from_numpy(np.random.default_rng(seed * 7919 + ep).permutation(n)).to("cuda")
        run = torch.zeros((), dtype=torch.float64, device="cuda")
        for b in range(steps_per_epoch):
            idx = perm[b * batch:(b + 1) * batch]
            f = step / max(1, total - 1)
            cur = lr * min(1.0, (step + 1) / max(1, warm)) * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * f)))
            loss = net.train_step(dX.index_select(0, idx), dY.index_select(0, idx), cur)
            run += loss.double().sum()
            step += 1
        tr = float(run) / (steps_per_epoch * batch)
        vl, acc = evaluate_heads(net, Xv, Yv)
        hist.append({"epoch": ep, "train_loss": tr, "val_loss": vl, "val_acc": acc,
                     "s": time.perf_counter() - t0})
        if progress:
            log(f"epoch {ep}: train {tr:.4f} val {vl:.4f} acc pred "
                f"{np.mean(acc[:3]):.3f} coder {np.mean(acc[3:]):.3f} ({time.perf_counter() - t0:.1f} s)")
    return net, hist


def evaluate_encodes(specs, net, workers, tau_pred=0.2, tau_coder=0.9):
    """Per held-out mesh: bytes, search seconds and decode equality for the
    exhaustive, learned, learned+fallback and baseline encoders."""
    import torch
    from ..encoder import STRIDEEncoder
    from ..utils.paradelta_v5_cuda import ParaDeltaV5GpuDecoder
    from .proposal import BaselineProposal, LearnedProposal
    kw = dict(precision_error=bbox_error(12), precision_mode="bbox_frac")
    prop = LearnedProposal(mlp=net, tau_pred=tau_pred, tau_coder=tau_coder)
    encs = {"exhaustive": STRIDEEncoder(**kw),
            "learned": STRIDEEncoder(selector="learned", proposal=prop, **kw),
            "learned+fallback": STRIDEEncoder(selector="learned+fallback", proposal=prop, **kw),
            "baseline": STRIDEEncoder(selector="learned", proposal=BaselineProposal(), **kw)}
    # the encoder's own thread count (the default), as a caller would run it
    preps = [_prep_one(s, n_threads=None) for s in specs]
    # warm-up: kernels loaded and allocator primed before timing
    for e in encs.values():
        e.encode_prepared(preps[0])
    rows = []
    for spec, prep in zip(specs, preps):
        row = {"family": spec[0], "n_tris": int(prep["n_t"]), "n_v": int(prep["n_v"]),
               "meshlets": int(prep["n_meshlets"])}
        ref = None
        for name, e in encs.items():
            torch.cuda.synchronize()
            data = e.encode_prepared(prep)
            st = e.last_stats
            dec = ParaDeltaV5GpuDecoder(data)
            V, T = dec.decode_to_host()
            if ref is None:
                ref = (V.view(np.uint32).copy(), T.copy())
                same = True
            else:
                same = bool(np.array_equal(V.view(np.uint32), ref[0]) and np.array_equal(T, ref[1]))
            row[name] = {"bytes": len(data), "search_s": st["search_s"],
                         "predictor_s": st["predictor_s"], "coder_s": st["coder_s"],
                         "decode_equal": same, "pred_n": st["pred_n"], "pred_K": st["pred_K"],
                         "irlp_evals": st.get("irlp_evals"),
                         "pred_fallback_axes": st.get("pred_fallback_axes"),
                         "coder_fallback_streams": st.get("coder_fallback_streams")}
        rows.append(row)
    return rows


def summarize(rows):
    out = {}
    names = [k for k in rows[0] if isinstance(rows[0][k], dict)]
    for fam in (None, "uv_sphere", "torus", "heightfield"):
        rs = [r for r in rows if fam is None or r["family"] == fam]
        if not rs:
            continue
        ex = sum(r["exhaustive"]["bytes"] for r in rs)
        te = sum(r["exhaustive"]["search_s"] for r in rs)
        s = {}
        for nme in names:
            b = sum(r[nme]["bytes"] for r in rs)
            t = sum(r[nme]["search_s"] for r in rs)
            s[nme] = {"bytes": b, "regret": (b - ex) / ex, "search_s": t,
                      "speedup_vs_exhaustive": te / t if t > 0 else None,
                      "decode_equal": all(r[nme]["decode_equal"] for r in rs),
                      "max_mesh_regret": max((r[nme]["bytes"] - r["exhaustive"]["bytes"])
                                             / r["exhaustive"]["bytes"] for r in rs)}
        bl = s["baseline"]["bytes"]
        for nme in ("learned", "learned+fallback"):
            s[nme]["gain_vs_baseline"] = (bl - s[nme]["bytes"]) / bl
        out[fam or "all"] = {"meshes": len(rs), **s}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--train", type=int, default=500)
    ap.add_argument("--eval", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--t-max", type=int, default=500000)
    ap.add_argument("--save-dataset", action="store_true")
    a = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        sys.exit("training needs a ROCm GPU")
    from .. import workloads
    out = Path(a.out)
    out.mkdir(parents=True, exist_ok=True)
    tr_specs = workloads.learned_specs(a.train, 0, t_max=a.t_max)
    ev_specs = workloads.learned_specs(a.eval, 1, t_max=a.t_max)
    X, Y, tr_stats, t_lab = label_set(tr_specs, a.workers, "train")
    Xv, Yv, ev_stats, _ = label_set(ev_specs, a.workers, "held-out")
    log(f"dataset: {len(X) // 2:,} train meshlets ({len(X):,} rows) from {len(tr_specs)} meshes, "
        f"{len(Xv) // 2:,} held-out meshlets; labelling {t_lab:.0f} s")
    if a.save_dataset:
        np.savez_compressed(out / "dataset.npz", X=X.astype(np.float16), Y=Y,
                            Xv=Xv.astype(np.float16), Yv=Yv)
    net, hist = train(X, Y, Xv, Yv, epochs=a.epochs, batch=a.batch, lr=a.lr)
    net.save(out / "proposal.npz", epochs=a.epochs, batch=a.batch, lr=a.lr,
             train_meshes=len(tr_specs), train_meshlets=len(X) // 2)
    label_dist = {f"head{h}": np.bincount(Y[:, h][Y[:, h] != 255], minlength=20).tolist()
                  for h in range(9)}
    (out / "train_log.json").write_text(json.dumps(
        {"history": hist, "train_meshes": len(tr_specs), "train_meshlets": int(len(X) // 2),
         "train_rows": int(len(X)), "heldout_meshes": len(ev_specs),
         "heldout_meshlets": int(len(Xv) // 2),
         "label_s": t_lab, "label_distribution": label_dist}, indent=1))
    log("evaluating held-out encodes")
    rows = evaluate_encodes(ev_specs, net, a.workers)
    summ = summarize(rows)
    (out / "eval.json").write_text(json.dumps({"summary": summ, "meshes": rows}, indent=1))
    s = summ["all"]
    for nme in ("learned", "learned+fallback", "baseline"):
        log(f"{nme}: regret {100 * s[nme]['regret']:.3f}%  search speedup "
            f"{s[nme]['speedup_vs_exhaustive']:.1f}x  decode equal {s[nme]['decode_equal']}")
    print(json.dumps(summ), flush=True)


if __name__ == "__main__":
    main()

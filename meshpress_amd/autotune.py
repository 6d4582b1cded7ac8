"""Launch-shape autotuner for the STRIDE decode kernels.

Times candidate launch shapes on a decoder (ParaDeltaV5GpuDecoder or
BatchDecoder, anything with set_tune/get_tune/decode) with HIP events on the
launch stream and keeps the fastest.  The search space:

  mode 2 (split groups, 4 lanes per meshlet): group_size x lds_bytes, and
                                     persistent grids with a device work queue
  mode 4 (two passes: token lanes or token chains over tok_group meshlets per wave,
         then the split groups' axis lanes): group_size x lds_bytes x tok_group x
                                    tok_chains
  mode 1 (lane-per-meshlet groups):  group_size x lds_bytes x grid
  mode 0 (wave per meshlet):         waves x stage_words (x persistent queue)

group_size is the number of meshlets (lanes) a wave decodes together;
lds_bytes the LDS a group may use for its staged bits and packed codes
(smaller budgets give more resident waves per CU, larger ones fuller
waves).
"""
from __future__ import annotations

import itertools
import json

GROUP_SIZES = (16, 32, 48, 64)
LDS_BUDGETS = (16384, 24576, 32768, 40960, 49152, 65536)


SPLIT_GROUP_SIZES = (4, 8, 12, 16)
QUEUE_GRIDS = (1024, 1536, 3072)


def candidates(include_wave_mode: bool = True, grids=(0,)):
    out = []
    for gs, lds, grid in itertools.product(SPLIT_GROUP_SIZES, LDS_BUDGETS, grids):
        out.append(dict(mode=2, group_size=gs, lds_bytes=lds, grid=grid, use_queue=0))
    # persistent grids pulling groups from the device work queue
    for g in QUEUE_GRIDS:
        out.append(dict(mode=2, group_size=16, lds_bytes=24576, grid=g, use_queue=1))
    # two passes: token lanes over tok_group meshlets per wave, or token
    # chains (tok_chains per meshlet), then the axis lanes
    for gs, lds, tg in itertools.product((8, 16), (16384, 24576), (32, 64)):
        out.append(dict(mode=4, group_size=gs, lds_bytes=lds, tok_group=tg, tok_chains=0, grid=0,
                        use_queue=0))
    for (kc, tg), (gs, lds) in itertools.product(((4, 16), (8, 8), (16, 4)),
                                                 ((8, 24576), (16, 16384), (21, 24576))):
        out.append(dict(mode=4, group_size=gs, lds_bytes=lds, tok_group=tg, tok_chains=kc,
                        grid=0, use_queue=0))
    for gs, lds, grid in itertools.product(GROUP_SIZES, LDS_BUDGETS, grids):
        out.append(dict(mode=1, group_size=gs, lds_bytes=lds, grid=grid, use_queue=0))
    if include_wave_mode:
        for w, sw in itertools.product((1, 4), (512, 1024)):
            out.append(dict(mode=0, waves=w, stage_words=sw, grid=0, use_queue=0))
    return out


# The compact search bench.py and the full-size tests run: split groups over
# LDS budgets, the split kernel's persistent queue grids, the two-pass shapes
# and lane-per-meshlet groups.
BENCH_CANDIDATES = ([dict(mode=2, group_size=gs, lds_bytes=lds, grid=0, use_queue=0)
                     for gs in (8, 16) for lds in (16384, 24576, 32768, 40960, 49152)]
                    + [dict(mode=2, group_size=16, lds_bytes=24576, grid=g, use_queue=1)
                       for g in (1536, 3072)]
                    + [dict(mode=4, group_size=gs, lds_bytes=lds, tok_group=32, tok_chains=0, grid=0,
                            use_queue=0) for gs, lds in ((16, 16384), (8, 16384))]
                    # the chain token pass: chains per meshlet x meshlets per token wave,
                    # axis groups of 6 to 21 meshlets (which pair wins depends on
                    # the mesh: 8 x 8 with groups of 16 on the 12-bit heightfield, the
                    # default 16 x 4 with groups of 21 on its 16-bit twin, 16 x 2
                    # with groups of 6 on the 1.3M-triangle icosphere)
                    + [dict(mode=4, group_size=gs, lds_bytes=lds, tok_chains=kc, tok_group=tg,
                            grid=0, use_queue=0)
                       for kc, tg in ((4, 16), (8, 8), (16, 4))
                       for gs, lds in ((8, 24576), (16, 16384), (21, 24576))]
                    # meshes of a few thousand meshlets (fewer waves than SIMDs):
                    # half-filled token waves and axis groups of 6 keep more
                    # waves in flight
                    + [dict(mode=4, group_size=gs, lds_bytes=16384, tok_chains=16, tok_group=tg,
                            grid=0, use_queue=0)
                       for tg, gs in ((2, 6), (2, 8), (4, 6))]
                    + [dict(mode=1, group_size=gs, lds_bytes=lds, grid=0, use_queue=0)
                       for gs in (32, 64) for lds in (16384, 40960)])


def time_once(dec, steps: int, mode_kw=None) -> float:
    """Average ms per decode over `steps` launches (HIP events)."""
    import torch
    kw = mode_kw or {}
    dec.decode(**kw)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        dec.decode(**kw)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def tune(dec, steps: int = 10, cands=None, mode_kw=None, log=None):
    """Tries every candidate, leaves the decoder on the fastest, returns
    (best_tune, [(tune, ms), ...])."""
    base = dec.get_tune()
    results = []
    for c in cands or candidates():
        try:
            dec.set_tune(**c)
        except ValueError:
            continue
        ms = time_once(dec, steps, mode_kw)
        results.append((dict(c), ms))
        if log:
            log(f"tune {json.dumps(c)}: {ms:.4f} ms")
    if not results:
        dec.set_tune(**base)
        return base, results
    best = min(results, key=lambda r: r[1])[0]
    dec.set_tune(**best)
    return best, results

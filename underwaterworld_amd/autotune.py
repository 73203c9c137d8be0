"""Autotuner for the batched persistent kernel's launch shape.

    python -m underwaterworld_amd.autotune [--quick] [--out underwaterworld_amd/tuning.json]

For each workload (the 4096-chunk 12^3 grid with 4 octaves, the 8192-chunk mixed-resolution set
with 6 octaves, the 512^3 volume as 4096 bricks of 32^3 with 8 octaves) it times, on the device, every combination of
    block   threads per workgroup: 128, 192, 256, 512
    budget  LDS bytes allowed for a chunk's density slab ring: a chunk whose (n+1)^3 densities fit
            is built in one slab (one density evaluation); larger chunks use the largest x-slab
            that fits (densities evaluated twice)
    grid    persistent workgroups: all that fit (0)
    bulk_min  chunks of at least this many cells per axis go to the bulk path (0: off); swept last
            over the workload's resolutions
with max_res set to the workload's largest resolution (it sizes the launch's LDS). The budget ties
every resolution's x-slab size to one LDS figure; the fastest shape is then refined one x-slab
resolution of the workload at a time (coordinate descent over that resolution's planes per slab,
within 4 of the current value), since the best slab of one size need not share the others'
budget; the descent repeats while a pass improves. The fastest shape per octave count goes to tuning.json, which `Perlin` applies when it
creates a context. Every candidate's mesh is checked against the first candidate's (all shapes
must produce identical meshes) before its time is accepted.

    --refine-from FILE   skip the sweep where FILE has a shape for the workload's octave count: start
                         from it and refine only
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tuning.json")

BLOCKS = (128, 192, 256, 512)
BUDGETS = (6 * 1024, 10 * 1024, 14 * 1024, 20 * 1024, 28 * 1024, 40 * 1024, 56 * 1024)
GRIDS = (0,)


def slabs_for_budget(budget: int):
    s = [1]
    for n in range(1, 33):
        L = n + 1
        if 4 * L ** 3 <= budget:
            s.append(n)
        else:
            s.append(max(1, min(n, budget // (4 * L * L) - 1)))
    return s


def _mesh_digest(batch, nchunks):
    info = batch.info()
    h = []
    for i in range(nchunks):
        v, ix = batch.chunk_mesh(i, info)
        h.append((int(info[i]["n_verts"]), int(info[i]["n_inds"]), hash(v.tobytes()), hash(ix.tobytes())))
    return h


def _time_shape(perlin, coords, res, t, steps, stream, ref, log, label):
    """Mean graph-replay time of shape t (ms), or None when the shape does not fit; ref is a
    one-element list holding the first accepted mesh digest."""
    import torch

    from .engine import BatchedGraph, set_batch_tuning

    try:
        set_batch_tuning(perlin, t)
        g = BatchedGraph(perlin, coords, res)
    except Exception as e:  # shape does not fit (LDS) -> skip
        log(f"  {label}: skipped ({e})")
        return None
    g.launch(stream.cuda_stream)
    torch.cuda.synchronize()
    dig = _mesh_digest(g.batch, len(coords))
    if ref[0] is None:
        ref[0] = dig
    elif dig != ref[0]:
        raise RuntimeError(f"tuning {label} changed the mesh")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        g.launch(stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    g.close()
    log(f"  {label}: {ms:.4f} ms")
    return ms


def refine_slabs(perlin, coords, res, best_ms, best, steps, stream, ref, log=print, passes=3):
    """Coordinate descent over the planes per slab of each x-slab resolution of the workload,
    repeated while a pass improves (one resolution's slab changes the LDS, and so the cell ring
    and the occupancy, that the others' best values depend on)."""
    results = []
    tried = set()
    for _ in range(passes):
        improved = False
        for n in sorted(set(int(r) for r in res)):
            s0 = int(best["slab"][n])
            if s0 >= n:  # one slab: nothing to refine
                continue
            for cand in range(max(1, s0 - 4), min(n - 1, s0 + 4) + 1):
                t = dict(best, slab=list(best["slab"]))
                t["slab"][n] = cand
                key = tuple(t["slab"])
                if cand == s0 or key in tried:
                    continue
                tried.add(key)
                ms = _time_shape(perlin, coords, res, t, steps, stream, ref, log, f"slab{n}={cand}")
                if ms is None:
                    continue
                results.append((ms, t))
                if ms < best_ms:
                    best_ms, best, improved = ms, t, True
        if not improved:
            break
    return best_ms, best, results


def refine_bulk(perlin, coords, res, best_ms, best, steps, stream, ref, log=print):
    """The bulk path's threshold: off (0), or, for each distinct resolution r of the workload above
    the one-pass sizes, one more than the largest resolution below r (chunks of at least bulk_min
    cells per axis leave the persistent queue, whose LDS is then sized for bulk_min - 1: the
    tightest threshold that moves r and above gives the persistent kernel its smallest shape)."""
    results = []
    rs = sorted({int(n) for n in res})
    cands = [0] + [max([q for q in rs if q < r], default=0) + 1 for r in rs if 13 <= r <= int(best["max_res"])]
    for b in cands:
        if b == int(best.get("bulk_min", -1)):
            continue
        t = dict(best, slab=list(best["slab"]), bulk_min=b)
        ms = _time_shape(perlin, coords, res, t, steps, stream, ref, log, f"bulk_min={b}")
        if ms is None:
            continue
        results.append((ms, t))
        if ms < best_ms:
            best_ms, best = ms, t
    return best_ms, best, results


def tune_workload(perlin, coords, res, steps=10, quick=False, log=print, start=None):
    import torch

    from .engine import get_batch_tuning, set_batch_tuning

    base = get_batch_tuning(perlin)
    results = []
    ref = [None]
    stream = torch.cuda.current_stream()
    if start is None:
        blocks = (256,) if quick else BLOCKS
        budgets = (10 * 1024, 20 * 1024) if quick else BUDGETS
        for block in blocks:
            for budget in budgets:
                for grid in GRIDS:
                    t = dict(base, block=block, grid=grid, max_res=int(max(res)), slab=slabs_for_budget(budget))
                    ms = _time_shape(perlin, coords, res, t, steps, stream, ref, log,
                                     f"block={block} budget={budget} grid={grid}")
                    if ms is not None:
                        results.append((ms, t))
        results.sort(key=lambda r: r[0])
        best_ms, best = results[0]
    else:
        best = dict(start, slab=list(start["slab"]))
        best_ms = _time_shape(perlin, coords, res, best, steps, stream, ref, log, "start")
        results.append((best_ms, best))
    best_ms, best, refined = refine_slabs(perlin, coords, res, best_ms, best, steps, stream, ref, log)
    results += refined
    best_ms, best, bulk = refine_bulk(perlin, coords, res, best_ms, best, steps, stream, ref, log)
    results += bulk
    set_batch_tuning(perlin, base)
    results.sort(key=lambda r: r[0])
    return results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--refine-from", default=None)
    args = ap.parse_args()
    start_file = json.load(open(args.refine_from)) if args.refine_from else None
    sys.path.insert(0, os.path.dirname(HERE))
    from . import params
    from .engine import Perlin

    report = {"by_octaves": {}, "timings": {}}
    jobs = [("chunks12", params.grid_coords(-8, 8), [12] * 4096)]
    mc, mr = params.mixed_workload(8192, 7)
    jobs.append(("mixed", mc, mr))
    jobs.append(("volume_bricks", params.volume_bricks(16), [32] * 4096))
    for name, coords, res in jobs:
        p = params.preset(name)
        perlin = Perlin(p, apply_tuning=False)
        print(f"tuning {name} ({len(coords)} chunks, {p['octaves']} octaves)", flush=True)
        start = start_file["by_octaves"].get(str(p["octaves"])) if start_file else None
        r = tune_workload(perlin, np.asarray(coords, np.int32), res, steps=args.steps, quick=args.quick,
                          start=start)
        best_ms, best = r[0]
        report["by_octaves"][str(p["octaves"])] = best
        report["timings"][name] = [{"ms": round(ms, 5), "block": t["block"], "grid": t["grid"],
                                    "slab12": t["slab"][12], "slab24": t["slab"][24], "slab32": t["slab"][32],
                                    "bulk_min": t.get("bulk_min", 0)}
                                   for ms, t in r]
        print(f"best {name}: {best_ms:.4f} ms block={best['block']} slab12={best['slab'][12]} "
              f"slab24={best['slab'][24]} slab32={best['slab'][32]}", flush=True)
    with open(args.out, "w") as f:
        json.dump(report, f, indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()

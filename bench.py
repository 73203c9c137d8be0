"""Chunk-meshing throughput on one MI355X.

Workload (one "step"): the 4096 chunks of the 16^3 chunk grid [-8, 8)^3, each 12^3 cells (13^3 f32
density samples), density = 4-octave Perlin fBm of seed 42 (base frequency 1/32 per voxel,
lacunarity 2, persistence 0.5) plus a linear z-bias, marched at isovalue 0. Chunk coordinates are
resident in HBM before timing starts; the mesh (xyz f32 vertices, u32 indices, per-chunk records)
is written to preallocated HBM buffers.

Paths: "plain" = one launch per builder stage over all chunks (density, classify + compact,
offsets, emit); "batched" = one persistent launch (plus two set-up kernels that write the z tables and
order the queue), also replayed from a captured hipGraph. `value` = chunks/s of the batched path. Extra fields: plain chunks/s,
voxel samples/s of both, and the mixed-resolution workload (8192 chunks, 25 % each at 8^3, 12^3,
24^3, 32^3, 6 octaves) through the batched path.

`roofline`: the persistent kernel of the headline path, FP32-VALU bound, algorithmic FLOPs
F(O) = 78 O + 6 per sample (per launch F(4) * 13^3 * 4096 = 2.86 GFLOP), timed with HIP events
recorded by the library on the launch stream around that kernel. `cpu_baseline`: the reference CPU
path (the oracle's reference-faithful C restatement: f64 noise, literal linear-search dedup), one
thread, repeated passes over all 4096 chunks of the same workload for at least 10 s;
`cpu_baseline_all_cores` is the same build chunk-parallel over the host threads this process may
use, passes for at least 2 s.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def flops_per_sample(octaves: int) -> int:
    return 78 * octaves + 6


def read_traffic(name: str):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary (profiles/<name>), if any."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-resolution workload")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint workload")
    ap.add_argument("--no-volume", action="store_true", help="skip the 512^3 volume (forward + adjoint) workload")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; sharding across GPUs is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    from underwaterworld_amd import BatchedGraph, Perlin, _lib, get_batch_tuning, params, set_batch_tuning
    from underwaterworld_amd.engine import tuned_shape
    from underwaterworld_amd.engine import _alloc_out, _mesh_out, batch_capacity

    L = _lib.lib()
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    def kernel_ms(perlin, fn, steps):
        L.uw_profile_enable(perlin.ctx, 1)
        for _ in range(steps):
            fn()
        tot, cnt = ctypes.c_double(), ctypes.c_uint64()
        _lib.check(L.uw_profile_read(perlin.ctx, ctypes.byref(tot), ctypes.byref(cnt)), perlin.ctx)
        L.uw_profile_enable(perlin.ctx, 0)
        return tot.value / max(cnt.value, 1)

    # ---- headline workload: 4096 chunks of 12^3 ----
    p = params.preset("chunks12")
    n = 12
    coords = np.array(params.grid_coords(-8, 8), np.int32)
    nch = len(coords)
    S = (n + 1) ** 3
    perlin = Perlin(p)
    if tuned_shape(p["octaves"]) is None:  # no autotuned shape: size the launch for 12^3 chunks
        t = get_batch_tuning(perlin)
        t["max_res"] = n
        set_batch_tuning(perlin, t)
    ct = torch.as_tensor(coords, device=dev)
    rt = torch.full((nch,), n, dtype=torch.int32, device=dev)
    vcap, icap = int(L.uw_max_verts(n)) * nch, int(L.uw_max_inds(n)) * nch
    verts, inds, info, totals = _alloc_out(nch, vcap, icap, dev)
    out = _mesh_out(verts, inds, info, totals, vcap, icap)

    def plain_step():
        rc = L.uw_build_chunks_plain(perlin.ctx, ctypes.c_void_p(ct.data_ptr()), nch, n, ctypes.byref(out),
                                     ctypes.c_void_p(0), sp)
        if rc:
            _lib.check(rc, perlin.ctx)

    def batched_step():
        rc = L.uw_build_chunks_batched(perlin.ctx, ctypes.c_void_p(ct.data_ptr()), ctypes.c_void_p(rt.data_ptr()),
                                       nch, ctypes.byref(out), sp)
        if rc:
            _lib.check(rc, perlin.ctx)

    ms_plain = timed(plain_step, args.steps, args.warmup)
    print(f"# plain path: {ms_plain:.4f} ms/step", file=sys.stderr, flush=True)
    ms_batched = timed(batched_step, args.steps, args.warmup)
    print(f"# batched path (eager): {ms_batched:.4f} ms/step", file=sys.stderr, flush=True)
    graph = BatchedGraph(perlin, coords, rt)
    ms_graph = timed(lambda: graph.launch(stream.cuda_stream), args.steps, args.warmup)
    print(f"# batched path (graph): {ms_graph:.4f} ms/step", file=sys.stderr, flush=True)
    gt = graph.batch.totals()

    # the reference's call pattern: one plain build per chunk (chunk.rs build_full per chunk)
    n_single = 256
    def per_chunk_calls():
        for i in range(n_single):
            rc = L.uw_build_chunks_plain(perlin.ctx, ctypes.c_void_p(ct.data_ptr() + 12 * i), 1, n, ctypes.byref(out),
                                         ctypes.c_void_p(0), sp)
            if rc:
                _lib.check(rc, perlin.ctx)
    ms_single = timed(per_chunk_calls, 3, 1) / n_single
    print(f"# plain path, one chunk per call: {ms_single * 1e3:.2f} us/chunk", file=sys.stderr, flush=True)
    ms_kernel = kernel_ms(perlin, batched_step, args.steps)
    ms_noise_plain = kernel_ms(perlin, plain_step, args.steps)
    print(f"# persistent kernel {ms_kernel:.4f} ms; plain density kernel {ms_noise_plain:.4f} ms",
          file=sys.stderr, flush=True)
    graph.close()

    if args.copy_mib:
        a = torch.empty(args.copy_mib << 18, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()

    # ---- mixed-resolution workload: 8192 chunks, 6 octaves ----
    mixed = None
    if not args.no_mixed:
        pm = params.preset("mixed")
        mc, mr = params.mixed_workload(8192, 7)
        pmix = Perlin(pm)
        mv, mi = batch_capacity(mr)
        g2 = BatchedGraph(pmix, mc, mr, vert_cap=mv, ind_cap=mi)
        ms_mixed = timed(lambda: g2.launch(stream.cuda_stream), max(3, args.steps // 4), 1)
        mt = g2.batch.totals()
        samples = int(sum((r + 1) ** 3 for r in mr))
        flops_m = flops_per_sample(pm["octaves"]) * samples
        mixed = {"workload": "8192 chunks, 2048 each at 8^3/12^3/24^3/32^3 cells, coords uniform in 32x32x8 "
                             "(SplitMix64 seed 7), 6 octaves",
                 "ms_per_step": round(ms_mixed, 4), "chunks_per_s": round(8192 / (ms_mixed / 1e3), 1),
                 "samples_per_s": round(samples / (ms_mixed / 1e3), 1), "samples": samples,
                 "algorithmic_tflops": round(flops_m / (ms_mixed / 1e3) / 1e12, 3),
                 "nonblank_chunks": int(mt["n_nonblank"]), "verts": int(mt["n_verts"]), "inds": int(mt["n_inds"])}
        print(f"# mixed workload: {ms_mixed:.4f} ms/step", file=sys.stderr, flush=True)
        g2.close()

    # ---- adjoint: 256 chunks of 24^3, forward with tape + backward (densities, amplitudes, offset) ----
    adjoint = None
    if not args.no_adjoint:
        from underwaterworld_amd.adjoint import ChunkMesher
        pa = params.preset("adjoint24")
        ac = [(x, y, z) for x in range(-4, 4) for y in range(-4, 4) for z in range(-3, 1)]
        pad = Perlin(pa)
        mesher = ChunkMesher(pad, ac, 24, want_dens_grad=True)
        mesher.forward()
        torch.cuda.synchronize()
        nvt = int(mesher.batch.totals()["n_verts"])
        gv = torch.zeros((mesher.vert_cap, 3), dtype=torch.float32, device=dev)
        gv[:nvt, 2] = 1.0

        def adj_step():
            mesher.forward(stream.cuda_stream)
            if mesher.grad_dens is not None:
                mesher.grad_dens.zero_()
            mesher.backward(gv, stream.cuda_stream)

        ms_adj = timed(adj_step, max(3, args.steps // 4), 1)
        L.uw_profile_enable(pad.ctx, 1)
        mesher.backward(gv, stream.cuda_stream)
        tot_b, cnt_b = ctypes.c_double(), ctypes.c_uint64()
        L.uw_profile_read(pad.ctx, ctypes.byref(tot_b), ctypes.byref(cnt_b))
        L.uw_profile_enable(pad.ctx, 0)
        adjoint = {"workload": "256 chunks (x,y in [-4,4), z in [-3,1)) x 24^3 cells, 6 octaves, seed 42; "
                               "forward with vertex tape + backward to 25^3 densities, 6 amplitudes, 3 offsets",
                   "ms_per_step": round(ms_adj, 4), "grads_per_s": round(len(ac) / (ms_adj / 1e3), 1),
                   "backward_kernel_ms": round(tot_b.value / max(cnt_b.value, 1), 4), "vertices": nvt}
        print(f"# adjoint: {ms_adj:.4f} ms/step (forward + backward)", file=sys.stderr, flush=True)

    # ---- volume: 512^3 cells as 16^3 bricks of 32^3, 8 octaves; forward + adjoint, peak HBM ----
    volume = None
    if not args.no_volume and not args.no_adjoint:
        from underwaterworld_amd.adjoint import ChunkMesher
        pv = params.preset("volume_bricks")
        bricks = params.volume_bricks(16)
        pvol = Perlin(pv)
        tv = get_batch_tuning(pvol)
        tv["max_res"] = 32
        set_batch_tuning(pvol, tv)
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats(dev)
        base_mem = torch.cuda.memory_allocated(dev)
        vm = ChunkMesher(pvol, bricks, 32, want_dens_grad=True, exact_capacity=True)
        vm.forward(stream.cuda_stream)
        torch.cuda.synchronize()
        vt = vm.batch.totals()
        nvv = int(vt["n_verts"])
        gvv = torch.zeros((vm.vert_cap, 3), dtype=torch.float32, device=dev)
        gvv[:nvv, 2] = 1.0
        ms_vfwd = timed(lambda: vm.forward(stream.cuda_stream), 2, 1)

        def vol_step():
            vm.forward(stream.cuda_stream)
            if vm.grad_dens is not None:
                vm.grad_dens.zero_()
            vm.backward(gvv, stream.cuda_stream)

        ms_vadj = timed(vol_step, 2, 1)
        peak = torch.cuda.max_memory_allocated(dev) - base_mem
        vinfo = vm.batch.info()
        act_cells = int(vinfo["n_active"].sum())
        nb = len(bricks)
        samples_v = nb * 33 ** 3
        o_v = int(pv["octaves"])
        # Without recomputation the backward would keep every sample's density and, per octave, the
        # noise value and its gradient (4 floats): the stored-field alternative to recomputing.
        stored_dens = samples_v * 4
        stored_oct = samples_v * o_v * 4 * 4
        volume = {"workload": "512^3 cells (16^3 bricks of 32^3, 33^3 samples each), seed 1234, 8 octaves, "
                              "linear z-bias (0.01, -0.05); forward + adjoint to densities, amplitudes, offset",
                  "forward_ms": round(ms_vfwd, 3), "forward_plus_adjoint_ms": round(ms_vadj, 3),
                  "samples": samples_v, "samples_per_s_forward": round(samples_v / (ms_vfwd / 1e3), 1),
                  "verts": nvv, "inds": int(vt["n_inds"]), "active_cells": act_cells,
                  "active_fraction": round(act_cells / (nb * 32 ** 3), 4),
                  "peak_hbm_bytes_recompute": int(peak),
                  "hbm_bytes_without_recompute": int(peak + stored_dens + stored_oct),
                  "note": "peak = torch allocations over the forward + adjoint steps (outputs, vertex tape, "
                          "density gradients); densities and octave values stay in LDS and are recomputed "
                          "in the backward. The alternative adds the stored densities (4 B per sample) and "
                          "per-octave value + gradient (16 B per sample per octave)."}
        print(f"# volume: forward {ms_vfwd:.3f} ms, forward + adjoint {ms_vadj:.3f} ms, peak {peak / 2**20:.1f} MiB",
              file=sys.stderr, flush=True)
        del vm, gvv

    flops = flops_per_sample(p["octaves"]) * S * nch
    achieved = flops / (ms_kernel / 1e3) / 1e12
    roof = {"bound": "fp32-valu", "kernel": "k_batched (persistent: density + mesh)", "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": read_traffic("k_batched_chunks12.md"),
            "algorithmic_flop_per_launch": flops, "ms_per_launch": round(ms_kernel, 5),
            "plain_density_kernel": {"ms_per_launch": round(ms_noise_plain, 5),
                                     "achieved_tflops": round(flops / (ms_noise_plain / 1e3) / 1e12, 3)}}

    cpu = cpu_all = None
    if not args.no_cpu:
        from oracle import oracle as O

        def timed_passes(nthreads: int, min_s: float, max_passes: int = 40):
            """Whole-workload passes until `min_s` seconds have elapsed (bounded)."""
            passes, t0 = 0, time.perf_counter()
            while True:
                O.build_chunks(p, coords, n, nthreads=nthreads, dedup="linear")
                passes += 1
                dt = time.perf_counter() - t0
                if dt >= min_s or passes >= max_passes:
                    return passes, dt

        k, dt = timed_passes(1, 10.0)
        cpu = {"value": round(k * nch / dt, 2), "unit": "chunks/s", "cores": 1, "kind": "port",
               "sample": f"{k} passes over all {nch} chunks of the workload ({k * nch} chunk builds), "
                         "reference-faithful C restatement (f64 noise, linear-search dedup), 1 thread",
               "seconds": round(dt, 3)}
        nthr = max(1, min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 64))
        k, dt = timed_passes(nthr, 2.0)
        cpu_all = {"value": round(k * nch / dt, 2), "unit": "chunks/s", "cores": nthr, "kind": "port",
                   "sample": f"{k} passes over all {nch} chunks, same build, chunk-parallel over {nthr} threads",
                   "seconds": round(dt, 3)}

    # headline: the faster of the two batched launch forms (both are reported)
    headline_ms = min(ms_graph, ms_batched)
    headline_path = ("batched (persistent kernel, hipGraph replay)" if ms_graph <= ms_batched
                     else "batched (persistent kernel, stream launch)")
    res = {
        "metric": "chunks/s at 12^3 cells (batched vs plain GPU); voxel samples/s; adjoint grads/s",
        "value": round(nch / (headline_ms / 1e3), 1),
        "unit": "chunks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(headline_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural density field, seed 42)",
        "config": {"workload": "4096 chunks (16^3 grid [-8,8)^3) x 12^3 cells, 13^3 samples each; "
                               "seed 42, 4 octaves, freq 1/32 per voxel, linear z-bias, iso 0",
                   "path": headline_path, "chunks": nch, "resolution": n},
        "samples_per_s": round(nch * S / (headline_ms / 1e3), 1),
        "batched_eager_ms": round(ms_batched, 5),
        "batched_graph_ms": round(ms_graph, 5),
        "plain_one_chunk_per_call_chunks_per_s": round(1e3 / ms_single, 1),
        "batched_over_one_chunk_per_call": round(ms_single * nch / headline_ms, 1),
        "plain_ms": round(ms_plain, 5),
        "plain_chunks_per_s": round(nch / (ms_plain / 1e3), 1),
        "plain_samples_per_s": round(nch * S / (ms_plain / 1e3), 1),
        "batched_over_plain": round(ms_plain / headline_ms, 3),
        "total_verts": int(gt["n_verts"]), "total_inds": int(gt["n_inds"]), "nonblank": int(gt["n_nonblank"]),
        "mixed": mixed,
        "adjoint": adjoint,
        "volume": volume,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

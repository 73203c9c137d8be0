"""Chunk-meshing throughput on one MI355X.

Workload (one "step"): the 4096 chunks of the 16^3 chunk grid [-8, 8)^3, each 12^3 cells (13^3 f32
density samples), density = 4-octave Perlin fBm of seed 42 (base frequency 1/32 per voxel,
lacunarity 2, persistence 0.5) plus a linear z-bias, marched at isovalue 0. Chunk coordinates are
resident in HBM before timing starts; the mesh (xyz f32 vertices, u32 indices, per-chunk records)
is written to preallocated HBM buffers.

Paths: "plain" = one launch per builder stage over all chunks (density, classify + compact,
offsets, emit); "batched" = one persistent launch (plus two set-up kernels that write the z tables
and order the queue), also replayed from a captured hipGraph. `value` = chunks/s of the batched
path. Beside it: the plain path with per-stage kernel times (HBM GB/s of classify/compaction and
emit against the 8 TB/s peak), the reference's one-chunk-per-call pattern, the mixed-resolution
workload (8192 chunks, 25 % each at 8^3, 12^3, 24^3, 32^3, 6 octaves), the adjoint workload (256
chunks of 24^3, forward + backward) and the single 512^3 volume (513^3 samples, 8 octaves,
forward + adjoint with the noise recomputed in the backward) with its device-wide memory use.

`roofline`: the persistent kernel of the headline path, FP32-VALU bound, algorithmic FLOPs
F(O) = 78 O + 6 per sample (per launch F(4) * 13^3 * 4096 = 2.86 GFLOP), timed with HIP events
recorded by the library on the launch stream around that kernel. `cpu_baseline`: the reference CPU
path (the oracle's reference-faithful C restatement: f64 noise, literal linear-search dedup), one
thread, repeated passes over all 4096 chunks of the same workload for at least 10 s;
`cpu_baseline_all_cores` is the same build chunk-parallel over the host threads this process may
use, passes for at least 2 s. The same leg checks the GPU output of the 4096 chunks against the
oracle (its reference-faithful and FP64 variants) and reports the errors and tie counts (`parity`).

--workload {headline,plain,mixed,adjoint,volume} runs one workload only (for kernel profiling) and
prints a short JSON line; the default runs everything and prints the full line.
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
WORKLOADS = ("all", "headline", "plain", "mixed", "adjoint", "volume")


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


def host_cpu():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=WORKLOADS, default="all")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity leg")
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

    from underwaterworld_amd import (BatchedGraph, Perlin, _lib, build_chunks_batched, get_batch_tuning, params,
                                     sample_density)
    from underwaterworld_amd import build_chunks_plain, set_batch_tuning
    from underwaterworld_amd.engine import _alloc_out, _mesh_out, batch_capacity, tuned_shape

    L = _lib.lib()
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    wl = args.workload
    run = lambda name: wl in ("all", name)  # noqa: E731

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

    def stage_ms(perlin, fn, steps):
        """Average ms per launch of every profiling stage over `steps` calls of fn."""
        L.uw_profile_enable(perlin.ctx, 1)
        for _ in range(steps):
            fn()
        k = _lib.STAGE_COUNT
        tot, cnt = (ctypes.c_double * k)(), (ctypes.c_uint64 * k)()
        _lib.check(L.uw_profile_read_stages(perlin.ctx, k, tot, cnt), perlin.ctx)
        L.uw_profile_enable(perlin.ctx, 0)
        return [tot[i] / cnt[i] if cnt[i] else None for i in range(k)]

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    res = {}
    if args.copy_mib:  # one plain device copy of known size: the HBM counters' calibration
        a = torch.empty(args.copy_mib << 18, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

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

    if wl == "plain":
        ms = timed(plain_step, args.steps, args.warmup)
        print(json.dumps({"workload": "plain 4096 x 12^3", "ms_per_step": round(ms, 5)}), flush=True)
        return
    if wl == "headline":
        ms = timed(batched_step, args.steps, args.warmup)
        print(json.dumps({"workload": "batched 4096 x 12^3", "ms_per_step": round(ms, 5)}), flush=True)
        return

    if run("headline"):
        ms_plain = timed(plain_step, args.steps, args.warmup)
        log(f"# plain path: {ms_plain:.4f} ms/step")
        ms_batched = timed(batched_step, args.steps, args.warmup)
        log(f"# batched path (eager): {ms_batched:.4f} ms/step")
        graph = BatchedGraph(perlin, coords, rt)
        ms_graph = timed(lambda: graph.launch(stream.cuda_stream), args.steps, args.warmup)
        log(f"# batched path (graph): {ms_graph:.4f} ms/step")
        gt = graph.batch.totals()
        graph.close()

        # The same step with host buffers on both sides, as the host facades use it: coordinates and
        # resolutions up from pinned memory, the launch totals down, then that many vertices and
        # indices down. Reported beside the HBM-resident figure, never as `value`.
        nv_ref, ni_ref = int(gt["n_verts"]), int(gt["n_inds"])
        hc = torch.as_tensor(coords, dtype=torch.int32).pin_memory()
        hr = torch.full((nch,), n, dtype=torch.int32).pin_memory()
        htot = torch.empty((24,), dtype=torch.uint8).pin_memory()
        hv = torch.empty((max(nv_ref, 1), 3), dtype=torch.float32).pin_memory()
        hi = torch.empty((max(ni_ref, 1),), dtype=inds.dtype).pin_memory()

        def host_step():
            ct.copy_(hc, non_blocking=True)
            rt.copy_(hr, non_blocking=True)
            batched_step()
            htot.copy_(totals, non_blocking=True)
            stream.synchronize()
            tv = htot.numpy().view(np.uint64)
            nv_, ni_ = min(int(tv[0]), nv_ref), min(int(tv[1]), ni_ref)
            hv[:nv_].copy_(verts[:nv_], non_blocking=True)
            hi[:ni_].copy_(inds[:ni_], non_blocking=True)
            stream.synchronize()

        for _ in range(args.warmup):
            host_step()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            host_step()
        ms_host = (time.perf_counter() - t0) * 1e3 / args.steps
        host_bytes = 12 * nch + 4 * nch + 24 + 12 * nv_ref + 4 * ni_ref
        log(f"# batched path with host buffers (PCIe both ways): {ms_host:.4f} ms/step")

        # the reference's call pattern: one plain build per chunk (chunk.rs build_full per chunk)
        n_single = 256

        def per_chunk_calls():
            for i in range(n_single):
                rc = L.uw_build_chunks_plain(perlin.ctx, ctypes.c_void_p(ct.data_ptr() + 12 * i), 1, n,
                                             ctypes.byref(out), ctypes.c_void_p(0), sp)
                if rc:
                    _lib.check(rc, perlin.ctx)
        ms_single = timed(per_chunk_calls, 3, 1) / n_single
        log(f"# plain path, one chunk per call: {ms_single * 1e3:.2f} us/chunk")
        st_b = stage_ms(perlin, batched_step, args.steps)
        st_p = stage_ms(perlin, plain_step, args.steps)
        ms_kernel, ms_noise_plain = st_b[_lib.STAGE_DOMINANT], st_p[_lib.STAGE_DOMINANT]
        log(f"# persistent kernel {ms_kernel:.4f} ms; plain stages {st_p[:4]}")

        # plain-path stage bytes (the algorithmic HBM traffic of the builder stages): compaction
        # reads the densities of every chunk it classifies (4 S) and writes the active list (4 A)
        # and one record (4 B) per chunk; emit writes the mesh once (12 V + 4 I), reads the active
        # list back (8 A: cell and index offset) and one 16-byte header per chunk
        pinfo = np.frombuffer(info.cpu().numpy().tobytes()[: nch * 32], dtype=np.dtype(
            [("flags", "<u4"), ("n_verts", "<u4"), ("n_inds", "<u4"), ("n_active", "<u4"),
             ("vert_off", "<u8"), ("ind_off", "<u8")]))
        classified = int(((pinfo["flags"] & 2) == 0).sum())
        A = int(pinfo["n_active"].sum())
        V, I = int(pinfo["n_verts"].sum()), int(pinfo["n_inds"].sum())
        b_classify = 4 * S * classified + 4 * A + 4 * nch
        b_emit = 12 * V + 4 * I + 8 * A + 16 * nch
        b_noise = 4 * S * nch

        def gbs(b, ms):
            return None if not ms else round(b / (ms / 1e3) / 1e9, 1)

        plain_stages = {
            "k_noise_grid": {"ms": round(ms_noise_plain, 5), "bound": "fp32-valu",
                             "achieved_tflops": round(flops_per_sample(p["octaves"]) * S * nch / (ms_noise_plain / 1e3)
                                                      / 1e12, 3),
                             "hbm_write_bytes": b_noise},
            "k_classify": {"ms": round(st_p[_lib.STAGE_PLAIN_CLASSIFY], 5), "bound": "hbm",
                           "algorithmic_bytes": b_classify, "achieved_gbs": gbs(b_classify, st_p[1]),
                           "frac_of_hbm_peak": round(gbs(b_classify, st_p[1]) / PEAK_HBM_GBS, 4),
                           "traffic": read_traffic("k_classify_emit_chunks12.md")},
            "k_offsets": {"ms": round(st_p[_lib.STAGE_PLAIN_OFFSETS], 5)},
            "k_emit": {"ms": round(st_p[_lib.STAGE_PLAIN_EMIT], 5), "bound": "hbm",
                       "algorithmic_bytes": b_emit, "achieved_gbs": gbs(b_emit, st_p[3]),
                       "frac_of_hbm_peak": round(gbs(b_emit, st_p[3]) / PEAK_HBM_GBS, 4)},
            "note": "classify: 4 S per classified chunk + 4 A + 4 per chunk; emit: 12 V + 4 I + 8 A + 16 per chunk "
                    f"(V={V}, I={I}, A={A}, classified chunks={classified})"}

        flops = flops_per_sample(p["octaves"]) * S * nch
        achieved = flops / (ms_kernel / 1e3) / 1e12
        headline_ms = min(ms_graph, ms_batched)
        headline_path = ("batched (persistent kernel, hipGraph replay)" if ms_graph <= ms_batched
                         else "batched (persistent kernel, stream launch)")
        res.update({
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
            "host_buffers": {"ms_per_step": round(ms_host, 5), "chunks_per_s": round(nch / (ms_host / 1e3), 1),
                             "bytes_moved": host_bytes,
                             "note": "coords + resolutions up from pinned host memory, totals down, then the "
                                     "launch's vertices and indices down; two host synchronisations per step "
                                     "(wall clock). Not the headline value: inputs there are HBM-resident."},
            "batched_setup_ms": round(st_b[_lib.STAGE_BATCH_SETUP], 5),
            "plain_one_chunk_per_call_chunks_per_s": round(1e3 / ms_single, 1),
            "batched_over_one_chunk_per_call": round(ms_single * nch / headline_ms, 1),
            "plain_ms": round(ms_plain, 5),
            "plain_chunks_per_s": round(nch / (ms_plain / 1e3), 1),
            "plain_samples_per_s": round(nch * S / (ms_plain / 1e3), 1),
            "batched_over_plain": round(ms_plain / headline_ms, 3),
            "plain_stages": plain_stages,
            "total_verts": int(gt["n_verts"]), "total_inds": int(gt["n_inds"]), "nonblank": int(gt["n_nonblank"]),
            "roofline": {"bound": "fp32-valu", "kernel": "k_batched (persistent: density + mesh)",
                         "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                         "traffic": read_traffic("k_batched_chunks12.md"),
                         "algorithmic_flop_per_launch": flops, "ms_per_launch": round(ms_kernel, 5),
                         "plain_density_kernel": {"ms_per_launch": round(ms_noise_plain, 5),
                                                  "achieved_tflops": round(flops / (ms_noise_plain / 1e3) / 1e12, 3)}},
        })

    # ---- mixed-resolution workload: 8192 chunks, 6 octaves ----
    if run("mixed") and not args.no_mixed:
        pm = params.preset("mixed")
        mc, mr = params.mixed_workload(8192, 7)
        pmix = Perlin(pm)
        mv, mi = batch_capacity(mr)
        g2 = BatchedGraph(pmix, mc, mr, vert_cap=mv, ind_cap=mi)
        ms_mixed = timed(lambda: g2.launch(stream.cuda_stream), max(3, args.steps // 4), 1)
        mt = g2.batch.totals()
        # per-stage kernel times of the same build launched eagerly (a graph records no timers)
        st_m = stage_ms(pmix, lambda: build_chunks_batched(pmix, mc, mr, vert_cap=mv, ind_cap=mi,
                                                           stream=stream.cuda_stream), 3)
        mixed_stages = {name: round(st_m[i], 4) for name, i in (
            ("setup", _lib.STAGE_BATCH_SETUP), ("bulk_front", _lib.STAGE_BULK_FRONT),
            ("persistent", _lib.STAGE_DOMINANT), ("bulk_emit", _lib.STAGE_BULK_EMIT)) if st_m[i] is not None}
        samples = int(sum((r + 1) ** 3 for r in mr))
        flops_m = flops_per_sample(pm["octaves"]) * samples
        res["mixed"] = {"workload": "8192 chunks, 2048 each at 8^3/12^3/24^3/32^3 cells, coords uniform in 32x32x8 "
                                    "(SplitMix64 seed 7), 6 octaves",
                        "ms_per_step": round(ms_mixed, 4), "chunks_per_s": round(8192 / (ms_mixed / 1e3), 1),
                        "samples_per_s": round(samples / (ms_mixed / 1e3), 1), "samples": samples,
                        "algorithmic_tflops": round(flops_m / (ms_mixed / 1e3) / 1e12, 3),
                        "nonblank_chunks": int(mt["n_nonblank"]), "verts": int(mt["n_verts"]),
                        "inds": int(mt["n_inds"]), "eager_stages_ms": mixed_stages,
                        "bulk_min": get_batch_tuning(pmix)["bulk_min"]}
        log(f"# mixed workload: {ms_mixed:.4f} ms/step")
        g2.close()
        del g2

    # ---- adjoint: 256 chunks of 24^3, forward with tape + backward (densities, amplitudes, offset) ----
    if run("adjoint") and not args.no_adjoint:
        from underwaterworld_amd.adjoint import ChunkMesher
        pa = params.preset("adjoint24")
        ac = [(x, y, z) for x in range(-4, 4) for y in range(-4, 4) for z in range(-3, 1)]
        pad = Perlin(pa)
        mesher = ChunkMesher(pad, ac, 24, want_dens_grad=True)
        rec = mesher.forward()
        torch.cuda.synchronize()
        nvt = int(rec.batch.totals()["n_verts"])
        gv = torch.zeros((mesher.vert_cap, 3), dtype=torch.float32, device=dev)
        gv[:nvt, 2] = 1.0

        def adj_step():
            r = mesher.forward(stream.cuda_stream)
            mesher.backward(gv, stream.cuda_stream, record=r)

        ms_adj = timed(adj_step, max(10, args.steps // 2), 2)
        st_f = stage_ms(pad, lambda: mesher.forward(stream.cuda_stream), 3)
        st_a = stage_ms(pad, lambda: mesher.backward(gv, stream.cuda_stream, record=rec), 3)
        # the forward's launches: queue set-up, the bulk path's front and emission (24^3 chunks leave
        # the persistent queue when the tuning's bulk_min allows) and the persistent kernel
        fwd_stages = {name: round(st_f[i], 4) for name, i in (
            ("setup", _lib.STAGE_BATCH_SETUP), ("bulk_front", _lib.STAGE_BULK_FRONT),
            ("persistent", _lib.STAGE_DOMINANT), ("bulk_emit", _lib.STAGE_BULK_EMIT)) if st_f[i] is not None}
        fwd_ms = sum(fwd_stages.values())
        res["adjoint"] = {"workload": "256 chunks (x,y in [-4,4), z in [-3,1)) x 24^3 cells, 6 octaves, seed 42; "
                                      "forward with vertex tape + backward to 25^3 densities, 6 amplitudes, 3 offsets",
                          "ms_per_step": round(ms_adj, 4), "grads_per_s": round(len(ac) / (ms_adj / 1e3), 1),
                          "forward_kernels_ms": round(fwd_ms, 4), "forward_stages_ms": fwd_stages,
                          "backward_kernel_ms": round(st_a[_lib.STAGE_DOMINANT], 4), "vertices": nvt,
                          "forward_algorithmic_tflops": round(
                              flops_per_sample(pa["octaves"]) * 25 ** 3 * len(ac) / (fwd_ms / 1e3) / 1e12, 3)}
        log(f"# adjoint: {ms_adj:.4f} ms/step (forward + backward)")
        del mesher, rec, gv

    # ---- volume: one 512^3-cell grid (513^3 samples), 8 octaves; forward + adjoint, device memory ----
    if run("volume") and not args.no_volume:
        from underwaterworld_amd.adjoint import VolumeMesher
        pv = params.preset("volume512")
        nvox = 512
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        free0, total_mem = torch.cuda.mem_get_info(dev)
        pvol = Perlin(pv)
        vm = VolumeMesher(pvol, nvox, want_dens_grad=True)
        rec = vm.forward(stream.cuda_stream)
        torch.cuda.synchronize()
        vinfo = rec.batch.info()[0]
        nvv = int(vinfo["n_verts"])
        gvv = torch.zeros((vm.vert_cap, 3), dtype=torch.float32, device=dev)
        gvv[:nvv, 2] = 1.0
        steps_v = 3 if wl == "all" else max(2, args.steps // 4)
        ms_vfwd = timed(lambda: vm.forward(stream.cuda_stream), steps_v, 1)

        def vol_step():
            r = vm.forward(stream.cuda_stream)
            vm.backward(gvv, stream.cuda_stream, record=r)

        ms_vadj = timed(vol_step, steps_v, 1)
        st_vf = stage_ms(pvol, lambda: vm.forward(stream.cuda_stream), 2)
        st_vb = stage_ms(pvol, lambda: vm.backward(gvv, stream.cuda_stream, record=rec), 2)
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info(dev)
        used = free0 - free1
        samples_v = (nvox + 1) ** 3
        o_v = int(pv["octaves"])
        scratch = int(L.uw_volume_scratch_bytes(pvol.ctx, nvox))
        dens_ms = st_vf[_lib.STAGE_DOMINANT]
        # stored-field alternative to recomputation: the backward would keep, per sample and octave,
        # the noise value and its gradient (4 floats) -- an estimate from that storage, not a run
        stored_oct = samples_v * o_v * 4 * 4
        res["volume"] = {
            "workload": "one 512^3-cell grid (513^3 samples, one index space), seed 1234, 8 octaves, "
                        "linear z-bias (0.01, -0.05); forward + adjoint to densities, amplitudes, offset",
            "forward_ms": round(ms_vfwd, 3), "forward_plus_adjoint_ms": round(ms_vadj, 3),
            "samples": samples_v, "samples_per_s_forward": round(samples_v / (ms_vfwd / 1e3), 1),
            "stages_ms": {"tables": st_vf[_lib.STAGE_VOLUME_TABLES], "density": dens_ms,
                          "count": st_vf[_lib.STAGE_VOLUME_COUNT], "scan": st_vf[_lib.STAGE_VOLUME_SCAN],
                          "emit": st_vf[_lib.STAGE_VOLUME_EMIT], "backward": st_vb[_lib.STAGE_DOMINANT]},
            "density_algorithmic_tflops": round(flops_per_sample(o_v) * samples_v / (dens_ms / 1e3) / 1e12, 3),
            "verts": nvv, "inds": int(vinfo["n_inds"]), "active_cells": int(vinfo["n_active"]),
            "active_fraction": round(int(vinfo["n_active"]) / nvox ** 3, 4),
            "device_memory_used_bytes": int(used), "library_scratch_bytes": scratch,
            "device_memory_without_recompute_estimate_bytes": int(used + stored_oct),
            "note": "device memory used = hipMemGetInfo before the workload minus after the forward and adjoint "
                    "steps (library scratch: tables, density field, row counts; torch's cached outputs, tape and "
                    "gradients). Without recomputation the backward would also keep each sample's per-octave "
                    "noise value and gradient (16 B per sample per octave): that figure is an estimate."}
        log(f"# volume: forward {ms_vfwd:.3f} ms, forward + adjoint {ms_vadj:.3f} ms, "
            f"device memory {used / 2**20:.1f} MiB, stages {res['volume']['stages_ms']}")
        del vm, rec, gvv

    if wl != "all":
        print(json.dumps({k: v for k, v in res.items()}), flush=True)
        return

    cpu = cpu_all = parity = None
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

        hc = host_cpu()
        k, dt = timed_passes(1, 10.0)
        cpu = {"value": round(k * nch / dt, 2), "unit": "chunks/s", "cores": 1, "kind": "port",
               "sample": f"{k} passes over all {nch} chunks of the workload ({k * nch} chunk builds), "
                         "reference-faithful C restatement (f64 noise, linear-search dedup), 1 thread",
               "seconds": round(dt, 3), "host_cpu": hc}
        # every host thread this process may run on; OMP_NUM_THREADS, when set, is the host's CPU share
        nthr = hc["affinity"]
        if hc["omp_num_threads"]:
            nthr = max(1, min(nthr, int(hc["omp_num_threads"])))
        k, dt = timed_passes(nthr, 2.0)
        cpu_all = {"value": round(k * nch / dt, 2), "unit": "chunks/s", "cores": nthr, "kind": "port",
                   "sample": f"{k} passes over all {nch} chunks, same build, chunk-parallel over {nthr} threads "
                             f"(nproc {hc['nproc']}, affinity {hc['affinity']}, OMP_NUM_THREADS "
                             f"{hc['omp_num_threads']})",
                   "seconds": round(dt, 3)}
        parity = parity_vs_oracle(O, p, coords, n, perlin, np, torch)

    res["cpu_baseline"] = cpu
    res["cpu_baseline_all_cores"] = cpu_all
    res["parity"] = parity
    print(json.dumps(res), flush=True)


def parity_vs_oracle(O, p, coords, n, perlin, np, torch):
    """The GPU output of the headline workload against the oracle: densities against its
    reference-faithful (f64 noise, f32 density) and FP64 variants; cube-case bits that differ (all
    must sit at corners within 1e-6 of the isovalue: ties); vertex positions, in cell edges, against
    the reference-faithful mesher (all vertices, and those on edges whose densities differ by at
    least 0.1) and against the FP64 mesher. cfg1 is chunk (0, 0, -1) of the same grid."""
    from underwaterworld_amd import build_chunks_plain, sample_density
    from underwaterworld_amd.adjoint import ChunkMesher
    nch = len(coords)
    S, C = (n + 1) ** 3, n ** 3
    dens, _ = sample_density(perlin, coords, n)
    pb = build_chunks_plain(perlin, coords, n, want_cases=True)
    mesher = ChunkMesher(perlin, coords, n)
    rec = mesher.forward()
    torch.cuda.synchronize()
    d_gpu = dens.cpu().numpy()
    cases = pb.cases.cpu().numpy().reshape(nch, C)
    pinfo = pb.info()
    pverts, pinds = pb.verts.cpu().numpy(), pb.inds.cpu().numpy().view(np.uint32)
    tinfo = rec.batch.info()
    tape = rec.tape.cpu().numpy()
    iso = float(np.float32(p["iso"]))
    h = float(np.float32(p["chunk_extent"]) / np.float32(n))
    corners = [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)]

    def case_of(dd):
        D = dd.reshape(n + 1, n + 1, n + 1) < np.float32(iso)
        c = np.zeros((n, n, n), np.uint8)
        for i, (a, b, e) in enumerate(corners):
            c |= (D[a:a + n, b:b + n, e:e + n].astype(np.uint8) << i)
        return c.ravel()

    def acc():
        return {"max_drho_ref": 0.0, "max_drho_fp64": 0.0, "tie_corners": 0, "case_bits_differ_ref": 0,
                "case_bits_differ_fp64": 0, "chunks_with_ties": 0, "early_skip_differs": 0,
                "surface_chunks_compared": 0, "index_buffers_equal": 0,
                "max_dp_over_h_ref": 0.0, "max_dp_over_h_ref_well_conditioned": 0.0,
                "ill_conditioned_vertices": 0, "max_dp_over_h_fp64": 0.0, "vertices_compared": 0}

    out = {"cfg1": acc(), "cfg2": acc()}
    k1 = int(np.nonzero((coords == np.array([0, 0, -1])).all(1))[0][0])
    for i in range(nch):
        c = tuple(int(v) for v in coords[i])
        dref = O.density_ref(p, c, n)
        d64 = O.density_fp64(p, c, n)
        dg = d_gpu[i].astype(np.float64)
        r = {"max_drho_ref": float(np.max(np.abs(dg - dref))), "max_drho_fp64": float(np.max(np.abs(dg - d64)))}
        ties = int(np.sum(np.abs(dref - iso) <= 1e-6))
        per = dict(r, tie_corners=ties)
        ni = int(pinfo[i]["n_inds"])
        early_gpu = bool(pinfo[i]["flags"] & 2)  # no cases are written for an early-skipped chunk
        early_ref = bool(O.early_blank(dref.astype(np.float32), p["iso"]))
        bits_ref = 0
        if early_gpu or early_ref:
            per["early_skip_differs"] = int(early_gpu != early_ref)
        else:
            cg = cases[i]
            cr = case_of(dref.astype(np.float32))
            c64 = (d64.reshape(n + 1, n + 1, n + 1) < iso)
            c64v = np.zeros((n, n, n), np.uint8)
            for b, (x0, y0, z0) in enumerate(corners):
                c64v |= (c64[x0:x0 + n, y0:y0 + n, z0:z0 + n].astype(np.uint8) << b)
            bits_ref = int(np.unpackbits((cg ^ cr)[:, None], axis=1).sum())
            bits_64 = int(np.unpackbits((cg ^ c64v.ravel())[:, None], axis=1).sum())
            per.update(case_bits_differ_ref=bits_ref, case_bits_differ_fp64=bits_64)
        if not early_gpu and not early_ref and bits_ref == 0 and ni > 0:
            mr = O.mesh_ref(p, c, n, dref.astype(np.float32))
            v0, nv = int(pinfo[i]["vert_off"]), int(pinfo[i]["n_verts"])
            i0 = int(pinfo[i]["ind_off"])
            per["surface_chunks_compared"] = 1
            per["index_buffers_equal"] = int(np.array_equal(pinds[i0:i0 + ni], mr["inds"]))
            vg = pverts[v0:v0 + nv].astype(np.float64)
            e = np.max(np.abs(vg - mr["verts"]), axis=1) / h
            t0 = int(tinfo[i]["vert_off"])
            tp = tape[t0:t0 + nv]
            gap = np.abs(tp[:, 3].view(np.float32).astype(np.float64) - tp[:, 2].view(np.float32))
            well = gap >= 0.1
            per["max_dp_over_h_ref"] = float(e.max())
            per["max_dp_over_h_ref_well_conditioned"] = float(e[well].max()) if well.any() else 0.0
            per["ill_conditioned_vertices"] = int(np.sum((e > 1e-5) & ~well))
            per["vertices_compared"] = nv
            if bits_64 == 0:
                m64 = O.mesh_fp64(p, c, n, d64)
                if len(m64["verts"]) == nv:
                    per["max_dp_over_h_fp64"] = float(np.max(np.abs(vg - m64["verts"])) / h)
        for key, a in (("cfg2", out["cfg2"]),) + ((("cfg1", out["cfg1"]),) if i == k1 else ()):
            for f in ("max_drho_ref", "max_drho_fp64", "max_dp_over_h_ref", "max_dp_over_h_ref_well_conditioned",
                      "max_dp_over_h_fp64"):
                a[f] = max(a[f], per.get(f, 0.0))
            for f in ("tie_corners", "case_bits_differ_ref", "case_bits_differ_fp64", "ill_conditioned_vertices",
                      "vertices_compared", "early_skip_differs", "surface_chunks_compared", "index_buffers_equal"):
                a[f] += per.get(f, 0)
            a["chunks_with_ties"] += int(ties > 0)
    for a in out.values():
        for f, v in list(a.items()):
            if isinstance(v, float):
                a[f] = float(f"{v:.3g}")
    out["note"] = ("GPU (plain path densities, cases and meshes; batched tape for the edge density gaps) against "
                   "the oracle. Case bits that differ from the reference-faithful oracle must sit at tie corners; "
                   "positions are compared in chunks without differing case bits. Bar: |drho| <= 1e-6, "
                   "|dp|/h <= 1e-5, or within 1e-5 + 2e-6 / |rho_b - rho_a| on edges with |rho_b - rho_a| < 0.1 "
                   "(ill-conditioned). Parity with the noise crate itself is unpinned.")
    return out


if __name__ == "__main__":
    main()

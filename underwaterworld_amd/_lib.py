"""ctypes binding of the engine's C ABI (include/uwmesh.h) -> underwaterworld_amd/libuwmesh.so.

The library is loaded from the package directory only; there is no other implementation to fall
back on, so a missing library is an error.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libuwmesh.so")

MAX_OCTAVES = 16
MAX_RESOLUTION = 64

UW_OK = 0
UW_ERR_INVALID_ARG = -1
UW_ERR_OVERFLOW = -2
UW_ERR_HIP = -3
UW_ERR_NO_DEVICE = -4
UW_ERR_OOM = -5

CHUNK_BLANK = 1
CHUNK_EARLY_SKIP = 2
CHUNK_OVERFLOW = 4
CHUNK_INDEX_WRAP = 16

P = ctypes.POINTER
c_void_p = ctypes.c_void_p


class UwParams(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint32),
        ("octaves", ctypes.c_int32),
        ("frequency", ctypes.c_double),
        ("lacunarity", ctypes.c_double),
        ("amplitudes", ctypes.c_double * MAX_OCTAVES),
        ("iso", ctypes.c_float),
        ("bias_kind", ctypes.c_int32),
        ("bias_a", ctypes.c_float),
        ("bias_b", ctypes.c_float),
        ("bias_c", ctypes.c_float),
        ("chunk_extent", ctypes.c_int32),
        ("offset", ctypes.c_double * 3),
    ]


class UwChunkInfo(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_uint32), ("n_verts", ctypes.c_uint32), ("n_inds", ctypes.c_uint32),
                ("n_active", ctypes.c_uint32), ("vert_off", ctypes.c_uint64), ("ind_off", ctypes.c_uint64)]


class UwTotals(ctypes.Structure):
    _fields_ = [("n_verts", ctypes.c_uint64), ("n_inds", ctypes.c_uint64), ("n_nonblank", ctypes.c_uint32),
                ("overflow", ctypes.c_uint32)]


BATCH_MAX_RESOLUTION = 32
VOLUME_MAX_RESOLUTION = 640
NOALLOC_SCRATCH_BYTES = 32 << 20
# enum uw_profile_stage
STAGE_DOMINANT, STAGE_PLAIN_CLASSIFY, STAGE_PLAIN_OFFSETS, STAGE_PLAIN_EMIT, STAGE_BATCH_SETUP = 0, 1, 2, 3, 4
STAGE_VOLUME_COUNT, STAGE_VOLUME_SCAN, STAGE_VOLUME_EMIT, STAGE_VOLUME_TABLES = 5, 6, 7, 8
STAGE_BULK_FRONT, STAGE_BULK_EMIT, STAGE_COUNT = 9, 10, 11
CHUNK_INVALID = 8


class UwBatchTuning(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("grid", ctypes.c_int32), ("max_res", ctypes.c_int32),
                ("slab", ctypes.c_int8 * (BATCH_MAX_RESOLUTION + 1)), ("bulk_min", ctypes.c_int32)]


class UwMeshOut(ctypes.Structure):
    _fields_ = [("verts", c_void_p), ("inds", c_void_p), ("info", c_void_p), ("totals", c_void_p),
                ("vert_cap", ctypes.c_uint64), ("ind_cap", ctypes.c_uint64), ("vertex_format", ctypes.c_uint32),
                ("index_format", ctypes.c_uint32), ("tri_normals", c_void_p), ("tri_cells", c_void_p)]


class UwHostOut(ctypes.Structure):
    _fields_ = [("verts", c_void_p), ("vert_cap", ctypes.c_uint64), ("inds", c_void_p), ("ind_cap", ctypes.c_uint64),
                ("vertex_format", ctypes.c_uint32), ("index_format", ctypes.c_uint32), ("tri_normals", c_void_p),
                ("tri_cells", c_void_p)]


class UwWorldConfig(ctypes.Structure):
    _fields_ = [("view_dist", ctypes.c_int32), ("generation_dist", ctypes.c_int32), ("keep_dist", ctypes.c_int32),
                ("max_z", ctypes.c_int32), ("min_z", ctypes.c_int32), ("recheck_dist", ctypes.c_float),
                ("recheck_angle", ctypes.c_float), ("resolution", ctypes.c_int32), ("max_batch", ctypes.c_int32),
                ("vertex_format", ctypes.c_uint32), ("index_format", ctypes.c_uint32), ("collision", ctypes.c_int32)]


class UwWorldCandidate(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_int32 * 3), ("dist", ctypes.c_float), ("z", ctypes.c_float),
                ("in_view", ctypes.c_int32), ("in_gen", ctypes.c_int32)]


class UwWorldChunk(ctypes.Structure):
    _fields_ = [("info", UwChunkInfo), ("verts", c_void_p), ("inds", c_void_p), ("tri_normals", c_void_p),
                ("tri_cells", c_void_p)]


# name -> (restype, argtypes); every symbol include/uwmesh.h declares.
SIGNATURES = {
    "uw_params_reference": (None, [P(UwParams), ctypes.c_uint32]),
    "uw_ctx_create": (ctypes.c_int, [P(UwParams), ctypes.c_int, P(c_void_p)]),
    "uw_ctx_destroy": (None, [c_void_p]),
    "uw_last_error": (ctypes.c_char_p, [c_void_p]),
    "uw_ctx_permutation": (ctypes.c_int, [c_void_p, P(ctypes.c_uint8)]),
    "uw_permutation_table": (None, [ctypes.c_uint32, P(ctypes.c_uint8)]),
    "uw_profile_enable": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "uw_profile_read": (ctypes.c_int, [c_void_p, P(ctypes.c_double), P(ctypes.c_uint64)]),
    "uw_profile_read_stages": (ctypes.c_int, [c_void_p, ctypes.c_int32, P(ctypes.c_double), P(ctypes.c_uint64)]),
    "uw_max_verts": (ctypes.c_uint64, [ctypes.c_int32]),
    "uw_max_inds": (ctypes.c_uint64, [ctypes.c_int32]),
    "uw_sample_density": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int32, ctypes.c_int32, c_void_p, c_void_p,
                                         c_void_p]),
    "uw_build_chunks_plain": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int32, ctypes.c_int32, P(UwMeshOut),
                                             c_void_p, c_void_p]),
    "uw_chunk_build_host": (ctypes.c_int, [c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           c_void_p, ctypes.c_uint64, c_void_p, ctypes.c_uint64, P(UwChunkInfo)]),
    "uw_chunk_build_host_out": (ctypes.c_int, [c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, P(UwHostOut), P(UwChunkInfo)]),
    "uw_tris_around": (ctypes.c_int, [c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_int32, c_void_p, ctypes.c_uint32,
                                      P(ctypes.c_uint32)]),
    "uw_get_batch_tuning": (ctypes.c_int, [c_void_p, P(UwBatchTuning)]),
    "uw_set_batch_tuning": (ctypes.c_int, [c_void_p, P(UwBatchTuning)]),
    "uw_build_chunks_batched": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int32, P(UwMeshOut), c_void_p]),
    "uw_build_chunks_batched_noalloc": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int32, P(UwMeshOut),
                                                       c_void_p]),
    "uw_batched_graph_create": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int32, P(UwMeshOut),
                                               P(c_void_p)]),
    "uw_graph_launch": (ctypes.c_int, [c_void_p, c_void_p]),
    "uw_ctx_set_field": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "uw_density_at": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int64, c_void_p, c_void_p]),
    "uw_build_chunks_batched_tape": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int32, P(UwMeshOut),
                                                    c_void_p, c_void_p]),
    "uw_adjoint_backward": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int32, c_void_p, c_void_p,
                                           c_void_p, c_void_p, ctypes.c_int64, c_void_p, c_void_p, c_void_p]),
    "uw_graph_destroy": (None, [c_void_p]),
    "uw_build_volume": (ctypes.c_int, [c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       P(UwMeshOut), c_void_p, c_void_p]),
    "uw_volume_backward": (ctypes.c_int, [c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "uw_volume_scratch_bytes": (ctypes.c_uint64, [c_void_p, ctypes.c_int32]),
    "uw_world_config_reference": (None, [P(UwWorldConfig)]),
    "uw_world_plan": (ctypes.c_int, [P(UwWorldConfig), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     ctypes.c_int32, c_void_p, ctypes.c_int32, P(ctypes.c_int32), P(UwWorldCandidate),
                                     ctypes.c_int32, P(ctypes.c_int32)]),
    "uw_world_create": (ctypes.c_int, [c_void_p, P(UwWorldConfig), P(c_void_p)]),
    "uw_world_destroy": (None, [c_void_p]),
    "uw_world_update": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int,
                                       P(ctypes.c_int32)]),
    "uw_world_generate_count": (ctypes.c_int32, [c_void_p]),
    "uw_world_render_count": (ctypes.c_int32, [c_void_p]),
    "uw_world_total_count": (ctypes.c_int32, [c_void_p]),
    "uw_world_chunks_to_render": (ctypes.c_int32, [c_void_p, c_void_p, ctypes.c_int32]),
    "uw_world_get_chunk": (ctypes.c_int, [c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          P(UwWorldChunk)]),
    "uw_world_read_chunk": (ctypes.c_int, [c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, P(UwHostOut),
                                           P(UwChunkInfo)]),
}

_lib = None


def lib():
    """Load libuwmesh.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("underwaterworld_amd/libuwmesh.so is not built: run "
                               "`python -m underwaterworld_amd._build` (or __graft_entry__.build())")
        # torch bundles its own HIP runtime under the same soname (libamdhip64.so.7). Loading torch
        # first makes this library bind to that runtime, so the process has one HIP runtime and
        # torch's streams, allocations and synchronisation cover this library's launches. (Loaded
        # the other way round, a second runtime comes up and torch finds no GPU.)
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def make_params(p: dict) -> UwParams:
    s = UwParams()
    s.seed = int(p["seed"]) & 0xFFFFFFFF
    s.octaves = int(p["octaves"])
    s.frequency = float(p["frequency"])
    s.lacunarity = float(p["lacunarity"])
    amps = list(p["amplitudes"])
    for k in range(MAX_OCTAVES):
        s.amplitudes[k] = float(amps[k]) if k < len(amps) else 0.0
    s.iso = float(p["iso"])
    s.bias_kind = int(p["bias_kind"])
    s.bias_a, s.bias_b, s.bias_c = (float(v) for v in p["bias"])
    s.chunk_extent = int(p["chunk_extent"])
    for i, v in enumerate(p.get("offset", (0.0, 0.0, 0.0))):
        s.offset[i] = float(v)
    return s


class UwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("uwmesh error %d: %s" % (code, msg))
        self.code = code


def check(rc, ctx=None):
    if rc != UW_OK:
        msg = lib().uw_last_error(ctx).decode() if ctx else ""
        raise UwError(rc, msg)

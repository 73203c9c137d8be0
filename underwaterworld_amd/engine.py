"""Host-side mirror of the reference chunk-builder interface, over the C ABI.

    Perlin(seed, params)            noise::Perlin::new(seed)            state.rs:358-359
    Chunk(pos)                      Chunk::new(pos)                     chunk.rs:89-103
    Chunk.build_full(perlin)        Chunk::build_full                   chunk.rs:266-268
    Chunk.build_partial(perlin)     Chunk::build_partial -> finished    chunk.rs:270-313
    Chunk.not_blank()               Chunk::not_blank                    chunk.rs:344
    Chunk.verts_buffer_slice()      Chunk::verts_buffer_slice (xyz)     chunk.rs:346
    Chunk.inds_buffer_slice()       Chunk::inds_buffer_slice (u32)      chunk.rs:347
    Chunk.num_inds()                Chunk::num_inds                     chunk.rs:348

and the device-resident multi-chunk forms (torch tensors on the context's device):

    build_chunks_plain(perlin, coords, n)   one launch per builder stage over all chunks
    sample_density(perlin, coords, n)       iso_at over every sample of every chunk

The reference's `device: &wgpu::Device` argument is the renderer's upload target; the engine's
output buffers are HIP device memory (or host arrays for the Chunk facade), so it is accepted and
ignored for signature parity.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import params as _params


def _stream_ptr(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(int(stream))


class Perlin:
    """Density-field context: the permutation table of `seed` plus the field parameters, uploaded
    to one GPU (noise::Perlin::new, state.rs:358-359)."""

    def __init__(self, params: dict | None = None, seed: int | None = None, device: int = 0,
                 apply_tuning: bool = True):
        p = dict(params) if params is not None else _params.preset("reference")
        if seed is not None:
            p["seed"] = int(seed)
        self.params = p
        self.device = device
        self._ctx = ctypes.c_void_p()
        s = _lib.make_params(p)
        rc = _lib.lib().uw_ctx_create(ctypes.byref(s), device, ctypes.byref(self._ctx))
        if rc != _lib.UW_OK:
            raise _lib.UwError(rc, "uw_ctx_create failed (parameters or device)")
        if apply_tuning:
            t = tuned_shape(int(p["octaves"]))
            if t is not None:
                set_batch_tuning(self, t)

    @classmethod
    def new(cls, seed: int):
        """noise::Perlin::new(seed) with the reference crate's constants."""
        return cls(None, seed=seed)

    @property
    def ctx(self):
        return self._ctx

    def field(self):
        """(amplitudes float64 (octaves,), offset float64 (3,)) of the context's current field."""
        p = self.params
        return (np.array(p["amplitudes"][: int(p["octaves"])], np.float64),
                np.array(p.get("offset", (0.0, 0.0, 0.0)), np.float64))

    def set_field(self, amplitudes=None, offset=None):
        """Replace the octave amplitudes and/or the world-space sample offset (uw_ctx_set_field);
        None leaves a part unchanged. Later builds on this context use the new field."""
        a = None if amplitudes is None else np.ascontiguousarray(amplitudes, np.float64).ravel()
        o = None if offset is None else np.ascontiguousarray(offset, np.float64).ravel()
        if a is not None and len(a) != int(self.params["octaves"]):
            raise ValueError(f"expected {self.params['octaves']} amplitudes, got {len(a)}")
        if o is not None and len(o) != 3:
            raise ValueError("offset needs 3 values")
        rc = _lib.lib().uw_ctx_set_field(self._ctx, a.ctypes.data_as(ctypes.c_void_p) if a is not None else None,
                                         o.ctypes.data_as(ctypes.c_void_p) if o is not None else None)
        _lib.check(rc, self._ctx)
        if a is not None:
            self.params["amplitudes"] = [float(v) for v in a]
        if o is not None:
            self.params["offset"] = tuple(float(v) for v in o)

    def permutation(self) -> np.ndarray:
        out = np.zeros(256, np.uint8)
        _lib.check(_lib.lib().uw_ctx_permutation(self._ctx, out.ctypes.data_as(_lib.P(ctypes.c_uint8))), self._ctx)
        return out

    def close(self):
        if self._ctx:
            _lib.lib().uw_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_TUNING = None


def tuned_shape(octaves: int):
    """The autotuner's launch shape for this octave count (underwaterworld_amd/tuning.json)."""
    global _TUNING
    if _TUNING is None:
        import json
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
        _TUNING = json.load(open(path)) if os.path.exists(path) else {}
    return _TUNING.get("by_octaves", {}).get(str(octaves))


def permutation_table(seed: int) -> np.ndarray:
    """Host-only PermutationTable::new(seed)."""
    out = np.zeros(256, np.uint8)
    _lib.lib().uw_permutation_table(int(seed) & 0xFFFFFFFF, out.ctypes.data_as(_lib.P(ctypes.c_uint8)))
    return out


def tri_ids_around(tri_cells: np.ndarray, n: int, local_pos_percent, range_: int) -> np.ndarray:
    """uw_tris_around over one chunk's tri_cells: the chunk-local ids of the triangles that
    Chunk::tris_around (chunk.rs:315-342) returns, in its order."""
    tc = np.ascontiguousarray(tri_cells, np.uint32)
    px, py, pz = (float(v) for v in local_pos_percent)
    L = _lib.lib()
    count = ctypes.c_uint32(0)
    cap = len(tc)
    ids = np.empty(max(cap, 1), np.uint32)
    rc = L.uw_tris_around(tc.ctypes.data, len(tc), int(n), px, py, pz, int(range_), ids.ctypes.data, cap,
                          ctypes.byref(count))
    _lib.check(rc)
    return ids[: count.value].copy()


class Tri:
    """Collision triangle (util::Tri, util.rs:7-20): verts (3, 3) f32 world positions, normal (3,)."""
    __slots__ = ("verts", "normal")

    def __init__(self, verts, normal):
        self.verts = verts
        self.normal = normal


class Chunk:
    """One chunk, built through the engine in a single call (chunk.rs:80-349).

    colors=True gives the reference's VertColor vertices (6 f32: position, sRGB colour); u16=True
    its u16 index buffer (chunk.rs:243). Collision triangles (Build.tris) are always built, for
    tris_around."""

    def __init__(self, pos, resolution: int | None = None, colors: bool = False, u16: bool = False):
        self.pos = tuple(int(v) for v in pos)
        self.resolution = resolution
        self.colors = colors
        self.u16 = u16
        self._verts = None
        self._inds = None
        self._info = None
        self._normals = None
        self._tri_cells = None
        self._n = None
        self._done = False

    @classmethod
    def new(cls, pos):
        return cls(pos)

    def build_full(self, perlin: Perlin, device=None):
        while not self.build_partial(perlin, device):
            pass

    def build_partial(self, perlin: Perlin, device=None) -> bool:
        """The reference slices a build over several frames (chunk.rs:19-20,270-313); the GPU
        build completes in one call, which returns True ("finished")."""
        if self._done:
            return True
        n = int(self.resolution or perlin.params.get("resolution", 12))
        L = _lib.lib()
        vcap, icap = int(L.uw_max_verts(n)), int(L.uw_max_inds(n))
        verts = np.empty((vcap, 6 if self.colors else 3), np.float32)
        inds = np.empty(icap, np.uint16 if self.u16 else np.uint32)
        normals = np.empty((icap // 3, 3), np.float32)
        cells = np.empty(icap // 3, np.uint32)
        ho = _lib.UwHostOut(verts.ctypes.data, vcap, inds.ctypes.data, icap,
                            VERTEX_XYZ_RGB if self.colors else VERTEX_XYZ, INDEX_U16 if self.u16 else INDEX_U32,
                            normals.ctypes.data, cells.ctypes.data)
        info = _lib.UwChunkInfo()
        rc = L.uw_chunk_build_host_out(perlin.ctx, self.pos[0], self.pos[1], self.pos[2], n, ctypes.byref(ho),
                                       ctypes.byref(info))
        _lib.check(rc, perlin.ctx)
        self._info = info
        self._n = n
        if info.n_inds > 0:
            nt = info.n_inds // 3
            self._verts = verts[: info.n_verts].copy()
            self._inds = inds[: info.n_inds].copy()
            self._normals = normals[:nt].copy()
            self._tri_cells = cells[:nt].copy()
        self._done = True
        return True

    def tris_around(self, local_pos_percent, range_: int) -> list:
        """Chunk::tris_around (chunk.rs:315-342): collision triangles of the cells within range_
        cells of local_pos_percent (chunk-local position / chunk extent), in the reference's order."""
        if self._tri_cells is None:
            return []
        ids = tri_ids_around(self._tri_cells, self._n, local_pos_percent, range_)
        pos = self._verts[:, :3]
        tri = pos[self._inds.astype(np.int64)].reshape(-1, 3, 3)
        return [Tri(tri[t], self._normals[t]) for t in ids]

    def collision(self):
        """(tri_cells (T,) u32, normals (T, 3) f32) of the chunk's collision triangles."""
        return self._tri_cells, self._normals

    def not_blank(self) -> bool:
        return self._verts is not None

    def early_skipped(self) -> bool:
        return bool(self._info and self._info.flags & _lib.CHUNK_EARLY_SKIP)

    def verts_buffer_slice(self) -> np.ndarray:
        # the reference's accessor panics on a blank chunk (chunk.rs:345-346)
        if self._verts is None:
            raise RuntimeError("verts_buffer_slice() on a blank chunk")
        return self._verts

    def inds_buffer_slice(self) -> np.ndarray:
        if self._inds is None:
            raise RuntimeError("inds_buffer_slice() on a blank chunk")
        return self._inds

    def num_inds(self) -> int:
        return int(self._info.n_inds) if self._info is not None else 0

    def num_active(self) -> int:
        return int(self._info.n_active) if self._info is not None else 0


# ---------------------------------------------------------------------------------------------
# Device-resident multi-chunk forms.

INFO_DTYPE = np.dtype([("flags", "<u4"), ("n_verts", "<u4"), ("n_inds", "<u4"), ("n_active", "<u4"),
                       ("vert_off", "<u8"), ("ind_off", "<u8")])
VERTEX_XYZ = 0      # 3 floats per vertex: world position
VERTEX_XYZ_RGB = 1  # 6 floats: position + sRGB colour, the reference's VertColor (draw.rs:4-9)
INDEX_U32 = 0
INDEX_U16 = 1       # the reference's index buffer type (chunk.rs:243)
TOTALS_DTYPE = np.dtype([("n_verts", "<u8"), ("n_inds", "<u8"), ("n_nonblank", "<u4"), ("overflow", "<u4")])


@dataclass
class MeshBatch:
    """Output of a multi-chunk build: device tensors plus host copies of the small records."""
    verts: "torch.Tensor"   # (vert_cap, 3) f32 xyz, or (vert_cap, 6) xyz + sRGB colour
    inds: "torch.Tensor"    # (ind_cap,) i32 (u32 bits)
    info_raw: "torch.Tensor"  # (nchunks * 32,) u8
    totals_raw: "torch.Tensor"  # (24,) u8
    cases: "torch.Tensor | None"
    tri_normals: "torch.Tensor | None" = None  # (ind_cap / 3, 3) f32, triangle ind_off / 3 + t
    tri_cells: "torch.Tensor | None" = None    # (ind_cap / 3,) i32 (u32 bits), chunk-local cell

    def info(self) -> np.ndarray:
        return self.info_raw.cpu().numpy().view(INFO_DTYPE)

    def totals(self) -> np.ndarray:
        return self.totals_raw.cpu().numpy().view(TOTALS_DTYPE)[0]

    def chunk_mesh(self, i: int, info=None):
        """(verts (V,3) or (V,6) f32, inds (I,) u32) of chunk i, as numpy arrays."""
        r = (info if info is not None else self.info())[i]
        v0, nv, i0, ni = int(r["vert_off"]), int(r["n_verts"]), int(r["ind_off"]), int(r["n_inds"])
        v = self.verts[v0:v0 + nv].cpu().numpy()
        ix = self.inds[i0:i0 + ni].cpu().numpy()
        ix = ix.view(np.uint16 if ix.dtype == np.int16 else np.uint32)
        return v, ix

    def chunk_collision(self, i: int, info=None):
        """(tri_cells (T,) u32, normals (T, 3) f32) of chunk i (built with collision=True)."""
        r = (info if info is not None else self.info())[i]
        t0, nt = int(r["ind_off"]) // 3, int(r["n_inds"]) // 3
        cells = self.tri_cells[t0:t0 + nt].cpu().numpy().view(np.uint32)
        return cells, self.tri_normals[t0:t0 + nt].cpu().numpy()


def _alloc_out(nchunks: int, vert_cap: int, ind_cap: int, device, colors: bool = False, u16: bool = False):
    import torch
    verts = torch.empty((max(vert_cap, 1), 6 if colors else 3), dtype=torch.float32, device=device)
    inds = torch.empty((max(ind_cap, 1),), dtype=torch.int16 if u16 else torch.int32, device=device)
    info = torch.empty((max(nchunks, 1) * 32,), dtype=torch.uint8, device=device)
    totals = torch.empty((24,), dtype=torch.uint8, device=device)
    return verts, inds, info, totals


def _alloc_collision(ind_cap: int, device):
    import torch
    nt = max(ind_cap // 3, 1)
    return (torch.empty((nt, 3), dtype=torch.float32, device=device),
            torch.empty((nt,), dtype=torch.int32, device=device))


def _mesh_out(verts, inds, info, totals, vert_cap, ind_cap, collision=None):
    import torch
    fmt = VERTEX_XYZ_RGB if verts.shape[1] == 6 else VERTEX_XYZ
    ifmt = INDEX_U16 if inds.dtype == torch.int16 else INDEX_U32
    nrm, cells = collision if collision is not None else (None, None)
    return _lib.UwMeshOut(verts.data_ptr(), inds.data_ptr(), info.data_ptr(), totals.data_ptr(), vert_cap, ind_cap,
                          fmt, ifmt, nrm.data_ptr() if nrm is not None else None,
                          cells.data_ptr() if cells is not None else None)


def _coords_tensor(coords, device):
    import torch
    if isinstance(coords, torch.Tensor):
        t = coords.to(device=device, dtype=torch.int32).contiguous().view(-1, 3)
    else:
        t = torch.as_tensor(np.asarray(coords, np.int32).reshape(-1, 3), device=device)
    return t


def build_chunks_plain(perlin: Perlin, coords, n: int, vert_cap: int | None = None, ind_cap: int | None = None,
                       want_cases: bool = False, stream=None, colors: bool = False, u16: bool = False,
                       collision: bool = False) -> MeshBatch:
    import torch
    dev = torch.device("cuda", perlin.device)
    ct = _coords_tensor(coords, dev)
    nch = ct.shape[0]
    L = _lib.lib()
    if vert_cap is None:
        vert_cap = int(L.uw_max_verts(n)) * nch
    if ind_cap is None:
        ind_cap = int(L.uw_max_inds(n)) * nch
    verts, inds, info, totals = _alloc_out(nch, vert_cap, ind_cap, dev, colors, u16)
    coll = _alloc_collision(ind_cap, dev) if collision else None
    cases = torch.empty((max(nch, 1) * n ** 3,), dtype=torch.uint8, device=dev) if want_cases else None
    out = _mesh_out(verts, inds, info, totals, vert_cap, ind_cap, coll)
    rc = L.uw_build_chunks_plain(perlin.ctx, ctypes.c_void_p(ct.data_ptr()), nch, n, ctypes.byref(out),
                                 ctypes.c_void_p(cases.data_ptr() if cases is not None else 0), _stream_ptr(stream))
    _lib.check(rc, perlin.ctx)
    return MeshBatch(verts, inds, info[: nch * 32], totals, cases, *(coll or (None, None)))


def _res_tensor(res, nch, device):
    import torch
    if isinstance(res, int):
        return torch.full((max(nch, 1),), int(res), dtype=torch.int32, device=device)
    if isinstance(res, torch.Tensor):
        return res.to(device=device, dtype=torch.int32).contiguous()
    return torch.as_tensor(np.asarray(res, np.int32), device=device)


def batch_capacity(res) -> tuple[int, int]:
    """Worst-case vertex and index counts of a chunk set (resolution per chunk)."""
    r = np.asarray(res, np.int64)
    return int((5 * r * (r + 1) ** 2).sum()), int((15 * r ** 3).sum())


def build_chunks_batched(perlin: Perlin, coords, res, vert_cap: int | None = None, ind_cap: int | None = None,
                         stream=None, colors: bool = False, u16: bool = False, collision: bool = False) -> MeshBatch:
    """One persistent launch over a mixed-resolution chunk set (res: cells per axis per chunk).
    colors=True writes the reference's VertColor layout (position + sRGB colour per vertex)."""
    import torch
    dev = torch.device("cuda", perlin.device)
    ct = _coords_tensor(coords, dev)
    nch = ct.shape[0]
    rt = _res_tensor(res, nch, dev)
    if vert_cap is None or ind_cap is None:
        vc, ic = batch_capacity(rt[:nch].cpu().numpy() if nch else [])
        vert_cap = vc if vert_cap is None else vert_cap
        ind_cap = ic if ind_cap is None else ind_cap
    verts, inds, info, totals = _alloc_out(nch, vert_cap, ind_cap, dev, colors, u16)
    coll = _alloc_collision(ind_cap, dev) if collision else None
    out = _mesh_out(verts, inds, info, totals, vert_cap, ind_cap, coll)
    rc = _lib.lib().uw_build_chunks_batched(perlin.ctx, ctypes.c_void_p(ct.data_ptr()), ctypes.c_void_p(rt.data_ptr()),
                                            nch, ctypes.byref(out), _stream_ptr(stream))
    _lib.check(rc, perlin.ctx)
    return MeshBatch(verts, inds, info[: nch * 32], totals, None, *(coll or (None, None)))


class BatchedGraph:
    """The batched build captured once over fixed buffers and replayed (uw_batched_graph_create)."""

    def __init__(self, perlin: Perlin, coords, res, vert_cap: int | None = None, ind_cap: int | None = None,
                 colors: bool = False):
        import torch
        dev = torch.device("cuda", perlin.device)
        self.perlin = perlin
        self.coords = _coords_tensor(coords, dev)
        nch = self.coords.shape[0]
        self.res = _res_tensor(res, nch, dev)
        if vert_cap is None or ind_cap is None:
            vc, ic = batch_capacity(self.res[:nch].cpu().numpy())
            vert_cap = vc if vert_cap is None else vert_cap
            ind_cap = ic if ind_cap is None else ind_cap
        verts, inds, info, totals = _alloc_out(nch, vert_cap, ind_cap, dev, colors)
        self.batch = MeshBatch(verts, inds, info[: nch * 32], totals, None)
        self._out = _mesh_out(verts, inds, info, totals, vert_cap, ind_cap)
        self._g = ctypes.c_void_p()
        torch.cuda.synchronize()
        rc = _lib.lib().uw_batched_graph_create(perlin.ctx, ctypes.c_void_p(self.coords.data_ptr()),
                                                ctypes.c_void_p(self.res.data_ptr()), nch, ctypes.byref(self._out),
                                                ctypes.byref(self._g))
        _lib.check(rc, perlin.ctx)

    def launch(self, stream=None) -> MeshBatch:
        rc = _lib.lib().uw_graph_launch(self._g, _stream_ptr(stream))
        if rc != _lib.UW_OK:
            raise _lib.UwError(rc, "uw_graph_launch failed")
        return self.batch

    def close(self):
        if self._g:
            _lib.lib().uw_graph_destroy(self._g)
            self._g = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_volume(perlin: Perlin, n: int, chunk=(0, 0, 0), vert_cap: int | None = None, ind_cap: int | None = None,
                 stream=None, colors: bool = False, u16: bool = False, collision: bool = False,
                 tape: bool = False):
    """One chunk of n cells per axis (up to 640) meshed as one grid with one index space
    (uw_build_volume): Chunk::new(chunk) + build_full at resolution n. Capacities default to the
    exact sizes, found by a counting build first (one host synchronisation). Returns the MeshBatch
    (one info record) and, with tape=True, the (vert_cap, 4) int32 vertex tape for the backward."""
    import torch
    dev = torch.device("cuda", perlin.device)
    if vert_cap is None or ind_cap is None:
        v, i = volume_required_sizes(perlin, n, chunk)
        vert_cap = v if vert_cap is None else vert_cap
        ind_cap = i if ind_cap is None else ind_cap
    verts, inds, info, totals = _alloc_out(1, vert_cap, ind_cap, dev, colors, u16)
    coll = _alloc_collision(ind_cap, dev) if collision else None
    tp = torch.empty((max(vert_cap, 1), 4), dtype=torch.int32, device=dev) if tape else None
    out = _mesh_out(verts, inds, info, totals, vert_cap, ind_cap, coll)
    rc = _lib.lib().uw_build_volume(perlin.ctx, int(chunk[0]), int(chunk[1]), int(chunk[2]), int(n), ctypes.byref(out),
                                    ctypes.c_void_p(tp.data_ptr()) if tp is not None else None, _stream_ptr(stream))
    _lib.check(rc, perlin.ctx)
    b = MeshBatch(verts, inds, info[:32], totals, None, *(coll or (None, None)))
    return (b, tp) if tape else b


def volume_required_sizes(perlin: Perlin, n: int, chunk=(0, 0, 0)) -> tuple[int, int]:
    """(vertices, indices) of the volume: a build with zero capacity reports the totals without
    writing a mesh."""
    import torch
    dev = torch.device("cuda", perlin.device)
    verts, inds, info, totals = _alloc_out(1, 0, 0, dev)
    out = _mesh_out(verts, inds, info, totals, 0, 0)
    rc = _lib.lib().uw_build_volume(perlin.ctx, int(chunk[0]), int(chunk[1]), int(chunk[2]), int(n), ctypes.byref(out),
                                    None, _stream_ptr(None))
    _lib.check(rc, perlin.ctx)
    t = MeshBatch(verts, inds, info, totals, None).totals()
    return int(t["n_verts"]), int(t["n_inds"])


def get_batch_tuning(perlin: Perlin) -> dict:
    t = _lib.UwBatchTuning()
    _lib.check(_lib.lib().uw_get_batch_tuning(perlin.ctx, ctypes.byref(t)), perlin.ctx)
    return {"block": t.block, "grid": t.grid, "max_res": t.max_res, "slab": [int(v) for v in t.slab],
            "bulk_min": t.bulk_min}


def set_batch_tuning(perlin: Perlin, tuning: dict):
    t = _lib.UwBatchTuning()
    t.block = int(tuning["block"])
    t.grid = int(tuning.get("grid", 0))
    t.max_res = int(tuning.get("max_res", _lib.BATCH_MAX_RESOLUTION))
    for i, v in enumerate(tuning["slab"]):
        t.slab[i] = int(v)
    t.bulk_min = int(tuning.get("bulk_min", 0))
    _lib.check(_lib.lib().uw_set_batch_tuning(perlin.ctx, ctypes.byref(t)), perlin.ctx)


def density_at(perlin: Perlin, points, stream=None):
    """iso_at at arbitrary noise-space points (N, 3) float64 -> (N,) float32 tensor."""
    import torch
    dev = torch.device("cuda", perlin.device)
    pts = torch.as_tensor(np.ascontiguousarray(points, np.float64).reshape(-1, 3), device=dev)
    out = torch.empty((max(pts.shape[0], 1),), dtype=torch.float32, device=dev)
    rc = _lib.lib().uw_density_at(perlin.ctx, ctypes.c_void_p(pts.data_ptr()), pts.shape[0],
                                  ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream))
    _lib.check(rc, perlin.ctx)
    return out[: pts.shape[0]]


def sample_density(perlin: Perlin, coords, n: int, stream=None):
    """Density grids (nchunks, (n+1)^3) f32 and the not-early-blank flags (nchunks,) as tensors."""
    import torch
    dev = torch.device("cuda", perlin.device)
    ct = _coords_tensor(coords, dev)
    nch = ct.shape[0]
    dens = torch.empty((max(nch, 1), (n + 1) ** 3), dtype=torch.float32, device=dev)
    below = torch.empty((max(nch, 1),), dtype=torch.int32, device=dev)
    rc = _lib.lib().uw_sample_density(perlin.ctx, ctypes.c_void_p(ct.data_ptr()), nch, n,
                                      ctypes.c_void_p(dens.data_ptr()), ctypes.c_void_p(below.data_ptr()),
                                      _stream_ptr(stream))
    _lib.check(rc, perlin.ctx)
    return dens[:nch], below[:nch]

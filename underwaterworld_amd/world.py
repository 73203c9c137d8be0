"""World streaming with a batched generation step (world.rs:64-247) over the C ABI's uw_world_*.

World mirrors the reference's World: update(sub_pos, bearing, view_proj, gen_proj, sub_reset),
get_chunk, chunks_to_render, generate_count, render_count, total_count. Matrices are 4x4 arrays
indexed [row, col] (what cgmath's Matrix4 means mathematically); they are handed to the engine
column-major. Each update builds up to max_batch pending chunks in one launch (0 = all of them)
where the reference builds one chunk per frame.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .engine import INDEX_U16, INDEX_U32, VERTEX_XYZ, VERTEX_XYZ_RGB, Perlin, tri_ids_around


UwWorldConfig = _lib.UwWorldConfig
UwWorldCandidate = _lib.UwWorldCandidate
UwWorldChunk = _lib.UwWorldChunk


def _L():
    return _lib.lib()


def reference_config(**overrides) -> UwWorldConfig:
    """world.rs:5-12 constants (plus engine options), with overrides by field name."""
    c = UwWorldConfig()
    _L().uw_world_config_reference(ctypes.byref(c))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def _mat(m) -> np.ndarray:
    """4x4 [row, col] -> 16 f32 column-major."""
    return np.ascontiguousarray(np.asarray(m, np.float32).reshape(4, 4).T).ravel()


def _vec(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, np.float32).reshape(3))


def plan(cfg: UwWorldConfig, sub_pos, view_proj, gen_proj, known: dict | None = None):
    """World::update_nearby (uw_world_plan): (render [(x,y,z)], candidates [(pos, dist, z, in_view,
    in_gen)] in pop order). known maps chunk position -> not_blank."""
    known = known or {}
    kp = np.array(list(known.keys()), np.int32).reshape(-1, 3)
    kb = np.array([1 if known[tuple(k)] else 0 for k in kp.tolist()], np.uint8)
    sp, vp, gp = _vec(sub_pos), _mat(view_proj), _mat(gen_proj)
    cap = (2 * cfg.generation_dist) ** 2 * (cfg.max_z - cfg.min_z + 1)
    render = np.zeros((cap, 3), np.int32)
    cands = (UwWorldCandidate * cap)()
    nr, nc = ctypes.c_int32(0), ctypes.c_int32(0)
    rc = _L().uw_world_plan(ctypes.byref(cfg), sp.ctypes.data, vp.ctypes.data, gp.ctypes.data,
                            kp.ctypes.data if len(kp) else None, kb.ctypes.data if len(kb) else None, len(kp),
                            render.ctypes.data, cap, ctypes.byref(nr), cands, cap, ctypes.byref(nc))
    _lib.check(rc)
    out = [(tuple(c.pos), np.float32(c.dist), np.float32(c.z), bool(c.in_view), bool(c.in_gen))
           for c in cands[:nc.value]]
    return [tuple(r) for r in render[:nr.value].tolist()], out


class WorldChunk:
    """A stored chunk read back to host memory (the upload source of chunk.rs:291-303)."""

    def __init__(self, info, verts, inds, tri_cells, normals, resolution):
        self.info = info
        self.verts = verts
        self.inds = inds
        self.tri_cells = tri_cells
        self.normals = normals
        self.resolution = resolution

    def not_blank(self) -> bool:
        return self.info.n_inds > 0

    def num_inds(self) -> int:
        return int(self.info.n_inds)

    def tris_around(self, local_pos_percent, range_: int):
        """(verts (T, 3, 3), normals (T, 3)) of Chunk::tris_around (chunk.rs:315-342)."""
        if self.tri_cells is None or not self.not_blank():
            return np.zeros((0, 3, 3), np.float32), np.zeros((0, 3), np.float32)
        ids = tri_ids_around(self.tri_cells, self.resolution, local_pos_percent, range_)
        tri = self.verts[:, :3][self.inds.astype(np.int64)].reshape(-1, 3, 3)
        return tri[ids], self.normals[ids]


class World:
    def __init__(self, perlin: Perlin, cfg: UwWorldConfig | None = None, **overrides):
        self.perlin = perlin
        self.cfg = cfg if cfg is not None else reference_config(**overrides)
        self._w = ctypes.c_void_p()
        _lib.check(_L().uw_world_create(perlin.ctx, ctypes.byref(self.cfg), ctypes.byref(self._w)), perlin.ctx)

    def update(self, sub_pos, bearing, view_proj, gen_proj, sub_reset: bool = False) -> int:
        """World::update (world.rs:93-111); returns the number of chunks built."""
        built = ctypes.c_int32(0)
        sp, sb, vp, gp = _vec(sub_pos), _vec(bearing), _mat(view_proj), _mat(gen_proj)
        rc = _L().uw_world_update(self._w, sp.ctypes.data, sb.ctypes.data, vp.ctypes.data, gp.ctypes.data,
                                  int(bool(sub_reset)), ctypes.byref(built))
        _lib.check(rc, self.perlin.ctx)
        return built.value

    def chunks_to_render(self) -> list:
        n = _L().uw_world_chunks_to_render(self._w, None, 0)
        out = np.zeros((max(n, 1), 3), np.int32)
        _L().uw_world_chunks_to_render(self._w, out.ctypes.data, n)
        return [tuple(r) for r in out[:n].tolist()]

    def generate_count(self) -> int:
        return int(_L().uw_world_generate_count(self._w))

    def render_count(self) -> int:
        return int(_L().uw_world_render_count(self._w))

    def total_count(self) -> int:
        return int(_L().uw_world_total_count(self._w))

    def has_chunk(self, pos) -> bool:
        c = UwWorldChunk()
        return _L().uw_world_get_chunk(self._w, *[int(v) for v in pos], ctypes.byref(c)) == _lib.UW_OK

    def get_chunk(self, pos) -> WorldChunk | None:
        """World::get_chunk (world.rs:89-91), read back to host arrays; None if not stored."""
        c = UwWorldChunk()
        if _L().uw_world_get_chunk(self._w, *[int(v) for v in pos], ctypes.byref(c)) != _lib.UW_OK:
            return None
        nv, ni = int(c.info.n_verts), int(c.info.n_inds)
        colors = self.cfg.vertex_format == VERTEX_XYZ_RGB
        u16 = self.cfg.index_format == INDEX_U16
        verts = np.zeros((max(nv, 1), 6 if colors else 3), np.float32)
        inds = np.zeros(max(ni, 1), np.uint16 if u16 else np.uint32)
        coll = bool(self.cfg.collision)
        normals = np.zeros((max(ni // 3, 1), 3), np.float32) if coll else None
        cells = np.zeros(max(ni // 3, 1), np.uint32) if coll else None
        ho = _lib.UwHostOut(verts.ctypes.data, nv, inds.ctypes.data, ni,
                            VERTEX_XYZ_RGB if colors else VERTEX_XYZ, INDEX_U16 if u16 else INDEX_U32,
                            normals.ctypes.data if coll else None, cells.ctypes.data if coll else None)
        info = _lib.UwChunkInfo()
        _lib.check(_L().uw_world_read_chunk(self._w, *[int(v) for v in pos], ctypes.byref(ho), ctypes.byref(info)),
                   self.perlin.ctx)
        return WorldChunk(info, verts[:nv], inds[:ni], cells[:ni // 3] if coll else None,
                          normals[:ni // 3] if coll else None, int(self.cfg.resolution))

    def close(self):
        if self._w:
            _L().uw_world_destroy(self._w)
            self._w = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

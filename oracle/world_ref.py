"""Pure-Python restatement of the reference's chunk streaming planner, used ONLY as a checker.

TEST INFRASTRUCTURE: imported by tests/ only. Restates underwater_world/src/world.rs:
  - World::update's recheck rule (world.rs:93-111): re-plan when reset, moved > 4 or turned
    > 0.33 rad (cgmath Vector3::angle = atan2(|a x b|, a . b), recalled);
  - World::update_nearby (world.rs:149-228): candidates over x, y in [-5, 5) around the sub's
    chunk and z in [max(sz - 5, -2), min(sz + 5, 2)], skipped beyond 5 chunk extents, flagged
    in_view / in_gen by util::in_frustum (util.rs:74-83) on the 8 chunk corners (VIEW 55 deg,
    GENERATE 90 deg frusta, world.rs:15-16); existing chunks go to the render list when
    dist < 4 extents, not blank and in view; missing ones become generation candidates;
  - GenPrio::compare (world.rs:35-60): in_view first, then in_gen, then dist^2 + z;
  - build_full_step / build_step render rule (world.rs:113-145);
  - remove_far_way (world.rs:230-241): one key per update, evicted at dist_sq >= 36.
f32 arithmetic is emulated with numpy.float32 scalars in the reference's operation order.
The candidate order among exact priority ties is unspecified in the reference
(sort_unstable_by); this restatement returns ties in reverse generation order.
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32

VIEW_DIST, GENERATION_DIST, KEEP_DIST, MAX_Z, MIN_Z = 4, 5, 6, 2, -2
CHUNK_SIZE = 16
RECHECK_NEARBY_DIST, RECHECK_NEARBY_ANGLE = f32(4.0), f32(0.33)
STOP_FULL_BUILD = GENERATION_DIST ** 3

_CORNERS = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (1.0, 1.0, 0.0),
            (0.0, 0.0, 1.0), (1.0, 0.0, 1.0), (0.0, 1.0, 1.0), (1.0, 1.0, 1.0)]


def magnitude(v) -> np.float32:
    x, y, z = (f32(c) for c in v)
    return f32(np.sqrt(f32(f32(x * x + y * y) + z * z)))


def in_frustum(pt, m) -> bool:
    """util::in_frustum: m is a column-major 4x4 (m[col][row]) cgmath Matrix4."""
    v = (f32(pt[0]), f32(pt[1]), f32(pt[2]), f32(1.0))
    r = []
    for row in range(4):
        acc = f32(m[0][row] * v[0])
        for col in range(1, 4):
            acc = f32(acc + f32(m[col][row] * v[col]))
        r.append(acc)
    w = r[3]
    x, y, z = f32(r[0] / w), f32(r[1] / w), f32(r[2] / w)
    return bool(abs(x) <= 1.0 and abs(y) <= 1.0 and z >= 0.0 and z <= 1.0)


def sub_chunk(pos):
    return tuple(int(math.floor(f32(f32(p) / f32(CHUNK_SIZE)))) for p in pos)


def dist_sq(a, b) -> int:
    return sum((x - y) * (x - y) for x, y in zip(a, b))


def angle(a, b) -> float:
    a = [f32(v) for v in a]
    b = [f32(v) for v in b]
    cx = f32(a[1] * b[2]) - f32(a[2] * b[1])
    cy = f32(a[2] * b[0]) - f32(a[0] * b[2])
    cz = f32(a[0] * b[1]) - f32(a[1] * b[0])
    dot = f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2]))
    return float(np.arctan2(magnitude((cx, cy, cz)), dot).astype(np.float32))


def update_nearby(sub_pos, view_proj, gen_proj, chunks: dict):
    """Returns (chunks_to_render, candidates in pop order [(pos, (dist, z, in_view, in_gen))])."""
    sp = tuple(f32(v) for v in sub_pos)
    sc = sub_chunk(sp)
    max_view = f32(VIEW_DIST * CHUNK_SIZE)
    max_gen = f32(GENERATION_DIST * CHUNK_SIZE)
    start_z = max(sc[2] - GENERATION_DIST, MIN_Z)
    end_z = min(sc[2] + GENERATION_DIST, MAX_Z)
    render, cands = [], []
    for x in range(-GENERATION_DIST, GENERATION_DIST):
        cx = sc[0] + x
        for y in range(-GENERATION_DIST, GENERATION_DIST):
            cy = sc[1] + y
            for cz in range(start_z, end_z + 1):
                center = (f32(f32(f32(cx) + f32(0.5)) * f32(CHUNK_SIZE)),
                          f32(f32(f32(cy) + f32(0.5)) * f32(CHUNK_SIZE)),
                          f32(f32(f32(cz) + f32(0.5)) * f32(CHUNK_SIZE)))
                d = magnitude([f32(sp[k] - center[k]) for k in range(3)])
                if d > max_gen:
                    continue
                in_view = in_gen = False
                for c in _CORNERS:
                    corner = (f32(f32(f32(cx) + f32(c[0])) * f32(CHUNK_SIZE)),
                              f32(f32(f32(cy) + f32(c[1])) * f32(CHUNK_SIZE)),
                              f32(f32(f32(cz) + f32(c[2])) * f32(CHUNK_SIZE)))
                    if in_frustum(corner, gen_proj):
                        in_gen = True
                        if in_frustum(corner, view_proj):
                            in_view = True
                            break
                pos = (cx, cy, cz)
                if pos in chunks:
                    if d < max_view and chunks[pos] and in_view:
                        render.append(pos)
                else:
                    cands.append((pos, (d, f32(f32(cz) * f32(CHUNK_SIZE)), in_view, in_gen)))

    def key(item):
        d, z, iv, ig = item[1]
        return (0 if iv else 1, 0 if ig else 1, float(f32(f32(d * d) + z)))

    # pop order: most important first; ties in reverse generation order
    order = sorted(range(len(cands)), key=lambda i: (key(cands[i]), -i))
    return render, [cands[i] for i in order]


class WorldRef:
    """World::update with the batched generation step: each update builds up to max_batch
    candidates in pop order (the reference builds one per frame)."""

    def __init__(self, max_batch: int = 0):
        self.chunks: dict = {}  # pos -> not_blank, insertion ordered
        self.render: list = []
        self.to_generate: list = []
        self.keys_left: list = []
        self.should_full_build = True
        self.last_pos = (f32(np.inf),) * 3
        self.last_bearing = (f32(np.inf),) * 3
        self.max_batch = max_batch

    def remove_far_way(self, sub_pos):
        if self.keys_left:
            pos = self.keys_left.pop()
            if dist_sq(pos, sub_chunk(sub_pos)) >= KEEP_DIST * KEEP_DIST:
                self.chunks.pop(pos, None)
        else:
            self.keys_left = list(self.chunks.keys())

    def update(self, sub_pos, bearing, view_proj, gen_proj, sub_reset, not_blank_of):
        """not_blank_of(list of pos) -> list of bool: the build of a batch."""
        self.remove_far_way(sub_pos)
        d = magnitude([f32(f32(sub_pos[k]) - self.last_pos[k]) for k in range(3)])
        ang = angle(bearing, self.last_bearing)
        if sub_reset or d > RECHECK_NEARBY_DIST or ang > RECHECK_NEARBY_ANGLE:
            self.render, self.to_generate = update_nearby(sub_pos, view_proj, gen_proj, self.chunks)
            self.last_pos = tuple(f32(v) for v in sub_pos)
            self.last_bearing = tuple(f32(v) for v in bearing)
        k = len(self.to_generate) if self.max_batch <= 0 else min(self.max_batch, len(self.to_generate))
        batch = [p for p, _ in self.to_generate[:k]]
        self.to_generate = self.to_generate[k:]
        nb = not_blank_of(batch) if batch else []
        sc = sub_chunk(sub_pos)
        full = self.should_full_build
        for pos, b in zip(batch, nb):
            if b and (full or dist_sq(pos, sc) <= (VIEW_DIST + 1) * (VIEW_DIST + 1)):
                self.render.append(pos)
            self.chunks[pos] = bool(b)
        if full:
            self.should_full_build = not (not self.to_generate or len(self.chunks) >= STOP_FULL_BUILD)
        return batch

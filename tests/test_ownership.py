"""The engine numbers vertices without a dedup search: a directed edge record belongs to the first
cell (in scan order) that contains it, and its id is a prefix sum of owned-record counts plus a
rank inside the owner (underwaterworld_amd/csrc/uw_mesh.h). This restates that rule in Python and
checks it against the oracle's literal first-seen ordered-pair dedup (chunk.rs:233-243)."""
import numpy as np
import pytest

import npref
from oracle import oracle as O
from test_tables import packed_rows
from underwaterworld_amd import params as PR


def owned_mask(x, y, z):
    m = 0x4F0
    if y == 0:
        m |= 0x00F
    if x == 0 and z == 0:
        m |= 1 << 8
    if z == 0:
        m |= 1 << 9
    if x == 0:
        m |= 1 << 11
    return m


def owner(x, y, z, e):
    if e < 4:
        return (x, y - 1, z), e + 4
    if e == 8:
        if x > 0 and z > 0:
            return (x - 1, y, z - 1), 10
        if x > 0:
            return (x - 1, y, z), 9
        return (x, y, z - 1), 11
    if e == 9:
        return (x, y, z - 1), 10
    return (x - 1, y, z), 10


def before_table(rows):
    t = np.zeros((256, 12), np.int64)
    for c, r in enumerate(rows):
        seen = 0
        b = [None] * 12
        for e in r:
            if not (seen >> e) & 1:
                b[e] = seen
                seen |= 1 << e
        for e in range(12):
            t[c, e] = seen if b[e] is None else b[e]
    return t


def ownership_ids(cases, n, rows):
    cases = cases.reshape(n, n, n)
    before = before_table(rows)
    used = [sum(1 << e for e in set(r)) for r in rows]
    counts = np.zeros((n, n, n), np.int64)
    for x in range(n):
        for y in range(n):
            for z in range(n):
                counts[x, y, z] = bin(used[cases[x, y, z]] & owned_mask(x, y, z)).count("1")
    voff = (np.cumsum(counts.ravel()) - counts.ravel()).reshape(n, n, n)
    inds = []
    for x in range(n):
        for y in range(n):
            for z in range(n):
                c = cases[x, y, z]
                om = owned_mask(x, y, z)
                for e in rows[c]:
                    if (om >> e) & 1:
                        inds.append(voff[x, y, z] + bin(before[c, e] & om).count("1"))
                    else:
                        (qx, qy, qz), qe = owner(x, y, z, e)
                        qc = cases[qx, qy, qz]
                        assert (used[qc] >> qe) & 1, "owner does not emit the record"
                        inds.append(voff[qx, qy, qz] + bin(before[qc, qe] & owned_mask(qx, qy, qz)).count("1"))
    return np.array(inds, np.uint32), int(counts.sum())


@pytest.mark.parametrize("n,seed,frac", [(1, 0, 0.5), (2, 1, 0.5), (3, 2, 0.3), (4, 3, 0.5), (5, 4, 0.7),
                                         (6, 5, 0.5)])
def test_ownership_numbering_equals_first_seen_random(n, seed, frac):
    rows = packed_rows()
    p = PR.preset("chunks12")
    d = (np.random.default_rng(seed).random((n + 1) ** 3) - frac).astype(np.float32)
    m = O.mesh_ref(p, (0, 0, 0), n, d)
    ids, nv = ownership_ids(m["cases"], n, rows)
    assert nv == len(m["verts"])
    assert np.array_equal(ids, m["inds"])


@pytest.mark.parametrize("preset,coord", [("reference", (0, 0, -1)), ("chunks12", (0, 0, -1)),
                                          ("chunks12", (3, 1, -1))])
def test_ownership_numbering_equals_first_seen_chunks(preset, coord):
    rows = packed_rows()
    p = PR.preset(preset)
    m = O.build_chunk(p, coord, 12)
    ids, nv = ownership_ids(m["cases"], 12, rows)
    assert nv == len(m["verts"])
    assert np.array_equal(ids, m["inds"])


def test_every_corner_cell_pattern_of_a_2cube():
    # exhaustive over sign patterns of a 2x2x2-cell block's boundary-heavy layouts (random subset)
    rows = packed_rows()
    p = PR.preset("chunks12")
    g = np.random.default_rng(9)
    for _ in range(300):
        d = np.where(g.random(27) < 0.5, -1.0, 1.0).astype(np.float32)
        m = O.mesh_ref(p, (0, 0, 0), 2, d)
        ids, nv = ownership_ids(m["cases"], 2, rows)
        assert nv == len(m["verts"]) and np.array_equal(ids, m["inds"])

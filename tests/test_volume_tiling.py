"""The single-volume path (csrc/uw_volume.hip) numbers the vertices of one large grid tile by tile:
per row segment (x, y, z range of one tile) it counts owned vertex records and triangles, scans
those counts in scan order, and each tile then numbers its cells from its rows' bases plus a
within-row prefix, with a one-cell halo on its low sides giving the ids of records owned by cells
of neighbouring tiles (the halo cell of a row's z range is the last cell of the previous segment,
so its base is the segment base minus its own count).

This restates that procedure in Python -- with the kernels' LDS extents asserted -- and checks it
against the oracle's literal first-seen ordered-pair dedup over the whole grid (chunk.rs:233-243)
on random grids whose sizes are not multiples of the tile, with tiles small enough that most
records lie on or next to a tile face."""
import numpy as np
import pytest

from oracle import oracle as O
from test_ownership import before_table, owned_mask, owner
from test_tables import packed_rows
from underwaterworld_amd import params as PR

CORNER_OFS = [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)]


def _cases(d, n, iso):
    D = d.reshape(n + 1, n + 1, n + 1) < np.float32(iso)
    c = np.zeros((n, n, n), np.int64)
    for i, (a, b, e) in enumerate(CORNER_OFS):
        c |= (D[a:a + n, b:b + n, e:e + n].astype(np.int64) << i)
    return c


def tiled_ids(cases, n, rows, tile, planes=None):
    """Index buffer and vertex count of the grid by the volume kernels' procedure (emission by
    x-slabs of `planes` cell planes of a tile, each with its own halo plane)."""
    planes = planes or tile
    before = before_table(rows)
    used = [sum(1 << e for e in set(r)) for r in rows]
    ntri = [len(r) // 3 for r in rows]
    T = (n + tile - 1) // tile
    nv_of = lambda x, y, z: bin(used[cases[x, y, z]] & owned_mask(x, y, z)).count("1")  # noqa: E731
    # k_vol_count: per row segment (x, y, tz): owned records | triangles
    cnt = np.zeros((n, n, T, 2), np.int64)
    for x in range(n):
        for y in range(n):
            for z in range(n):
                cnt[x, y, z // tile, 0] += nv_of(x, y, z)
                cnt[x, y, z // tile, 1] += ntri[cases[x, y, z]]
    # k_vol_scan_*: exclusive scan in (x, y, tz) order
    flat = cnt.reshape(-1, 2)
    base = (np.cumsum(flat, axis=0) - flat).reshape(n, n, T, 2)
    nverts, ntris = int(flat[:, 0].sum()), int(flat[:, 1].sum())
    inds = np.full(3 * ntris, -1, np.int64)
    # k_vol_emit, one x-slab of a tile at a time
    for tx in range(T):
        for ty in range(T):
            for tz, sl in [(tz, sl) for tz in range(T) for sl in range((tile + planes - 1) // planes)]:
                X0, Y0, Z0 = tx * tile + sl * planes, ty * tile, tz * tile
                X1 = min(n, tx * tile + tile, X0 + planes)
                Y1, Z1 = min(n, Y0 + tile), min(n, Z0 + tile)
                if X0 >= X1:
                    continue
                xs, ys, zs = max(X0 - 1, 0), max(Y0 - 1, 0), max(Z0 - 1, 0)
                ny, nz = Y1 - ys, Z1 - zs
                assert ny <= tile + 1 and nz <= tile + 1 and (ny + 1) <= tile + 2 and (nz + 1) <= tile + 2
                CB = {}  # (x slot, r * nz + zz) -> (vertex base, case)
                for x in range(xs, X1):
                    slot = x & 1
                    for key in [k for k in CB if k[0] == slot]:
                        del CB[key]  # the slot's previous plane (x - 2) is overwritten
                    for r in range(ny):
                        y = ys + r
                        # interior z cells of the row: segment (x, y, tz), within-row prefix
                        pre_v, pre_t = 0, 0
                        for z in range(Z0, Z1):
                            o = r * nz + (z - zs)
                            assert o < (tile + 1) ** 2
                            CB[(slot, o)] = (base[x, y, tz, 0] + pre_v, cases[x, y, z],
                                             3 * (base[x, y, tz, 1] + pre_t))
                            pre_v += nv_of(x, y, z)
                            pre_t += ntri[cases[x, y, z]]
                        if zs < Z0:  # halo cell: the last cell of segment (x, y, tz - 1)
                            CB[(slot, r * nz)] = (base[x, y, tz, 0] - nv_of(x, y, zs), cases[x, y, zs], None)
                    if x < X0:
                        continue
                    for y in range(Y0, Y1):
                        for z in range(Z0, Z1):
                            vb, cs, ib = CB[(slot, (y - ys) * nz + (z - zs))]
                            om = owned_mask(x, y, z)
                            for j, e in enumerate(rows[cs]):
                                if (om >> e) & 1:
                                    vid = vb + bin(before[cs, e] & om).count("1")
                                else:
                                    (qx, qy, qz), qe = owner(x, y, z, e)
                                    assert qx >= xs and qy >= ys and qz >= zs
                                    qvb, qc, _ = CB[(qx & 1, (qy - ys) * nz + (qz - zs))]
                                    vid = qvb + bin(before[qc, qe] & owned_mask(qx, qy, qz)).count("1")
                                assert inds[ib + j] == -1
                                inds[ib + j] = vid
    assert np.all(inds >= 0)
    return inds.astype(np.uint32), nverts


@pytest.mark.parametrize("n,tile,planes,seed,frac", [(5, 2, 1, 0, 0.5), (9, 4, 2, 1, 0.5), (10, 3, 3, 2, 0.35),
                                                     (7, 7, 7, 3, 0.6), (8, 4, 1, 4, 0.5), (11, 4, 4, 5, 0.5),
                                                     (1, 4, 1, 6, 0.5), (13, 8, 4, 7, 0.5)])
def test_tiled_numbering_equals_first_seen(n, tile, planes, seed, frac):
    rows = packed_rows()
    p = PR.preset("chunks12")
    d = (np.random.default_rng(seed).random((n + 1) ** 3) - frac).astype(np.float32)
    m = O.mesh_ref(p, (0, 0, 0), n, d)
    ids, nv = tiled_ids(_cases(d, n, 0.0), n, rows, tile, planes)
    assert nv == len(m["verts"])
    assert np.array_equal(ids, m["inds"])


def test_tiled_numbering_on_a_field():
    # a smooth field (the volume preset's) rather than noise: long connected surfaces across faces
    rows = packed_rows()
    p = PR.preset("volume512", chunk_extent=14, resolution=14)
    d = O.density_ref(p, (0, 0, 0), 14)
    m = O.mesh_ref(p, (0, 0, 0), 14, d)
    assert len(m["inds"]) > 0
    ids, nv = tiled_ids(_cases(d, 14, p["iso"]), 14, rows, 4)
    assert nv == len(m["verts"]) and np.array_equal(ids, m["inds"])

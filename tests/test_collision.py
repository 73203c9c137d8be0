"""Collision triangles (Build.tris, chunk.rs:42,167-175,245-250; util::Tri, util.rs:7-20) and
Chunk::tris_around (chunk.rs:315-342): the oracle's restatement pinned by numpy geometry and a
brute-force box filter, and the engine's host-side lookup (uw_tris_around, no GPU needed)
against the oracle on many queries, edge cases included."""
import numpy as np
import pytest

from oracle import oracle as O
from underwaterworld_amd import params as PR
from underwaterworld_amd.engine import tri_ids_around


def _np_normal(a, b, c):
    n = np.cross(np.float64(b) - a, np.float64(c) - a)
    m = np.linalg.norm(n)
    return n if m == 0 else n / m


def test_tri_normal_matches_numpy():
    rng = np.random.default_rng(5)
    for _ in range(500):
        a, b, c = rng.normal(size=(3, 3)).astype(np.float32) * rng.uniform(0.01, 30)
        got = O.tri_normal(a, b, c)
        want = _np_normal(a, b, c)
        assert np.allclose(got, want, atol=2e-5), (got, want)
        assert abs(np.linalg.norm(got) - 1.0) < 1e-5


def test_tri_normal_degenerate_and_winding():
    z = O.tri_normal([1, 2, 3], [1, 2, 3], [4, 5, 6])
    assert np.array_equal(z, np.zeros(3, np.float32))
    # counter-clockwise in the xy plane -> +z; swapping two vertices flips it
    assert np.array_equal(O.tri_normal([0, 0, 0], [1, 0, 0], [0, 1, 0]), np.float32([0, 0, 1]))
    assert np.array_equal(O.tri_normal([0, 0, 0], [0, 1, 0], [1, 0, 0]), np.float32([0, 0, -1]))


@pytest.mark.parametrize("coord", [(3, 1, -1), (0, 0, -1), (-2, 5, 0), (1, -1, -2)])
def test_mesh_tris_consistent(coord):
    p = PR.preset("reference")
    dens = O.density_ref(p, coord, 12)
    v, ix, cells, nrm = O.mesh_ref_tris(p, coord, 12, dens)
    m = O.mesh_ref(p, coord, 12, dens, with_tri_cells=True)
    assert np.array_equal(v, m["verts"]) and np.array_equal(ix, m["inds"])
    assert np.array_equal(cells, m["tri_cells"])
    assert (np.diff(cells.astype(np.int64)) >= 0).all()
    tri = v[ix].reshape(-1, 3, 3)
    for t in range(len(tri)):
        want = _np_normal(*tri[t])
        assert np.allclose(nrm[t], want, atol=2e-5)
        # cells hold the triangle: every vertex lies in the closed cell box
        g = int(cells[t])
        x, y, z = g // 144, (g // 12) % 12, g % 12
        loc = (tri[t] - np.float32(16) * np.float32(coord)) / np.float32(16 / 12)
        assert (loc[:, 0] >= x - 1e-4).all() and (loc[:, 0] <= x + 1 + 1e-4).all()
        assert (loc[:, 1] >= y - 1e-4).all() and (loc[:, 1] <= y + 1 + 1e-4).all()
        assert (loc[:, 2] >= z - 1e-4).all() and (loc[:, 2] <= z + 1 + 1e-4).all()


def _brute_around(cells, n, pct, rng_):
    box = []
    for k in range(3):
        m = int(np.floor(np.float32(pct[k]) * np.float32(n)))
        box.append((max(m - rng_, 0), min(m + rng_, n)))
    keep = []
    for t, g in enumerate(cells.tolist()):
        x, y, z = g // (n * n), (g // n) % n, g % n
        if all(lo <= c <= hi for c, (lo, hi) in zip((x, y, z), box)):
            keep.append((g, t))
    return np.array([t for _, t in sorted(keep)], np.uint32)


def _queries(rng, count):
    q = [((0.5, 0.5, 0.5), 3), ((0.0, 0.0, 0.0), 0), ((1.0, 1.0, 1.0), 2), ((-0.2, 0.5, 1.3), 3),
         ((0.99, 0.01, 0.5), 1), ((0.5, 0.5, 0.5), 20), ((-1.0, -1.0, -1.0), 3), ((2.0, 0.5, 0.5), 1)]
    for _ in range(count):
        q.append((tuple(rng.uniform(-0.3, 1.3, 3)), int(rng.integers(0, 5))))
    return q


def test_tris_around_oracle_vs_brute_force_and_engine():
    p = PR.preset("reference")
    rng = np.random.default_rng(11)
    checked = 0
    for coord in [(3, 1, -1), (0, 0, -1), (-2, 5, 0)]:
        dens = O.density_ref(p, coord, 12)
        _, _, cells, _ = O.mesh_ref_tris(p, coord, 12, dens)
        for pct, r in _queries(rng, 60):
            ref = O.tris_around(cells, 12, pct, r)
            assert np.array_equal(ref, _brute_around(cells, 12, pct, r)), (coord, pct, r)
            got = tri_ids_around(cells, 12, pct, r)
            assert np.array_equal(got, ref), (coord, pct, r)
            checked += len(ref) > 0
    assert checked > 50


def test_tris_around_edges_and_errors():
    from underwaterworld_amd import _lib
    import ctypes
    empty = np.zeros(0, np.uint32)
    assert len(tri_ids_around(empty, 12, (0.5, 0.5, 0.5), 3)) == 0
    cells = np.array([0, 0, 5, 143, 1727, 1727], np.uint32)
    assert np.array_equal(tri_ids_around(cells, 12, (0.0, 0.0, 0.0), 0), [0, 1])
    assert np.array_equal(tri_ids_around(cells, 12, (0.99, 0.99, 0.99), 0), [4, 5])
    assert np.array_equal(tri_ids_around(cells, 12, (float("nan"), 0.0, 0.0), 1), O.tris_around(cells, 12, (float("nan"), 0.0, 0.0), 1))
    # capacity: the count is reported and UW_ERR_OVERFLOW returned
    L = _lib.lib()
    ids = np.zeros(2, np.uint32)
    cnt = ctypes.c_uint32(0)
    rc = L.uw_tris_around(cells.ctypes.data, len(cells), 12, 0.5, 0.5, 0.5, 20, ids.ctypes.data, 2, ctypes.byref(cnt))
    assert rc == _lib.UW_ERR_OVERFLOW and cnt.value == 6 and list(ids) == [0, 1]
    rc = L.uw_tris_around(cells.ctypes.data, len(cells), 0, 0.5, 0.5, 0.5, 1, ids.ctypes.data, 2, ctypes.byref(cnt))
    assert rc == _lib.UW_ERR_INVALID_ARG

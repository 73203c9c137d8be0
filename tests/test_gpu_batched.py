"""GPU parity of the batched persistent path: bit-identical to the plain path chunk by chunk (both
use the same density evaluator and mesher), checked against the oracle, across tunings, under
graph replay, and on its error paths."""
import numpy as np
import pytest

from oracle import oracle as O
from parity import DENS_TOL, edge_density_gaps, position_check, tie_corners
from test_tables import packed_rows
from underwaterworld_amd import params as PR

pytestmark = pytest.mark.gpu


def _sync():
    import torch
    torch.cuda.synchronize()


def _plain_by_resolution(perlin, coords, res):
    """Plain-path meshes of a mixed set, grouped by resolution: {chunk index: (info, verts, inds)}."""
    from underwaterworld_amd import build_chunks_plain
    out = {}
    res = np.asarray(res)
    for n in sorted(set(res.tolist())):
        idx = np.nonzero(res == n)[0]
        b = build_chunks_plain(perlin, [coords[i] for i in idx], int(n))
        _sync()
        info = b.info()
        for j, i in enumerate(idx):
            v, ix = b.chunk_mesh(j, info)
            out[int(i)] = (info[j], v, ix)
    return out


def _assert_same(batch, ref, nchunks):
    info = batch.info()
    for i in range(nchunks):
        ri, rv, rx = ref[i]
        bi = info[i]
        assert (bi["flags"], bi["n_verts"], bi["n_inds"], bi["n_active"]) == \
               (ri["flags"], ri["n_verts"], ri["n_inds"], ri["n_active"]), i
        v, ix = batch.chunk_mesh(i, info)
        assert np.array_equal(v, rv) and np.array_equal(ix, rx), i


def _mixed_subset(count=96):
    coords, res = PR.mixed_workload(8192, 7)
    return coords[:count], res[:count]


def test_batched_equals_plain_mixed():
    from underwaterworld_amd import Perlin, build_chunks_batched
    p = PR.preset("mixed")
    perlin = Perlin(p)
    coords, res = _mixed_subset(128)
    ref = _plain_by_resolution(perlin, coords, res)
    b = build_chunks_batched(perlin, coords, res)
    _sync()
    _assert_same(b, ref, len(coords))
    _assert_totals(b, ref)
    t = b.totals()
    print(f"\nmixed 128: nonblank {int(t['n_nonblank'])}, verts {int(t['n_verts'])}, inds {int(t['n_inds'])}")


def _assert_totals(batch, ref):
    t = batch.totals()
    assert t["overflow"] == 0
    assert t["n_verts"] == sum(int(ref[i][0]["n_verts"]) for i in ref)
    assert t["n_inds"] == sum(int(ref[i][0]["n_inds"]) for i in ref)
    assert t["n_nonblank"] == sum(int(ref[i][0]["n_inds"]) > 0 for i in ref)


def test_batched_unpacked_reservations():
    # 9000 chunks under a max_res-32 launch shape: the worst-case index total exceeds 2^32, so the
    # launch reserves vertex and index ranges with separate atomics instead of one packed atomic
    from underwaterworld_amd import Perlin, build_chunks_batched, get_batch_tuning
    perlin = Perlin(PR.preset("mixed"))
    assert get_batch_tuning(perlin)["max_res"] == 32
    coords = [(x, y, z) for z in (-1, 0, 1) for y in range(-25, 25) for x in range(-30, 30)]
    assert len(coords) == 9000 and 9000 * 15 * 32 ** 3 >= 2 ** 32
    ref = _plain_by_resolution(perlin, coords, [8] * len(coords))
    b = build_chunks_batched(perlin, coords, [8] * len(coords))
    _sync()
    _assert_same(b, ref, len(coords))
    _assert_totals(b, ref)


def test_batched_queue_setup_forms():
    # Queue set-up at small and large launch sizes without the bulk path, where the ordering
    # workgroups place their own chunks into an interleaved queue (20000 chunks: 79 workgroups, the
    # last with 32 chunks; 300: two, the last with 44), alternating on one context so nothing of a
    # launch's set-up state leaks into the next; each equal to the plain path chunk by chunk. The
    # placement kernel's batched histogram loads (bulk launches of more than 4096 chunks) run in
    # tests/test_gpu_fullsize.py.
    from underwaterworld_amd import Perlin, build_chunks_batched
    perlin = Perlin(PR.preset("chunks12"))
    big = [(x, y, z) for z in (-1, 0) for y in range(-50, 50) for x in range(-50, 50)]
    assert len(big) == 20000 and (len(big) + 255) // 256 > 8
    small = PR.grid_coords(-4, 4)
    for coords, n in ((small, 12), (big, 6), (small, 12), (big[:300], 6)):
        ref = _plain_by_resolution(perlin, coords, [n] * len(coords))
        b = build_chunks_batched(perlin, coords, [n] * len(coords))
        _sync()
        _assert_same(b, ref, len(coords))
        _assert_totals(b, ref)


def test_batched_equals_plain_full_grid_12():
    from underwaterworld_amd import Perlin, build_chunks_batched
    p = PR.preset("chunks12")
    perlin = Perlin(p)
    coords = PR.grid_coords(-8, 8)
    ref = _plain_by_resolution(perlin, coords, [12] * len(coords))
    b = build_chunks_batched(perlin, coords, 12)
    _sync()
    _assert_same(b, ref, len(coords))
    _assert_totals(b, ref)


@pytest.mark.parametrize("preset", ["reference", "mixed"])
def test_batched_against_oracle(preset):
    from underwaterworld_amd import Perlin, build_chunks_batched
    p = PR.preset(preset)
    coords = [(0, 0, -1), (3, 1, -1), (1, 2, 0), (-4, 3, -1), (0, 0, 0), (5, 5, -2), (-9, 4, 0), (12, -15, -2)]
    res = [12, 12, 8, 24, 32, 12, 16, 5]
    b = build_chunks_batched(Perlin(p), coords, res)
    _sync()
    info = b.info()
    nonblank, worst, ill = 0, 0.0, 0
    for k, (c, n) in enumerate(zip(coords, res)):
        dref = O.density_ref(p, c, n)
        if O.early_blank(dref, p["iso"]):
            assert info[k]["flags"] & 2 or np.any(np.abs(dref - p["iso"]) <= 1e-6)
            continue
        mr = O.mesh_ref(p, c, n, dref)
        v, ix = b.chunk_mesh(k, info)
        if len(ix) != len(mr["inds"]) or not np.array_equal(ix, mr["inds"]):
            # only a corner within the tie tolerance may change the topology
            assert np.any(np.abs(dref - p["iso"]) <= 1e-6), f"chunk {c} n={n} differs without a tie"
            continue
        assert info[k]["n_active"] == mr["n_active"]
        if len(ix) == 0:
            continue
        nonblank += 1
        h = float(np.float32(p["chunk_extent"]) / np.float32(n))
        # 1e-5 of a cell edge, or the bound the 1e-6 density tolerance implies through the lerp on
        # edges whose densities differ by less than 0.1 (tests/parity.py)
        gaps = edge_density_gaps(mr["verts"], mr["inds"], mr["cases"], dref, n, packed_rows())
        e, nill, okp = position_check(v, mr["verts"], gaps, h)
        assert okp, f"chunk {c} n={n}: max |dp|/h = {e:.3g}"
        worst, ill = max(worst, e), ill + nill
    print(f"\n[{preset}] batched vs oracle: max |dp|/h = {worst:.3g}, ill-conditioned vertices {ill}")
    assert nonblank >= 3


@pytest.mark.parametrize("block,slab_budget", [(128, 8 * 1024), (256, 20 * 1024), (512, 40 * 1024), (256, 4 * 1024)])
def test_tunings_give_identical_meshes(block, slab_budget):
    from underwaterworld_amd import Perlin, build_chunks_batched, get_batch_tuning, set_batch_tuning
    p = PR.preset("mixed")
    perlin = Perlin(p)
    coords, res = _mixed_subset(64)
    base = build_chunks_batched(perlin, coords, res)
    _sync()
    ref = {}
    info = base.info()
    for i in range(len(coords)):
        v, ix = base.chunk_mesh(i, info)
        ref[i] = (info[i], v, ix)
    t = get_batch_tuning(perlin)
    t["block"] = block
    for n in range(1, 33):
        L = n + 1
        t["slab"][n] = n if L ** 3 * 4 <= slab_budget else max(1, min(n, slab_budget // (4 * L * L) - 1))
    set_batch_tuning(perlin, t)
    b = build_chunks_batched(perlin, coords, res)
    _sync()
    _assert_same(b, ref, len(coords))


def test_graph_replay_matches_eager():
    import torch
    from underwaterworld_amd import BatchedGraph, Perlin, build_chunks_batched
    p = PR.preset("mixed")
    perlin = Perlin(p)
    coords, res = _mixed_subset(96)
    eager = build_chunks_batched(perlin, coords, res)
    _sync()
    einfo = eager.info()
    ref = {}
    for i in range(len(coords)):
        v, ix = eager.chunk_mesh(i, einfo)
        ref[i] = (einfo[i], v, ix)
    g = BatchedGraph(perlin, coords, res)
    for rep in range(3):
        # eager work and synchronisation between replays
        x = torch.randn(1 << 20, device="cuda")
        float(x.sum())
        b = g.launch()
        _sync()
        _assert_same(b, ref, len(coords))
    g.close()


def test_graph_replay_sees_new_coordinates():
    from underwaterworld_amd import BatchedGraph, Perlin, build_chunks_batched
    p = PR.preset("chunks12")
    perlin = Perlin(p)
    c1 = [(0, 0, -1), (3, 1, -1), (0, 0, 0)]
    c2 = [(5, 5, -2), (0, 0, -2), (1, 1, -1)]
    g = BatchedGraph(perlin, c1, 12)
    g.launch()
    _sync()
    import torch
    g.coords.copy_(torch.as_tensor(np.asarray(c2, np.int32), device="cuda"))
    b = g.launch()
    _sync()
    e = build_chunks_batched(perlin, c2, 12)
    _sync()
    bi, ei = b.info(), e.info()
    for i in range(3):
        assert np.array_equal(b.chunk_mesh(i, bi)[1], e.chunk_mesh(i, ei)[1])
    g.close()


def test_batched_overflow_and_invalid_resolution():
    from underwaterworld_amd import Perlin, _lib, build_chunks_batched
    perlin = Perlin(PR.preset("chunks12"))
    coords = [(0, 0, -1), (3, 1, -1), (0, 0, 5), (0, 0, -1)]
    b = build_chunks_batched(perlin, coords, [12, 12, 12, 40], vert_cap=50, ind_cap=50)
    _sync()
    t, info = b.totals(), b.info()
    assert t["overflow"] == 1
    assert info[3]["flags"] & _lib.CHUNK_INVALID
    full = build_chunks_batched(perlin, coords[:3], 12)
    _sync()
    assert full.totals()["n_inds"] == t["n_inds"]
    empty = build_chunks_batched(perlin, np.zeros((0, 3), np.int32), 12)
    _sync()
    assert empty.totals()["n_inds"] == 0


def test_max_resolution_guard():
    from underwaterworld_amd import Perlin, _lib, build_chunks_batched, get_batch_tuning, set_batch_tuning
    perlin = Perlin(PR.preset("chunks12"))
    t = get_batch_tuning(perlin)
    t["max_res"] = 12
    set_batch_tuning(perlin, t)
    coords = [(0, 0, -1), (3, 1, -1), (0, 0, -1)]
    b = build_chunks_batched(perlin, coords, [12, 24, 8])
    _sync()
    info = b.info()
    assert info[1]["flags"] & _lib.CHUNK_INVALID
    assert not info[0]["flags"] & _lib.CHUNK_INVALID and info[0]["n_inds"] > 0
    assert not info[2]["flags"] & _lib.CHUNK_INVALID

"""GPU parity of the batched build's bulk path: chunks of at least bulk_min cells per axis leave the
persistent queue and are built by the plane-parallel kernels of csrc/uw_bulk.hip (densities per
z-line over the device, per-plane counts, a per-chunk scan and range reservation, per-plane
emission). Chunk records, meshes, colours, u16 indices, collision records and the adjoint's vertex
tape must be bit-identical to the same launch with the bulk path off, for several bulk_min; the
launch totals must agree; a chunk above the launch's max_res is flagged invalid inside the bulk
range; capacity overflow is flagged and leaves every chunk that fits intact."""
import numpy as np
import pytest

from underwaterworld_amd import params as PR

pytestmark = pytest.mark.gpu


def _sync():
    import torch
    torch.cuda.synchronize()


def _set_bulk(perlin, bulk_min, max_res=None):
    from underwaterworld_amd import get_batch_tuning, set_batch_tuning
    t = get_batch_tuning(perlin)
    t["bulk_min"] = int(bulk_min)
    if max_res is not None:
        t["max_res"] = int(max_res)
    set_batch_tuning(perlin, t)


def _snapshot(b, nchunks, collision=False):
    info = b.info()
    out = []
    for i in range(nchunks):
        r = info[i]
        v, ix = b.chunk_mesh(i, info)
        rec = [(int(r["flags"]), int(r["n_verts"]), int(r["n_inds"]), int(r["n_active"])), v, ix]
        if collision and r["n_inds"] > 0:
            rec += list(b.chunk_collision(i, info))
        out.append(rec)
    return out


def _assert_same(a, b, skip=()):
    assert len(a) == len(b)
    for i, (ra, rb) in enumerate(zip(a, b)):
        if i in skip:
            continue
        assert ra[0] == rb[0], (i, ra[0], rb[0])
        for xa, xb in zip(ra[1:], rb[1:]):
            assert xa.dtype == xb.dtype and np.array_equal(xa, xb), i


def _mixed(count):
    coords, res = PR.mixed_workload(8192, 7)
    return coords[:count], res[:count]


@pytest.mark.parametrize("bulk_min", [9, 17, 24, 32])
def test_bulk_equals_persistent(bulk_min):
    from underwaterworld_amd import Perlin, build_chunks_batched
    perlin = Perlin(PR.preset("mixed"))
    coords, res = _mixed(320)
    _set_bulk(perlin, 0)
    ref = build_chunks_batched(perlin, coords, res)
    _sync()
    _set_bulk(perlin, bulk_min)
    b = build_chunks_batched(perlin, coords, res)
    _sync()
    _assert_same(_snapshot(b, len(coords)), _snapshot(ref, len(coords)))
    ta, tb = b.totals(), ref.totals()
    assert all(int(ta[k]) == int(tb[k]) for k in ("n_verts", "n_inds", "n_nonblank", "overflow")), (ta, tb)
    nbulk = sum(1 for n in res if n >= bulk_min)
    print(f"\nbulk_min {bulk_min}: {nbulk} bulk chunks of {len(coords)}, verts {int(ta['n_verts'])}")


def test_bulk_formats_and_collision():
    from underwaterworld_amd import Perlin, build_chunks_batched
    perlin = Perlin(PR.preset("mixed"))
    coords, res = _mixed(96)
    # u16 indices wrap like the reference's `ind as u16`; colours are the VertColor layout
    _set_bulk(perlin, 0)
    ref = build_chunks_batched(perlin, coords, res, colors=True, u16=True, collision=True)
    _sync()
    _set_bulk(perlin, 17)
    b = build_chunks_batched(perlin, coords, res, colors=True, u16=True, collision=True)
    _sync()
    _assert_same(_snapshot(b, len(coords), True), _snapshot(ref, len(coords), True))


def test_bulk_tape_equals_persistent():
    import torch
    from underwaterworld_amd.adjoint import ChunkMesher
    p = PR.preset("adjoint24")
    perlin = __import__("underwaterworld_amd").Perlin(p)
    coords = [(x, y, z) for z in (-2, -1, 0) for y in range(-2, 2) for x in range(-2, 2)]
    res = [24] * len(coords)
    snaps = []
    for bulk_min in (0, 17):
        _set_bulk(perlin, bulk_min)
        m = ChunkMesher(perlin, coords, res)
        rec = m.forward()
        _sync()
        info = rec.batch.info()
        tape = rec.tape.cpu().numpy()
        per = []
        for i in range(len(coords)):
            v0, nv = int(info[i]["vert_off"]), int(info[i]["n_verts"])
            per.append((int(info[i]["flags"]), nv, rec.batch.chunk_mesh(i, info)[1], tape[v0:v0 + nv].copy()))
        snaps.append(per)
        del m, rec
        torch.cuda.empty_cache()
    for i, (a, b) in enumerate(zip(*snaps)):
        assert a[0] == b[0] and a[1] == b[1], i
        assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3]), i
    assert sum(s[1] for s in snaps[0]) > 0


def test_bulk_invalid_resolution_and_overflow():
    from underwaterworld_amd import Perlin, _lib, build_chunks_batched
    perlin = Perlin(PR.preset("mixed"))
    coords, res = _mixed(64)
    res = list(res)
    res[3] = 32  # above max_res 24 below: sorted into the bulk range, flagged invalid there
    _set_bulk(perlin, 0, max_res=24)
    ref = _snapshot(build_chunks_batched(perlin, coords, res), len(coords))
    _sync()
    _set_bulk(perlin, 17, max_res=24)
    b = build_chunks_batched(perlin, coords, res)
    _sync()
    snap = _snapshot(b, len(coords))
    assert snap[3][0][0] & _lib.CHUNK_INVALID
    _assert_same(snap, ref)
    # capacity overflow: which chunks fit depends on completion order; every chunk not flagged
    # keeps its full mesh
    total = b.totals()
    small = build_chunks_batched(perlin, coords, res, vert_cap=int(total["n_verts"]) // 3,
                                 ind_cap=int(total["n_inds"]) // 3)
    _sync()
    assert int(small.totals()["overflow"]) == 1
    info = small.info()
    kept = 0
    for i in range(len(coords)):
        if info[i]["flags"] & _lib.CHUNK_OVERFLOW or info[i]["n_inds"] == 0:
            continue
        v, ix = small.chunk_mesh(i, info)
        assert np.array_equal(v, ref[i][1]) and np.array_equal(ix, ref[i][2]), i
        kept += 1
    assert kept > 0

"""GPU collision triangles (tri_cells + normals, Build.tris of chunk.rs:167-175,245-250) and u16
index output (chunk.rs:243) of the plain and batched paths, against the oracle on the GPU's own
density grids (same topology), and the Chunk facade's tris_around against the oracle's."""
import numpy as np
import pytest

from oracle import oracle as O
from underwaterworld_amd import params as PR

pytestmark = pytest.mark.gpu

# Normals: the same f32 operations in the same order on both sides; sqrt and divide are correctly
# rounded on both, so they agree to rounding of the inputs (identical) -- allow 1 ulp-scale slack.
NORMAL_TOL = 1e-6


def _coords():
    return [(3, 1, -1), (0, 0, -1), (-2, 5, 0), (1, -1, -2), (4, 4, 1), (-3, 0, -2)]


def test_plain_collision_and_u16_match_oracle():
    import torch
    from underwaterworld_amd import Perlin, build_chunks_plain, sample_density
    p = PR.preset("reference")
    perlin = Perlin(p)
    coords, n = _coords(), 12
    dens, _ = sample_density(perlin, coords, n)
    b = build_chunks_plain(perlin, coords, n, collision=True, u16=True)
    bx = build_chunks_plain(perlin, coords, n)
    torch.cuda.synchronize()
    dens = dens.cpu().numpy()
    info, infox = b.info(), bx.info()
    nonblank = 0
    for i, c in enumerate(coords):
        assert not (info[i]["flags"] & 16)
        v, ix16 = b.chunk_mesh(i, info)
        vx, ix32 = bx.chunk_mesh(i, infox)
        assert ix16.dtype == np.uint16 and np.array_equal(ix16.astype(np.uint32), ix32), c
        if info[i]["n_inds"] == 0:
            continue
        nonblank += 1
        cells, nrm = b.chunk_collision(i, info)
        rv, rix, rcells, rnrm = O.mesh_ref_tris(p, c, n, dens[i])
        assert np.array_equal(v, rv) and np.array_equal(ix32, rix), c
        assert np.array_equal(cells, rcells), c
        err = np.abs(nrm - rnrm).max()
        assert err <= NORMAL_TOL, (c, err)
    assert nonblank >= 3


def test_batched_collision_equals_plain():
    import torch
    from underwaterworld_amd import Perlin, build_chunks_batched, build_chunks_plain
    p = PR.preset("mixed")
    perlin = Perlin(p)
    coords, res = PR.mixed_workload(8192, 7)
    coords, res = coords[:64], np.asarray(res[:64])
    b = build_chunks_batched(perlin, coords, res, collision=True, u16=True)
    torch.cuda.synchronize()
    info = b.info()
    checked = 0
    for n in sorted(set(res.tolist())):
        idx = np.nonzero(res == n)[0]
        pl = build_chunks_plain(perlin, [coords[i] for i in idx], int(n), collision=True)
        torch.cuda.synchronize()
        pinfo = pl.info()
        for j, i in enumerate(idx):
            v, ix = b.chunk_mesh(int(i), info)
            pv, pix = pl.chunk_mesh(j, pinfo)
            assert np.array_equal(v, pv) and np.array_equal(ix.astype(np.uint32), pix), int(i)
            if len(pix):
                c1, n1 = b.chunk_collision(int(i), info)
                c2, n2 = pl.chunk_collision(j, pinfo)
                assert np.array_equal(c1, c2) and np.array_equal(n1, n2), int(i)
                checked += 1
    assert checked > 8


def test_chunk_facade_tris_around_matches_oracle():
    from underwaterworld_amd import Chunk, Perlin, sample_density
    p = PR.preset("reference")
    perlin = Perlin(p)
    rng = np.random.default_rng(3)
    for c in [(3, 1, -1), (-2, 5, 0)]:
        ch = Chunk(c, colors=True, u16=True)
        ch.build_full(perlin)
        assert ch.not_blank()
        assert ch.verts_buffer_slice().shape[1] == 6 and ch.inds_buffer_slice().dtype == np.uint16
        dens = sample_density(perlin, [c], 12)[0].cpu().numpy()[0]
        rv, rix, rcells, rnrm = O.mesh_ref_tris(p, c, 12, dens)
        tri_pos = rv[rix].reshape(-1, 3, 3)
        for _ in range(25):
            pct = tuple(rng.uniform(-0.2, 1.2, 3))
            r = int(rng.integers(0, 4))
            ids = O.tris_around(rcells, 12, pct, r)
            got = ch.tris_around(pct, r)
            assert len(got) == len(ids)
            for t, tri in zip(ids, got):
                assert np.array_equal(tri.verts, tri_pos[t])
                assert np.abs(tri.normal - rnrm[t]).max() <= NORMAL_TOL

"""GPU VertColor output (position + sRGB colour per vertex, draw.rs:4-9, chunk.rs:215-222) of the
plain and batched paths against the oracle, meshed on the GPU's own density grids so that the
topology and positions are identical and only the colour arithmetic is compared."""
import numpy as np
import pytest

from oracle import oracle as O
from underwaterworld_amd import params as PR

pytestmark = pytest.mark.gpu

# GPU powf and the host libm powf may differ by a few ulp; colours lie in (0, 1).
COLOR_TOL = 2e-6


def _coords():
    return [(3, 1, -1), (0, 0, -1), (-2, 5, 0), (1, -1, -2), (4, 4, 1), (-3, 0, -2)]


def test_plain_colors_match_oracle():
    import torch
    from underwaterworld_amd import Perlin, build_chunks_plain, sample_density
    p = PR.preset("reference")
    perlin = Perlin(p)
    coords = _coords()
    n = 12
    dens, _ = sample_density(perlin, coords, n)
    bc = build_chunks_plain(perlin, coords, n, colors=True)
    bx = build_chunks_plain(perlin, coords, n)
    torch.cuda.synchronize()
    dens = dens.cpu().numpy()
    assert bc.verts.shape[1] == 6
    ic, ix = bc.info(), bx.info()
    nonblank = 0
    for i, c in enumerate(coords):
        vc, idc = bc.chunk_mesh(i, ic)
        vx, idx = bx.chunk_mesh(i, ix)
        assert np.array_equal(vc[:, :3], vx) and np.array_equal(idc, idx), c
        if ic[i]["n_inds"] == 0:
            continue
        nonblank += 1
        rv, rcol, rind = O.mesh_ref_rgb(p, c, n, dens[i])
        assert np.array_equal(vc[:, :3], rv) and np.array_equal(idc, rind), c
        err = np.abs(vc[:, 3:] - rcol).max()
        assert err <= COLOR_TOL, (c, err)
    assert nonblank >= 3


def test_batched_colors_equal_plain():
    import torch
    from underwaterworld_amd import Perlin, build_chunks_batched, build_chunks_plain
    p = PR.preset("mixed")
    perlin = Perlin(p)
    coords, res = PR.mixed_workload(8192, 7)
    coords, res = coords[:64], res[:64]
    b = build_chunks_batched(perlin, coords, res, colors=True)
    torch.cuda.synchronize()
    info = b.info()
    res = np.asarray(res)
    checked = 0
    for n in sorted(set(res.tolist())):
        idx = np.nonzero(res == n)[0]
        pl = build_chunks_plain(perlin, [coords[i] for i in idx], int(n), colors=True)
        torch.cuda.synchronize()
        pinfo = pl.info()
        for j, i in enumerate(idx):
            v, ix = b.chunk_mesh(int(i), info)
            pv, pix = pl.chunk_mesh(j, pinfo)
            assert np.array_equal(v, pv) and np.array_equal(ix, pix), int(i)
            checked += int(len(v) > 0)
    assert checked > 8


def test_bad_vertex_format_rejected():
    import ctypes
    import torch
    from underwaterworld_amd import Perlin, _lib
    from underwaterworld_amd.engine import _alloc_out, _mesh_out, _coords_tensor
    perlin = Perlin(PR.preset("reference"))
    dev = torch.device("cuda", 0)
    ct = _coords_tensor([(0, 0, -1)], dev)
    verts, inds, info, totals = _alloc_out(1, 4096, 4096, dev)
    out = _mesh_out(verts, inds, info, totals, 4096, 4096)
    out.vertex_format = 7
    rc = _lib.lib().uw_build_chunks_plain(perlin.ctx, ctypes.c_void_p(ct.data_ptr()), 1, 12, ctypes.byref(out),
                                          None, None)
    assert rc == _lib.UW_ERR_INVALID_ARG

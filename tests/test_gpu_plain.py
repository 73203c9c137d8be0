"""GPU parity of the plain path (one launch per builder stage) and of the Chunk facade against the
oracle, on the reference preset and the workload presets. Tolerances: tests/parity.py."""
import numpy as np
import pytest

from oracle import oracle as O
from parity import DENS_TOL, edge_density_gaps, position_check, tie_corners
from test_tables import packed_rows
from underwaterworld_amd import params as PR

pytestmark = pytest.mark.gpu


def _grid_sample(lo, hi, every, extra=()):
    coords = PR.grid_coords(lo, hi)
    return list(extra) + coords[::every]


CASES = [
    ("reference", 12, [(0, 0, -1), (0, 0, 0), (1, 2, 0), (-4, 3, -1), (3, 1, -1), (2, -3, 0), (5, 5, -2),
                       (0, 0, -2), (-7, 11, 1), (100, -200, 0), (-1, -1, -1), (0, 0, 2)]),
    ("chunks12", 12, _grid_sample(-8, 8, 97, [(0, 0, -1), (3, 1, -1), (5, 5, -2), (0, 0, 0), (0, 0, -2)])),
    ("mixed", 8, [(0, 0, -1), (3, 1, -1), (-9, 4, 0), (12, -15, -2), (1, 1, 3)]),
    ("mixed", 24, [(0, 0, -1), (3, 1, -1), (-9, 4, 0), (12, -15, -2)]),
    ("mixed", 32, [(0, 0, -1), (3, 1, -1), (-9, 4, 0)]),
    ("reference", 1, [(0, 0, -1), (0, 0, 0), (3, 1, -1)]),
]


def _run_plain(preset, n, coords):
    import torch
    from underwaterworld_amd import Perlin, build_chunks_plain, sample_density
    perlin = Perlin(PR.preset(preset))
    batch = build_chunks_plain(perlin, coords, n, want_cases=True)
    dens, below = sample_density(perlin, coords, n)
    torch.cuda.synchronize()
    return perlin, batch, dens.cpu().numpy(), below.cpu().numpy()


@pytest.mark.parametrize("preset,n,coords", CASES, ids=[f"{c[0]}-n{c[1]}" for c in CASES])
def test_plain_path_parity(preset, n, coords):
    p = PR.preset(preset)
    rows = packed_rows()
    perlin, batch, dens_gpu, below = _run_plain(preset, n, coords)
    info = batch.info()
    cases_gpu = batch.cases.cpu().numpy().reshape(len(coords), n ** 3)
    h = float(np.float32(p["chunk_extent"]) / np.float32(n))
    iso = float(np.float32(p["iso"]))
    max_dens, max_pos, ties, ill, nonblank = 0.0, 0.0, 0, 0, 0
    for k, c in enumerate(coords):
        dref = O.density_ref(p, c, n)
        max_dens = max(max_dens, float(np.max(np.abs(dens_gpu[k].astype(np.float64) - dref))))
        early_ref = O.early_blank(dref, p["iso"])
        r = info[k]
        vg, ig = batch.chunk_mesh(k, info)
        # the engine's mesher, given its own densities, is bit-identical to the oracle's
        mg = O.mesh_ref(p, c, n, dens_gpu[k]) if not (r["flags"] & 2) else None
        if mg is not None:
            assert np.array_equal(cases_gpu[k], mg["cases"])
            assert np.array_equal(ig, mg["inds"]) and np.array_equal(vg, mg["verts"])
            assert r["n_active"] == mg["n_active"]
        if early_ref:
            if not (r["flags"] & 2):
                assert np.any(np.abs(dref - p["iso"]) <= 1e-6)  # only a tie can flip the early skip
            continue
        mr = O.mesh_ref(p, c, n, dref)
        if r["flags"] & 2:
            assert np.any(np.abs(dref - p["iso"]) <= 1e-6)
            continue
        bits, ok = tie_corners(cases_gpu[k], mr["cases"], dref, n, iso)
        assert ok, f"chunk {c}: case bits differ away from the isovalue"
        ties += bits
        if bits:
            continue
        assert r["n_inds"] == len(mr["inds"]) and r["n_verts"] == len(mr["verts"])
        assert np.array_equal(ig, mr["inds"])
        assert bool(r["flags"] & 1) == (len(mr["inds"]) == 0)
        nonblank += len(mr["inds"]) > 0
        if len(mr["verts"]):
            gaps = edge_density_gaps(mr["verts"], mr["inds"], mr["cases"], dref, n, rows)
            e, nill, okp = position_check(vg, mr["verts"], gaps, h)
            assert okp, f"chunk {c}: vertex position beyond tolerance"
            max_pos, ill = max(max_pos, e), ill + nill
    print(f"\n[{preset} n={n}] chunks={len(coords)} nonblank={nonblank} max|drho|={max_dens:.3g} "
          f"max|dp|/h={max_pos:.3g} tie_bits={ties} ill_conditioned_vertices={ill}")
    assert max_dens <= DENS_TOL
    assert nonblank > 0 or n == 1


def test_density_error_vs_fp64():
    import torch
    from underwaterworld_amd import Perlin, sample_density
    p = PR.preset("chunks12")
    coords = [(0, 0, -1), (3, 1, -1), (-5, 2, 0)]
    dens, _ = sample_density(Perlin(p), coords, 12)
    torch.cuda.synchronize()
    d = dens.cpu().numpy().astype(np.float64)
    e64 = max(float(np.max(np.abs(d[k] - O.density_fp64(p, c, 12)))) for k, c in enumerate(coords))
    eref = max(float(np.max(np.abs(O.density_ref(p, c, 12) - O.density_fp64(p, c, 12)))) for c in coords)
    print(f"\nmax |rho_gpu - rho_fp64| = {e64:.3g}; reference-faithful f32 vs fp64: {eref:.3g}")
    assert e64 <= 1e-6


def test_chunk_facade_matches_batch():
    from underwaterworld_amd import Chunk, Perlin
    p = PR.preset("reference", seed=11)
    perlin = Perlin(p)
    coords = [(0, 0, -1), (0, 0, 2), (1, 2, 0)]
    _, batch, _, _ = _run_plain_with(perlin, coords, 12)
    info = batch.info()
    for k, c in enumerate(coords):
        ch = Chunk(c)
        assert ch.build_partial(perlin) is True
        assert ch.num_inds() == int(info[k]["n_inds"])
        assert ch.not_blank() == (info[k]["n_inds"] > 0)
        if ch.not_blank():
            v, i = batch.chunk_mesh(k, info)
            assert np.array_equal(ch.verts_buffer_slice(), v) and np.array_equal(ch.inds_buffer_slice(), i)
        else:
            with pytest.raises(RuntimeError):
                ch.verts_buffer_slice()


def _run_plain_with(perlin, coords, n):
    import torch
    from underwaterworld_amd import build_chunks_plain
    b = build_chunks_plain(perlin, coords, n)
    torch.cuda.synchronize()
    return perlin, b, None, None


def test_overflow_reports_required_sizes():
    import torch
    from underwaterworld_amd import Perlin, build_chunks_plain
    perlin = Perlin(PR.preset("chunks12"))
    coords = [(0, 0, -1), (3, 1, -1), (0, 0, 5)]
    full = build_chunks_plain(perlin, coords, 12)
    small = build_chunks_plain(perlin, coords, 12, vert_cap=100, ind_cap=100)
    torch.cuda.synchronize()
    tf, ts = full.totals(), small.totals()
    assert tf["overflow"] == 0 and ts["overflow"] == 1
    assert tf["n_verts"] == ts["n_verts"] and tf["n_inds"] == ts["n_inds"]
    fi, si = full.info(), small.info()
    assert (si["flags"][:2] & 4).all()
    assert np.array_equal(fi["n_inds"], si["n_inds"])


def test_empty_and_invalid_inputs():
    import torch
    from underwaterworld_amd import Perlin, build_chunks_plain, _lib
    perlin = Perlin(PR.preset("chunks12"))
    b = build_chunks_plain(perlin, np.zeros((0, 3), np.int32), 12)
    torch.cuda.synchronize()
    t = b.totals()
    assert t["n_verts"] == 0 and t["n_inds"] == 0
    with pytest.raises(_lib.UwError):
        build_chunks_plain(perlin, [(0, 0, 0)], 0)
    with pytest.raises(_lib.UwError):
        build_chunks_plain(perlin, [(0, 0, 0)], 65)


def test_full_grid_counts_match_oracle():
    """All 4096 chunks of the 16^3 grid of seed 42: per-chunk blank flag, vertex, index and
    active-cell counts against the oracle's builds."""
    import torch
    from underwaterworld_amd import Perlin, build_chunks_plain
    p = PR.preset("chunks12")
    coords = np.array(PR.grid_coords(-8, 8), np.int32)
    b = build_chunks_plain(Perlin(p), coords, 12)
    torch.cuda.synchronize()
    info = b.info()
    ref = O.build_chunks(p, coords, 12, nthreads=8, dedup="hash")
    same = (info["n_verts"] == ref[:, 0]) & (info["n_inds"] == ref[:, 1]) & (info["n_active"] == ref[:, 2])
    blank_ref = ref[:, 1] == 0
    assert np.array_equal((info["flags"] & 1) != 0, blank_ref) or not same.all()
    mism = np.nonzero(~same)[0]
    for k in mism:  # a mismatch is only allowed where a sample sits within 1e-6 of the isovalue
        d = O.density_ref(p, coords[k], 12)
        assert np.any(np.abs(d - p["iso"]) <= 1e-6), f"chunk {coords[k]} differs without a tie"
    print(f"\n4096 chunks: blank {blank_ref.mean():.3f}, tie-affected chunks {len(mism)}")

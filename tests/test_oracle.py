"""Pins for the CPU oracle (oracle/uw_oracle.c): known answers, mathematical invariants, and an
independent numpy restatement. The reference has no tests or golden vectors of its own and its
Perlin crate (noise 0.8.2) is not present, so these checks are what the oracle rests on; see the
oracle's header for which facts stay unpinned against the crate itself."""
import numpy as np
import pytest

import npref
from oracle import oracle as O
from underwaterworld_amd import params as PR

SCALE = 1.1547005383792515


# --- rand_xorshift / rand::gen_range / PermutationTable ---------------------------------------

def test_xorshift_known_answer():
    # Marsaglia (2003), "Xorshift RNGs", xor128 with its published seed: first output 3701687786.
    out = O.xorshift_stream(123456789, 362436069, 521288629, 88675123, 3)
    assert int(out[0]) == 3701687786


@pytest.mark.parametrize("rng", [1, 2, 3, 5, 7, 12, 100, 255, 256, 1000, 4097, 65535])
def test_widening_multiply_rejection_is_uniform(rng):
    # the accepted 16-bit words map to every output equally often (a wrong zone breaks this)
    h = O.gen_index_histogram16(rng)
    assert h.min() == h.max() > 0


@pytest.mark.parametrize("seed", [0, 1, 7, 42, 1234, 0xFFFFFFFF])
def test_permutation_is_bijective_and_deterministic(seed):
    p = O.perm(seed)
    assert sorted(p.tolist()) == list(range(256))
    assert np.array_equal(p, O.perm(seed))


@pytest.mark.parametrize("seed", [0, 3, 42, 1234, 0xDEADBEEF])
def test_permutation_matches_numpy_restatement(seed):
    assert np.array_equal(O.perm(seed), npref.perm(seed))


def test_permutations_differ_between_seeds():
    assert not np.array_equal(O.perm(1), O.perm(2))


# --- perlin_3d -----------------------------------------------------------------------------------

def _hash(P, x, y, z):
    P = P.astype(np.int64)
    return int(P[(P[(P[x & 255] ^ (y & 255))] ^ (z & 255))])


def test_noise_is_zero_on_lattice():
    P = O.perm(42)
    rng = np.random.default_rng(0)
    pts = rng.integers(-600, 600, size=(500, 3)).astype(np.float64)
    assert np.all(O.perlin3(P, pts) == 0.0)


def test_lattice_gradient_is_hashed_gradient_vector():
    # fade'(0) = 0, so the noise gradient at a lattice point is SCALE * G(hash(corner)).
    P = O.perm(42)
    rng = np.random.default_rng(1)
    h = 1e-4
    for c in rng.integers(-300, 300, size=(40, 3)):
        g = npref.GRAD[_hash(P, *c) & 15]
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            d = (O.perlin3(P, c + e)[0] - O.perlin3(P, c - e)[0]) / (2 * h)
            assert abs(d / SCALE - g[k]) < 1e-6


def test_noise_closed_form_on_lattice_axis():
    # with y, z integral only the x edge of the cell contributes:
    # n = SCALE * ((1 - s(t)) Gx0 t + s(t) Gx1 (t - 1))
    P = O.perm(7)
    rng = np.random.default_rng(2)
    for _ in range(200):
        c = rng.integers(-100, 100, size=3)
        t = rng.random()
        s = t ** 3 * (t * (t * 6 - 15) + 10)
        g0 = npref.GRAD[_hash(P, c[0], c[1], c[2]) & 15][0]
        g1 = npref.GRAD[_hash(P, c[0] + 1, c[1], c[2]) & 15][0]
        want = SCALE * ((1 - s) * g0 * t + s * g1 * (t - 1))
        got = O.perlin3(P, np.array([c[0] + t, c[1], c[2]], np.float64))[0]
        assert abs(got - np.clip(want, -1, 1)) < 1e-12


def test_noise_period_256():
    P = O.perm(42)
    rng = np.random.default_rng(3)
    pts = rng.integers(-4096, 4096, size=(300, 3)) / 1024.0
    base = O.perlin3(P, pts)
    for k in range(3):
        sh = pts.copy()
        sh[:, k] += 256.0
        assert np.array_equal(O.perlin3(P, sh), base)


def test_noise_bounded_and_nontrivial():
    P = O.perm(1234)
    pts = np.random.default_rng(4).uniform(-50, 50, size=(100000, 3))
    v = O.perlin3(P, pts)
    assert np.all(np.abs(v) <= 1.0)
    assert v.std() > 0.1


def test_noise_matches_numpy_restatement():
    P = O.perm(99)
    pts = np.random.default_rng(5).uniform(-300, 300, size=(20000, 3))
    assert np.array_equal(O.perlin3(P, pts), npref.perlin(P, pts))


# --- density (iso_at + build_iso) -------------------------------------------------------------

def test_reference_preset_lattice_sample_is_terrace():
    # every chunk's sample (0,0,0) sits on the lattice: fbm = 0 and rho = adj - fmod(adj, 0.25)
    p = PR.preset("reference", seed=5)
    for cz in (-2, -1, 0, 1, 2):
        d = O.density_ref(p, (3, -4, cz), 12)
        adj = np.float32(cz) * np.float32(16.0) / np.float32(32.0)
        assert d[0] == np.float32((adj + np.float32(0)) - np.fmod(adj, np.float32(0.25)))
    assert O.density_ref(p, (0, 0, -1), 12)[0] == np.float32(-0.5)


def test_workload_preset_lattice_samples_are_linear_bias():
    # 12 c + i = 0 (mod 32) on every axis puts a sample on the lattice of every octave
    p = PR.preset("chunks12")
    for coord, (i, j, k) in [((2, 5, -1), (8, 4, 12)), ((-3, 8, 0), (4, 0, 0)), ((0, 0, -3), (0, 0, 4))]:
        d = O.density_ref(p, coord, 12).reshape(13, 13, 13)
        pz = (12 * coord[2] + k) / 32.0
        assert pz == int(pz)
        want = (np.float32(pz) * np.float32(0.2) + np.float32(0.0)) + np.float32(0.25)
        assert d[i, j, k] == want


def test_sample_coordinates_reference():
    # chunk.rs:107-108 with SIZE_SCALE = 16f32/12f32: local 12 * 1.3333333730697632
    p = PR.preset("reference")
    c = O.coords(p, 0, 12)
    assert c[12] == (12 * 1.3333333730697632) / 16.0
    assert c[12] == 1.0000000298023224
    assert np.all(np.diff(c) > 0)
    assert O.step(p, 12) == np.float32(16.0) / np.float32(12.0)


@pytest.mark.parametrize("preset,coord,n", [("reference", (0, 0, -1), 4), ("reference", (3, -2, 1), 3),
                                             ("chunks12", (0, 0, -1), 4), ("mixed", (-5, 7, 2), 3),
                                             ("chunks12", (1, 1, 0), 5)])
def test_density_matches_numpy_restatement(preset, coord, n):
    p = PR.preset(preset)
    assert np.array_equal(O.density_ref(p, coord, n), npref.density(p, coord, n))


def test_fp64_density_close_to_reference_faithful():
    p = PR.preset("chunks12")
    a = O.density_ref(p, (0, 0, -1), 12).astype(np.float64)
    b = O.density_fp64(p, (0, 0, -1), 12)
    assert np.max(np.abs(a - b)) < 2e-7


# --- mesh (build_mesh) ---------------------------------------------------------------------------

def _one_cell(case, lo=-1.0, hi=1.0):
    d = np.zeros(8, np.float32)
    for i, (dx, dy, dz) in enumerate(npref.CORNERS):
        d[dx * 4 + dy * 2 + dz] = lo if (case >> i) & 1 else hi
    return d


def test_all_256_single_cell_patterns():
    from test_tables import packed_rows
    rows = packed_rows()
    p = PR.preset("chunks12")
    for case in range(256):
        m = O.mesh_ref(p, (0, 0, 0), 1, _one_cell(case))
        assert int(m["cases"][0]) == case
        assert len(m["inds"]) == len(rows[case])
        assert len(m["verts"]) == len(set(rows[case]))


def test_dyadic_lerp_is_exact():
    # rho_a = iso - 1/4, rho_b = iso + 1/4 -> t = 1/2; rho_b = iso + 3/4 -> t = 1/4
    p = PR.preset("chunks12", iso=0.0)
    step = np.float32(12.0)
    for lo, hi, t in [(-0.25, 0.25, 0.5), (-0.25, 0.75, 0.25)]:
        m = O.mesh_ref(p, (0, 0, 0), 1, _one_cell(1, lo, hi))  # corner 0 inside: edges 0, 8, 3
        v = m["verts"]
        # edge 0: (0,0,0)->(1,0,0) -> x = t * step
        assert v[0][0] == np.float32(t) * step and v[0][1] == 0 and v[0][2] == 0


@pytest.mark.parametrize("preset,coord", [("reference", (0, 0, -1)), ("reference", (1, 2, 0)),
                                          ("chunks12", (0, 0, -1)), ("chunks12", (3, 1, -1))])
def test_linear_and_hashed_dedup_agree(preset, coord):
    p = PR.preset(preset)
    d = O.density_ref(p, coord, 12)
    a = O.mesh_ref(p, coord, 12, d, dedup="linear")
    b = O.mesh_ref(p, coord, 12, d, dedup="hash")
    assert np.array_equal(a["verts"], b["verts"]) and np.array_equal(a["inds"], b["inds"])
    assert len(a["inds"]) > 0


def _random_grid(n, seed, frac=0.5):
    g = np.random.default_rng(seed)
    return (g.random((n + 1) ** 3) - frac).astype(np.float32)


@pytest.mark.parametrize("n,seed", [(2, 0), (3, 1), (4, 2), (5, 3)])
def test_mesh_matches_numpy_restatement(n, seed):
    from test_tables import packed_rows
    rows = packed_rows()
    p = PR.preset("chunks12")
    d = _random_grid(n, seed)
    m = O.mesh_ref(p, (1, -1, 2), n, d)
    v, i, c = npref.mesh(p, (1, -1, 2), n, d, rows)
    assert np.array_equal(m["cases"], c)
    assert np.array_equal(m["inds"], i)
    assert np.array_equal(m["verts"], v)


def test_vertex_count_is_distinct_ordered_pairs():
    from test_tables import packed_rows
    rows = packed_rows()
    p = PR.preset("chunks12")
    n = 4
    d = _random_grid(n, 11).reshape(n + 1, n + 1, n + 1)
    keys = set()
    for x in range(n):
        for y in range(n):
            for z in range(n):
                cc = [(x + a, y + b, z + c) for (a, b, c) in npref.CORNERS]
                case = sum(1 << i for i, q in enumerate(cc) if d[q] < 0)
                for e in rows[case]:
                    keys.add((cc[npref.EDGES[e][0]], cc[npref.EDGES[e][1]]))
    m = O.mesh_ref(p, (0, 0, 0), n, d.ravel())
    assert len(m["verts"]) == len(keys)
    assert len(m["inds"]) % 3 == 0

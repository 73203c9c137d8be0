"""Independent numpy restatement of the reference hot path, for small inputs.

Test infrastructure: written separately from oracle/uw_oracle.c (vectorised numpy for the noise,
Python dicts for the dedup) so that a transcription slip in either one shows up as a mismatch.
Follows the same sources: noise 0.8.2 (PermutationTable::new, perlin_3d), perlin_util.rs:6-29,
chunk.rs:105-264, marching_table.rs.
"""
from __future__ import annotations

import numpy as np

GRAD = np.array([
    (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
    (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1), (1, 1, 0), (-1, 1, 0), (0, -1, 1), (0, -1, -1),
], np.float64)

CORNERS = [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)]
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def perm(seed: int) -> np.ndarray:
    """XorShift128 state (1, s, s, s); top-down Fisher-Yates with rand 0.8's widening-multiply
    gen_range."""
    st = [1, seed & 0xFFFFFFFF, seed & 0xFFFFFFFF, seed & 0xFFFFFFFF]
    M = 0xFFFFFFFF

    def nxt():
        x = st[0]
        t = (x ^ (x << 11)) & M
        st[0], st[1], st[2] = st[1], st[2], st[3]
        w = st[3]
        st[3] = (w ^ (w >> 19) ^ (t ^ (t >> 8))) & M
        return st[3]

    a = list(range(256))
    for i in range(255, 0, -1):
        rng = i + 1
        lz = 32 - rng.bit_length()
        zone = ((rng << lz) - 1) & M
        while True:
            m = nxt() * rng
            if (m & M) <= zone:
                j = m >> 32
                break
        a[i], a[j] = a[j], a[i]
    return np.array(a, np.uint8)


def perlin(P: np.ndarray, pts: np.ndarray) -> np.ndarray:
    P = P.astype(np.int64)
    pts = np.asarray(pts, np.float64).reshape(-1, 3)
    fl = np.floor(pts)
    c = fl.astype(np.int64)
    d = pts - fl

    def g(o):
        h = P[(P[(P[(c[:, 0] + o[0]) & 255] ^ ((c[:, 1] + o[1]) & 255))] ^ ((c[:, 2] + o[2]) & 255))]
        v = d - np.array(o, np.float64)
        G = GRAD[h & 15]
        return (G[:, 0] * v[:, 0] + G[:, 1] * v[:, 1]) + G[:, 2] * v[:, 2]

    g000, g100, g010, g110 = g((0, 0, 0)), g((1, 0, 0)), g((0, 1, 0)), g((1, 1, 0))
    g001, g101, g011, g111 = g((0, 0, 1)), g((1, 0, 1)), g((0, 1, 1)), g((1, 1, 1))
    f = d * d * d * (d * (d * 6.0 - 15.0) + 10.0)
    a, b, cc = f[:, 0], f[:, 1], f[:, 2]
    k0 = g000
    k1 = g100 - g000
    k2 = g010 - g000
    k3 = g001 - g000
    k4 = g000 + g110 - g100 - g010
    k5 = g000 + g101 - g100 - g001
    k6 = g000 + g011 - g010 - g001
    k7 = g100 + g010 + g001 + g111 - g000 - g110 - g101 - g011
    r = k0 + k1 * a + k2 * b + k3 * cc + k4 * a * b + k5 * a * cc + k6 * b * cc + k7 * a * b * cc
    return np.clip(r * 1.1547005383792515, -1.0, 1.0)


def axis_coords(p: dict, c: int, n: int, axis: int = 0) -> np.ndarray:
    step = np.float32(p["chunk_extent"]) / np.float32(n)
    i = np.arange(n + 1, dtype=np.float64)
    off = float(p.get("offset", (0.0, 0.0, 0.0))[axis])
    return ((i * np.float64(step) + np.float64(p["chunk_extent"] * c)) + off) * p["frequency"]


def density(p: dict, coord, n: int) -> np.ndarray:
    P = perm(p["seed"])
    ax = [axis_coords(p, coord[k], n, k) for k in range(3)]
    X, Y, Z = np.meshgrid(ax[0], ax[1], ax[2], indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    total = np.zeros(len(pts))
    m = 0.0
    f = 1.0
    for k in range(p["octaves"]):
        total = total + perlin(P, pts * f) * p["amplitudes"][k]
        m += p["amplitudes"][k]
        f *= p["lacunarity"]
    pf = (total / m).astype(np.float32)
    z32 = pts[:, 2].astype(np.float32)
    a, b, c = (np.float32(v) for v in p["bias"])
    if p["bias_kind"] == 0:
        adj = (z32 * a) / b
        return ((adj + pf) - np.fmod(adj, c)).astype(np.float32)
    return ((z32 * a + pf) + b).astype(np.float32)


def mesh(p: dict, coord, n: int, dens: np.ndarray, tri_rows):
    """Marching cubes with the reference's first-seen ordered-pair dedup, Python dicts."""
    L = n + 1
    D = dens.reshape(L, L, L)
    iso = np.float32(p["iso"])
    step = np.float32(p["chunk_extent"]) / np.float32(n)
    off = [np.float32(p["chunk_extent"] * c) for c in coord]
    ids, verts, inds, cases = {}, [], [], []
    for x in range(n):
        for y in range(n):
            for z in range(n):
                cc = [(x + a, y + b, z + c) for (a, b, c) in CORNERS]
                case = sum(1 << i for i, q in enumerate(cc) if D[q] < iso)
                cases.append(case)
                for e in tri_rows[case]:
                    qa, qb = cc[EDGES[e][0]], cc[EDGES[e][1]]
                    key = (qa, qb)
                    if key not in ids:
                        ra, rb = D[qa], D[qb]
                        t = (iso - ra) / (rb - ra)
                        v = []
                        for k in range(3):
                            sa, sb = np.float32(qa[k]) * step, np.float32(qb[k]) * step
                            v.append((sa + t * (sb - sa)) + off[k])
                        ids[key] = len(verts)
                        verts.append(v)
                    inds.append(ids[key])
    return (np.array(verts, np.float32).reshape(-1, 3), np.array(inds, np.uint32), np.array(cases, np.uint8))

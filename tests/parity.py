"""Comparison of engine output with the oracle, shared by the GPU tests, smoke() and bench.py.

Tolerances (FP32 engine vs the reference-faithful oracle, which carries the noise in f64):
  densities      |d_gpu - d_ref| <= 1e-6 absolute
  cube cases     bit-exact, except bits of corners whose oracle density lies within 1e-6 of the
                 isovalue ("ties"): those are counted and reported
  topology       vertex count, index buffer and triangle count bit-exact in every chunk without ties
  positions      the engine and the oracle number vertices identically, so they are compared in
                 order: |p_gpu - p_ref| <= 1e-5 * h (h = cell edge length), or, for edges whose two
                 densities are so close that the 1e-6 density tolerance moves the crossing by
                 more (|rho_b - rho_a| < 0.1), within the bound that tolerance implies:
                 h * 2 * 1e-6 / |rho_b - rho_a| + 1e-5 * h. Such vertices are counted and reported.
  mesher         given the engine's own densities the oracle's f32 mesher reproduces the engine's
                 vertices and indices bit for bit (same operations, same order)
"""
from __future__ import annotations

import numpy as np

DENS_TOL = 1e-6
TIE_TOL = 1e-6
POS_TOL = 1e-5


def tie_corners(case_gpu: np.ndarray, case_ref: np.ndarray, dens_ref: np.ndarray, n: int, iso: float):
    """Return (number of differing case bits, whether all of them are ties)."""
    from npref import CORNERS
    diff = np.nonzero(case_gpu != case_ref)[0]
    if len(diff) == 0:
        return 0, True
    L = n + 1
    D = dens_ref.reshape(L, L, L)
    bits, ok = 0, True
    for cell in diff:
        x, y, z = cell // (n * n), (cell // n) % n, cell % n
        x_or = int(case_gpu[cell]) ^ int(case_ref[cell])
        for i in range(8):
            if (x_or >> i) & 1:
                bits += 1
                dx, dy, dz = CORNERS[i]
                if abs(float(D[x + dx, y + dy, z + dz]) - iso) > TIE_TOL:
                    ok = False
    return bits, ok


def edge_density_gaps(verts_ref: np.ndarray, inds_ref: np.ndarray, cases_ref: np.ndarray, dens_ref: np.ndarray,
                      n: int, rows):
    """|rho_b - rho_a| of the edge behind every oracle vertex (first-seen order)."""
    from npref import CORNERS, EDGES
    L = n + 1
    D = dens_ref.reshape(L, L, L).astype(np.float64)
    gaps = np.full(len(verts_ref), np.inf)
    seen = np.zeros(len(verts_ref), bool)
    k = 0
    for cell in range(n ** 3):
        c = int(cases_ref[cell])
        if not rows[c]:
            continue
        x, y, z = cell // (n * n), (cell // n) % n, cell % n
        for e in rows[c]:
            vid = int(inds_ref[k])
            k += 1
            if not seen[vid]:
                a, b = EDGES[e]
                qa = tuple(np.add((x, y, z), CORNERS[a]))
                qb = tuple(np.add((x, y, z), CORNERS[b]))
                gaps[vid] = abs(D[qb] - D[qa])
                seen[vid] = True
    return gaps


def position_check(vg: np.ndarray, vr: np.ndarray, gaps: np.ndarray, h: float):
    """Returns (max |dp| / h, number of ill-conditioned vertices, all within bound)."""
    err = np.max(np.abs(vg.astype(np.float64) - vr.astype(np.float64)), axis=1) / h
    plain = err <= POS_TOL
    bound = POS_TOL + 2 * DENS_TOL / np.maximum(gaps, 1e-30)
    ill = (~plain) & (gaps < 0.1)
    ok = bool(np.all(plain | (ill & (err <= bound))))
    return (float(err.max()) if len(err) else 0.0), int(ill.sum()), ok

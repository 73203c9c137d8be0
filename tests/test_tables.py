"""The packed marching-cubes table (underwaterworld_amd/csrc/mc_table.h) against the reference's
TRIANGULATION / EDGE_VERTEX_INDICES (underwater_world/src/marching_table.rs:1-272)."""
import hashlib
import os
import re

import numpy as np

from npref import CORNERS, EDGES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "underwaterworld_amd", "csrc", "mc_table.h")

# sha256 of the reference's [256][16] TRIANGULATION values (-1 padded), comma-joined in row order,
# computed from marching_table.rs:16-272.
REFERENCE_SHA256 = "446bbde4d59001da5148084f441baf3405c7e4c3b09b04b39b6db4403fc66db6"


def packed_rows():
    text = open(HEADER).read()
    body = text[text.index("UW_MC_TRI_PACKED[256]"):]
    body = body[body.index("{") + 1: body.index("};")]
    words = [int(w, 16) for w in re.findall(r"0x([0-9a-fA-F]+)ull", body)]
    assert len(words) == 256
    rows = []
    for w in words:
        t = w & 15
        rows.append([(w >> (4 + 4 * j)) & 15 for j in range(3 * t)])
    return rows


def test_unpacked_table_matches_reference_hash():
    rows = packed_rows()
    flat = []
    for r in rows:
        flat += r + [-1] * (16 - len(r))
    assert hashlib.sha256(",".join(map(str, flat)).encode()).hexdigest() == REFERENCE_SHA256


def test_table_structure():
    rows = packed_rows()
    lens = [len(r) for r in rows]
    assert all(l % 3 == 0 and l <= 15 for l in lens)
    assert sum(lens) // 3 == 820
    assert [c for c in range(256) if lens[c] == 0] == [0, 255]
    assert all(0 <= e < 12 for r in rows for e in r)


def test_rows_use_exactly_the_straddling_edges():
    """Every edge whose endpoints lie on opposite sides of the isovalue is used by its case's
    triangles, and no other edge: the property the engine's edge-ownership numbering rests on."""
    rows = packed_rows()
    for c in range(256):
        inside = [(c >> i) & 1 for i in range(8)]
        straddle = {e for e, (a, b) in enumerate(EDGES) if inside[a] != inside[b]}
        assert set(rows[c]) == straddle, c


def test_corner_and_edge_constants():
    text = open(HEADER).read()
    ofs = [int(v) for v in re.search(r"UW_MC_CORNER_OFS\[8\] = \{([^}]*)\}", text).group(1).split(",")]
    assert ofs == [dx | (dy << 1) | (dz << 2) for (dx, dy, dz) in CORNERS]
    a = [int(v) for v in re.search(r"UW_MC_EDGE_A_IDX\[12\] = \{([^}]*)\}", text).group(1).split(",")]
    b = [int(v) for v in re.search(r"UW_MC_EDGE_B_IDX\[12\] = \{([^}]*)\}", text).group(1).split(",")]
    assert list(zip(a, b)) == EDGES


def test_edges_are_unit_axis_steps():
    for a, b in EDGES:
        d = np.abs(np.array(CORNERS[a]) - np.array(CORNERS[b]))
        assert d.sum() == 1

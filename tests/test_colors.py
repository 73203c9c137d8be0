"""The reference's per-vertex terrain colour (chunk.rs:215-222, util.rs:93-153) in the oracle,
pinned against Python's colorsys (an independent HSV implementation) and the sRGB curve's
closed-form end points."""
import colorsys

import numpy as np
import pytest

from oracle import oracle as O
from underwaterworld_amd import params as PR


def _inv_srgb(c):
    """Inverse of ((v / 255 + 0.055) / 1.055) ** 2.4, giving v / 255."""
    return np.asarray(c, np.float64) ** (1.0 / 2.4) * 1.055 - 0.055


def _expected_hsv(world_z, corner_b, size=16.0):
    mix = (world_z / size + 2.0) / 4.0
    hue = (-150.0 + 210.0 * mix) % 360.0
    return hue, 0.6, 0.4 + (corner_b % 3) / 9.0


@pytest.mark.parametrize("world_z", [-40.0, -32.0, -17.5, -3.25, 0.0, 0.1, 7.9, 15.99, 31.0, 32.0])
@pytest.mark.parametrize("corner_b", range(8))
def test_vertex_color_matches_colorsys(world_z, corner_b):
    got = O.vertex_color(world_z, corner_b)
    hue, sat, val = _expected_hsv(world_z, corner_b)
    want = np.array(colorsys.hsv_to_rgb(hue / 360.0, sat, val))
    # to_srgb of the 0..255 scaled rgb
    want = ((want * 255.0 / 255.0 + 0.055) / 1.055) ** 2.4
    assert np.allclose(got, want, rtol=0, atol=2e-6), (got, want)


def test_srgb_end_points():
    # value 0.4 + 2/9 with saturation 0.6 keeps every channel inside (0, 1); the curve maps
    # 0 -> (0.055 / 1.055)^2.4 and 255 -> 1 exactly, and is increasing.
    lo = (0.055 / 1.055) ** 2.4
    cols = np.stack([O.vertex_color(z, b) for z in np.linspace(-32, 32, 33) for b in range(3)])
    assert (cols > lo).all() and (cols < 1.0).all()
    # hue sextant boundaries: world_z giving hue exactly 0 (mix = 150/210) -> r is the max channel
    z0 = (150.0 / 210.0 * 4.0 - 2.0) * 16.0
    c = O.vertex_color(z0, 0)
    assert c[0] > c[1] and abs(c[1] - c[2]) < 1e-6


def test_mesh_colors_decode_to_height_and_corner():
    """Every vertex colour of a reference mesh decodes (through colorsys) to the hue of its world
    height, saturation 0.6 and a value in {0.4, 0.4 + 1/9, 0.4 + 2/9}; positions equal mesh_ref."""
    p = PR.preset("reference")
    for coord in [(3, 1, -1), (0, 0, -1), (-2, 5, 0)]:
        dens = O.density_ref(p, coord, 12)
        v, col, ix = O.mesh_ref_rgb(p, coord, 12, dens)
        m = O.mesh_ref(p, coord, 12, dens)
        assert np.array_equal(v, m["verts"]) and np.array_equal(ix, m["inds"])
        if len(v) == 0:
            continue
        vals = set()
        for z, c in zip(v[:, 2], col):
            r, g, b = _inv_srgb(c)
            h, s, val = colorsys.rgb_to_hsv(r, g, b)
            hue, _, _ = _expected_hsv(float(z), 0)
            dh = abs(h * 360.0 - hue)
            assert min(dh, 360.0 - dh) < 1e-3 and abs(s - 0.6) < 1e-4
            k = int(round((val - 0.4) * 9.0))
            assert k in (0, 1, 2) and abs(val - (0.4 + k / 9.0)) < 1e-5
            vals.add(k)
        assert len(vals) >= 2

"""Density-field and chunk-geometry parameters, with the reference preset and the workload presets.

A parameter set is a plain dict (so that the engine's ctypes struct and the oracle's struct can
both be filled from it):

    seed          u32 Perlin permutation seed (noise::Perlin::new, state.rs:358-359)
    octaves       number of fBm octaves (chunk.rs:9 PERLIN_OCTAVES)
    frequency     noise units per world unit (chunk.rs:108: divide by CHUNK_SIZE)
    lacunarity    per-octave frequency multiplier (perlin_util.rs:18)
    amplitudes    per-octave amplitudes (perlin_util.rs:17: 1, 0.5, 0.25, ...)
    iso           isovalue (chunk.rs:10 ISO_LEVEL)
    bias_kind     0 = terrace  rho = (adj + p) - fmod(adj, c), adj = (z * a) / b  (perlin_util.rs:27-28)
                  1 = linear   rho = (z * a + p) + b
    bias          (a, b, c) of the bias formula above
    chunk_extent  world units per chunk edge (chunk.rs:5 CHUNK_SIZE)
    resolution    cells per chunk axis (chunk.rs:6 INTERNAL_SIZE)

Sample i of chunk c along an axis sits at noise coordinate
    ((f64) i * (f64) step + (f64)(extent * c)) * frequency,   step = (f32) extent / (f32) n
(chunk.rs:107-116). With a power-of-two frequency the product equals the reference's division.
"""
from __future__ import annotations

import copy

MAX_OCTAVES = 16
BIAS_TERRACE = 0
BIAS_LINEAR = 1


def _amps(octaves, persistence=0.5):
    return [persistence ** k for k in range(octaves)]


# The reference crate's constants (chunk.rs:5-12, perlin_util.rs:6-29). The reference seeds from
# the wall clock; the preset takes the seed as a parameter.
REFERENCE = dict(
    seed=0, octaves=3, frequency=1.0 / 16.0, lacunarity=2.0, amplitudes=_amps(3), iso=-0.1,
    bias_kind=BIAS_TERRACE, bias=(16.0, 32.0, 0.25), chunk_extent=16, resolution=12,
)

# Workload presets. One chunk spans 12 world units (one unit per 12^3 voxel) and the base
# frequency is 1/32 per voxel, so every sample coordinate is a dyadic rational at any of the
# resolutions 8, 12, 24, 32. The linear z-bias (slope 0.2, offset 0.25 per noise unit) is the
# committed calibration: on the 16^3-chunk grid of seed 42 it leaves 54.8 % of chunks blank and
# the rest 0.1-39 % active.
_LINEAR_BIAS = (0.2, 0.25, 0.0)

CHUNKS_12 = dict(
    seed=42, octaves=4, frequency=1.0 / 32.0, lacunarity=2.0, amplitudes=_amps(4), iso=0.0,
    bias_kind=BIAS_LINEAR, bias=_LINEAR_BIAS, chunk_extent=12, resolution=12,
)

MIXED = dict(CHUNKS_12, octaves=6, amplitudes=_amps(6))

ADJOINT_24 = dict(CHUNKS_12, octaves=6, amplitudes=_amps(6), resolution=24)

# One 512^3-cell volume (513^3 samples, one world unit per voxel, base frequency 1/32 per voxel):
# a single chunk of extent 512 and resolution 512 at chunk coordinate (0, 0, 0), meshed as one grid
# with one index space. The nearly flat bias (slope 0.01, offset -0.05) leaves about 15 % of the
# cells active (column-sampled calibration of seed 1234).
VOLUME_512 = dict(CHUNKS_12, seed=1234, octaves=8, amplitudes=_amps(8), chunk_extent=512, resolution=512,
                  bias=(0.01, -0.05, 0.0))

# The same field cut into 16^3 independent chunks ("bricks") of 32^3 cells whose sample coordinates
# coincide with the volume's (brick c covers samples 32 c .. 32 c + 32 per axis).
VOLUME_BRICKS = dict(VOLUME_512, chunk_extent=32, resolution=32)


def volume_bricks(bricks_per_axis: int = 16):
    """Brick coordinates of the cubic volume [0, 32 * bricks_per_axis)^3, x-major."""
    r = range(bricks_per_axis)
    return [(x, y, z) for x in r for y in r for z in r]


def preset(name: str, **overrides) -> dict:
    base = {"reference": REFERENCE, "chunks12": CHUNKS_12, "mixed": MIXED, "adjoint24": ADJOINT_24,
            "volume512": VOLUME_512, "volume_bricks": VOLUME_BRICKS}[name]
    p = copy.deepcopy(base)
    p.update(overrides)
    if "octaves" in overrides and "amplitudes" not in overrides:
        p["amplitudes"] = _amps(p["octaves"])
    return p


def grid_coords(lo: int, hi: int):
    """All chunk coordinates of the cube [lo, hi)^3 in x-major order."""
    return [(x, y, z) for x in range(lo, hi) for y in range(lo, hi) for z in range(lo, hi)]


def splitmix64(seed: int):
    """SplitMix64 stream (Steele, Lea, Flood 2014), used to draw workload chunk lists."""
    s = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def mixed_workload(count: int = 8192, seed: int = 7):
    """8192 chunks: a quarter each at 8^3, 12^3, 24^3, 32^3 cells, coordinates uniform over the
    32 x 32 x 8 chunk region x, y in [-16, 16), z in [-4, 4), drawn from SplitMix64(seed).
    Returns (coords list, resolutions list), interleaved so neighbouring entries differ in size."""
    g = splitmix64(seed)
    res_cycle = (8, 12, 24, 32)
    coords, res = [], []
    for i in range(count):
        r = next(g)
        x = (r & 31) - 16
        y = ((r >> 5) & 31) - 16
        z = ((r >> 10) & 7) - 4
        coords.append((int(x), int(y), int(z)))
        res.append(res_cycle[i % 4])
    return coords, res

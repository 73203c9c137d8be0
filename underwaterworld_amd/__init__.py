"""MI355X (gfx950) chunk-meshing engine for the UnderwaterWorld world generator's hot path:
multi-octave Perlin density over each chunk's voxel grid, then marching cubes with the
reference's vertex dedup. Compute runs in hand-written HIP kernels behind the C ABI in
include/uwmesh.h (libuwmesh.so); this package is the host-side mirror of the reference
interface (noise::Perlin + Chunk) and the device-resident multi-chunk forms.
"""
from . import params
from .engine import (Chunk, MeshBatch, Perlin, build_chunks_plain, permutation_table, sample_density)

__all__ = ["params", "Perlin", "Chunk", "MeshBatch", "build_chunks_plain", "sample_density", "permutation_table"]

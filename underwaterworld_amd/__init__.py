"""MI355X (gfx950) chunk-meshing engine for the UnderwaterWorld world generator's hot path:
multi-octave Perlin density over each chunk's voxel grid, then marching cubes with the
reference's vertex dedup. Compute runs in hand-written HIP kernels behind the C ABI in
include/uwmesh.h (libuwmesh.so); this package is the host-side mirror of the reference
interface (noise::Perlin + Chunk) and the device-resident multi-chunk forms.
"""
from . import params
from .engine import (BatchedGraph, Chunk, MeshBatch, Perlin, Tri, batch_capacity, build_chunks_batched,
                     build_chunks_plain, build_volume, density_at, get_batch_tuning, permutation_table,
                     sample_density, set_batch_tuning, tri_ids_around, volume_required_sizes)

__all__ = ["params", "Perlin", "Chunk", "Tri", "MeshBatch", "BatchedGraph", "build_chunks_plain",
           "build_chunks_batched", "batch_capacity", "sample_density", "density_at", "permutation_table",
           "get_batch_tuning", "set_batch_tuning", "tri_ids_around", "build_volume", "volume_required_sizes"]

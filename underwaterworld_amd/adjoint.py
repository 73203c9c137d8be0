"""Differentiable chunk meshing: the engine's forward (batched build with a vertex tape) and its
hand-written adjoint (uw_adjoint_backward), wrapped as a torch.autograd.Function.

    verts = ChunkMesh.apply(amplitudes, offset, mesher)

amplitudes: CPU float64 tensor (octaves,), offset: CPU float64 tensor (3,) (world-space shift of
every sample). The output is the (V, 3) float32 vertex tensor of the whole chunk set, laid out as
the forward placed it (each chunk's range is in mesher.batch.info()). Backward takes dL/dverts and
returns dL/damplitudes and dL/doffset (sums over chunks); per-chunk gradients and, on request, the
dense density gradients stay on the mesher. The mesh topology (cube cases) of the forward is held
fixed, as for any piecewise-smooth function.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .engine import MeshBatch, Perlin, _alloc_out, _coords_tensor, _mesh_out, _res_tensor, _stream_ptr, batch_capacity


class ChunkMesher:
    """Forward/backward state for one chunk set on one Perlin context."""

    def __init__(self, perlin: Perlin, coords, res, want_dens_grad: bool = False, exact_capacity: bool = False):
        """exact_capacity: size the output (and tape) buffers from a counting launch instead of the
        worst case 5 n (n+1)^2 vertices / 15 n^3 indices per chunk (large volumes)."""
        import torch
        self.perlin = perlin
        dev = torch.device("cuda", perlin.device)
        self.coords = _coords_tensor(coords, dev)
        self.nchunks = self.coords.shape[0]
        self.res = _res_tensor(res, self.nchunks, dev)
        rnp = self.res[: self.nchunks].cpu().numpy()
        self.vert_cap, self.ind_cap = batch_capacity(rnp)
        if exact_capacity and self.nchunks:
            self.vert_cap, self.ind_cap = self._required_sizes()
        self.max_samples = int(((rnp + 1) ** 3).max()) if self.nchunks else 1
        self.want_dens_grad = want_dens_grad
        self.batch: MeshBatch | None = None
        self.tape = None
        self.grad_amp = None   # (nchunks, octaves) float64
        self.grad_off = None   # (nchunks, 3) float64
        self.grad_dens = None  # (nchunks, max_samples) float32

    def _required_sizes(self):
        """Vertex and index counts of the set: a build with zero capacity overflows every chunk and
        reports the required totals without writing a mesh."""
        import torch
        dev = self.coords.device
        verts, inds, info, totals = _alloc_out(self.nchunks, 0, 0, dev)
        out = _mesh_out(verts, inds, info, totals, 0, 0)
        rc = _lib.lib().uw_build_chunks_batched(self.perlin.ctx, ctypes.c_void_p(self.coords.data_ptr()),
                                                ctypes.c_void_p(self.res.data_ptr()), self.nchunks, ctypes.byref(out),
                                                None)
        _lib.check(rc, self.perlin.ctx)
        torch.cuda.synchronize()
        t = MeshBatch(verts, inds, info, totals, None).totals()
        return int(t["n_verts"]), int(t["n_inds"])

    def set_field(self, amplitudes=None, offset=None):
        a = None if amplitudes is None else np.ascontiguousarray(amplitudes, np.float64)
        o = None if offset is None else np.ascontiguousarray(offset, np.float64)
        rc = _lib.lib().uw_ctx_set_field(self.perlin.ctx, a.ctypes.data_as(ctypes.c_void_p) if a is not None else None,
                                         o.ctypes.data_as(ctypes.c_void_p) if o is not None else None)
        _lib.check(rc, self.perlin.ctx)

    def forward(self, stream=None) -> MeshBatch:
        import torch
        dev = self.coords.device
        verts, inds, info, totals = _alloc_out(self.nchunks, self.vert_cap, self.ind_cap, dev)
        self.tape = torch.empty((max(self.vert_cap, 1), 4), dtype=torch.int32, device=dev)
        out = _mesh_out(verts, inds, info, totals, self.vert_cap, self.ind_cap)
        rc = _lib.lib().uw_build_chunks_batched_tape(self.perlin.ctx, ctypes.c_void_p(self.coords.data_ptr()),
                                                     ctypes.c_void_p(self.res.data_ptr()), self.nchunks,
                                                     ctypes.byref(out), ctypes.c_void_p(self.tape.data_ptr()),
                                                     _stream_ptr(stream))
        _lib.check(rc, self.perlin.ctx)
        self.batch = MeshBatch(verts, inds, info[: self.nchunks * 32], totals, None)
        return self.batch

    def backward(self, grad_verts, stream=None):
        import torch
        dev = self.coords.device
        o = int(self.perlin.params["octaves"])
        gv = grad_verts.to(device=dev, dtype=torch.float32).contiguous()
        if gv.shape[0] < self.vert_cap:
            full = torch.zeros((self.vert_cap, 3), dtype=torch.float32, device=dev)
            full[: gv.shape[0]] = gv
            gv = full
        self.grad_amp = torch.empty((max(self.nchunks, 1), o), dtype=torch.float64, device=dev)
        self.grad_off = torch.empty((max(self.nchunks, 1), 3), dtype=torch.float64, device=dev)
        gd_ptr, stride = None, 0
        if self.want_dens_grad:
            self.grad_dens = torch.zeros((max(self.nchunks, 1), self.max_samples), dtype=torch.float32, device=dev)
            gd_ptr, stride = ctypes.c_void_p(self.grad_dens.data_ptr()), self.max_samples
        rc = _lib.lib().uw_adjoint_backward(self.perlin.ctx, ctypes.c_void_p(self.coords.data_ptr()),
                                            ctypes.c_void_p(self.res.data_ptr()), self.nchunks,
                                            ctypes.c_void_p(self.batch.info_raw.data_ptr()),
                                            ctypes.c_void_p(self.tape.data_ptr()), ctypes.c_void_p(gv.data_ptr()),
                                            gd_ptr, stride, ctypes.c_void_p(self.grad_amp.data_ptr()),
                                            ctypes.c_void_p(self.grad_off.data_ptr()), _stream_ptr(stream))
        _lib.check(rc, self.perlin.ctx)
        return self.grad_amp[: self.nchunks], self.grad_off[: self.nchunks]


def _autograd():
    import torch

    class ChunkMesh(torch.autograd.Function):
        @staticmethod
        def forward(ctx, amplitudes, offset, mesher: ChunkMesher):
            mesher.set_field(amplitudes.detach().cpu().numpy(), offset.detach().cpu().numpy())
            b = mesher.forward()
            n = int(b.totals()["n_verts"])
            ctx.mesher = mesher
            ctx.amp_device, ctx.off_device = amplitudes.device, offset.device
            return b.verts[:n]

        @staticmethod
        def backward(ctx, grad_verts):
            ga, go = ctx.mesher.backward(grad_verts)
            return (ga.sum(0).to(ctx.amp_device), go.sum(0).to(ctx.off_device), None)

    return ChunkMesh


ChunkMesh = None


def chunk_mesh(amplitudes, offset, mesher: ChunkMesher):
    """verts = f(amplitudes, offset) with the engine's adjoint as backward."""
    global ChunkMesh
    if ChunkMesh is None:
        ChunkMesh = _autograd()
    return ChunkMesh.apply(amplitudes, offset, mesher)


def plane_loss(verts, mesher: ChunkMesher, target: float):
    """sum over vertices of (v_z - chunk_origin_z - target)^2, the workload's loss; chunk origins
    are looked up from the forward's per-chunk vertex ranges."""
    import torch
    info = mesher.batch.info()
    dev = verts.device
    oz = torch.zeros(verts.shape[0], dtype=torch.float32, device=dev)
    coords = mesher.coords.cpu().numpy()
    ext = float(mesher.perlin.params["chunk_extent"])
    for i in range(mesher.nchunks):
        nv = int(info[i]["n_verts"])
        if nv:
            v0 = int(info[i]["vert_off"])
            oz[v0:v0 + nv] = ext * float(coords[i][2])
    return ((verts[:, 2] - oz - target) ** 2).sum()

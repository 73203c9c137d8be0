"""Differentiable chunk meshing: the engine's forward (batched build with a vertex tape) and its
hand-written adjoint (uw_adjoint_backward), wrapped as a torch.autograd.Function.

    verts = ChunkMesh.apply(amplitudes, offset, mesher)

amplitudes: CPU float64 tensor (octaves,), offset: CPU float64 tensor (3,) (world-space shift of
every sample). The output is the (V, 3) float32 vertex tensor of the whole chunk set, laid out as
the forward placed it (each chunk's range is in the forward record's batch.info()). Backward takes
dL/dverts and returns dL/damplitudes and dL/doffset (sums over chunks); per-chunk gradients and, on
request, the dense density gradients are returned by ChunkMesher.backward. The mesh topology (cube
cases) of the forward is held fixed, as for any piecewise-smooth function.

Every forward returns its own ForwardRecord (output buffers, vertex tape and the field it was built
with), and the backward of a record uses that record only: several forwards on one mesher may be
outstanding at once, and the context's field is set to the record's for the backward launch and
restored afterwards, so neither call changes the field that later builds on the context see.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .engine import MeshBatch, Perlin, _alloc_out, _coords_tensor, _mesh_out, _res_tensor, _stream_ptr, batch_capacity


@dataclass
class ForwardRecord:
    """What one forward produced and what its backward needs."""
    batch: MeshBatch
    tape: "torch.Tensor"       # (vert_cap, 4) int32: uw_vertex_tape per vertex
    amplitudes: np.ndarray     # field the forward was built with
    offset: np.ndarray


@dataclass
class Gradients:
    """Per-chunk gradients of one backward."""
    grad_amp: "torch.Tensor"          # (nchunks, octaves) float64
    grad_off: "torch.Tensor"          # (nchunks, 3) float64
    grad_dens: "torch.Tensor | None"  # (nchunks, max_samples) float32, or None


class _FieldScope:
    """Sets the context's field for the duration of a launch and restores the previous one. Kernel
    launches take the field by value, so restoring right after an asynchronous launch is safe."""

    def __init__(self, perlin: Perlin, amplitudes, offset):
        self.perlin, self.amps, self.off = perlin, amplitudes, offset

    def __enter__(self):
        self.saved = self.perlin.field()
        # a part equal to the current field is left alone (a backward usually runs on the field of
        # its forward, which is still the context's): no library call, nothing to restore
        if self.amps is not None and np.array_equal(np.asarray(self.amps, np.float64), self.saved[0]):
            self.amps = None
        if self.off is not None and np.array_equal(np.asarray(self.off, np.float64), self.saved[1]):
            self.off = None
        if self.amps is not None or self.off is not None:
            self.perlin.set_field(self.amps, self.off)
        return self

    def __exit__(self, *exc):
        if self.amps is not None or self.off is not None:
            self.perlin.set_field(*self.saved)
        return False


class ChunkMesher:
    """Forward/backward launches for one chunk set on one Perlin context."""

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
        self.last: ForwardRecord | None = None  # the latest forward (for callers that keep one)
        self.grads: Gradients | None = None     # the latest backward

    # convenience views of the latest forward / backward
    @property
    def batch(self):
        return self.last.batch if self.last else None

    @property
    def tape(self):
        return self.last.tape if self.last else None

    @property
    def grad_amp(self):
        return self.grads.grad_amp if self.grads else None

    @property
    def grad_off(self):
        return self.grads.grad_off if self.grads else None

    @property
    def grad_dens(self):
        return self.grads.grad_dens if self.grads else None

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
        """Change the context's field for every later build (Perlin.set_field)."""
        self.perlin.set_field(amplitudes, offset)

    def forward(self, stream=None, amplitudes=None, offset=None) -> ForwardRecord:
        """Batched build with a vertex tape. amplitudes / offset (optional) give the field of this
        build only; the context's own field is unchanged afterwards."""
        import torch
        dev = self.coords.device
        verts, inds, info, totals = _alloc_out(self.nchunks, self.vert_cap, self.ind_cap, dev)
        tape = torch.empty((max(self.vert_cap, 1), 4), dtype=torch.int32, device=dev)
        out = _mesh_out(verts, inds, info, totals, self.vert_cap, self.ind_cap)
        with _FieldScope(self.perlin, amplitudes, offset):
            amps, off = self.perlin.field()
            rc = _lib.lib().uw_build_chunks_batched_tape(self.perlin.ctx, ctypes.c_void_p(self.coords.data_ptr()),
                                                         ctypes.c_void_p(self.res.data_ptr()), self.nchunks,
                                                         ctypes.byref(out), ctypes.c_void_p(tape.data_ptr()),
                                                         _stream_ptr(stream))
            _lib.check(rc, self.perlin.ctx)
        rec = ForwardRecord(MeshBatch(verts, inds, info[: self.nchunks * 32], totals, None), tape, amps, off)
        self.last = rec
        return rec

    def backward(self, grad_verts, stream=None, record: ForwardRecord | None = None) -> Gradients:
        """Adjoint of `record` (default: the latest forward) for dL/dverts laid out like its verts."""
        import torch
        rec = record if record is not None else self.last
        if rec is None:
            raise RuntimeError("backward() before any forward()")
        dev = self.coords.device
        o = int(self.perlin.params["octaves"])
        gv = grad_verts.to(device=dev, dtype=torch.float32).contiguous()
        if gv.shape[0] < self.vert_cap:
            full = torch.zeros((self.vert_cap, 3), dtype=torch.float32, device=dev)
            full[: gv.shape[0]] = gv
            gv = full
        g = Gradients(torch.empty((max(self.nchunks, 1), o), dtype=torch.float64, device=dev),
                      torch.empty((max(self.nchunks, 1), 3), dtype=torch.float64, device=dev), None)
        gd_ptr, stride = None, 0
        if self.want_dens_grad:
            # written by the backward (zeroed there first)
            g.grad_dens = torch.empty((max(self.nchunks, 1), self.max_samples), dtype=torch.float32, device=dev)
            gd_ptr, stride = ctypes.c_void_p(g.grad_dens.data_ptr()), self.max_samples
        with _FieldScope(self.perlin, rec.amplitudes, rec.offset):
            rc = _lib.lib().uw_adjoint_backward(self.perlin.ctx, ctypes.c_void_p(self.coords.data_ptr()),
                                                ctypes.c_void_p(self.res.data_ptr()), self.nchunks,
                                                ctypes.c_void_p(rec.batch.info_raw.data_ptr()),
                                                ctypes.c_void_p(rec.tape.data_ptr()), ctypes.c_void_p(gv.data_ptr()),
                                                gd_ptr, stride, ctypes.c_void_p(g.grad_amp.data_ptr()),
                                                ctypes.c_void_p(g.grad_off.data_ptr()), _stream_ptr(stream))
            _lib.check(rc, self.perlin.ctx)
        g.grad_amp = g.grad_amp[: self.nchunks]
        g.grad_off = g.grad_off[: self.nchunks]
        self.grads = g
        return g


class VolumeMesher:
    """Forward/backward launches for one large volume (uw_build_volume / uw_volume_backward): a
    chunk of n cells per axis at `chunk`, meshed as one grid. Output buffers are sized exactly
    (a counting build at construction, with the context's field at that time)."""

    def __init__(self, perlin: Perlin, n: int, chunk=(0, 0, 0), want_dens_grad: bool = False,
                 vert_cap: int | None = None, ind_cap: int | None = None):
        from .engine import volume_required_sizes
        self.perlin, self.n, self.chunk = perlin, int(n), tuple(int(v) for v in chunk)
        if vert_cap is None or ind_cap is None:
            v, i = volume_required_sizes(perlin, self.n, self.chunk)
            vert_cap = v if vert_cap is None else vert_cap
            ind_cap = i if ind_cap is None else ind_cap
        self.vert_cap, self.ind_cap = int(vert_cap), int(ind_cap)
        self.want_dens_grad = want_dens_grad
        self.last: ForwardRecord | None = None
        self.grads: Gradients | None = None

    def forward(self, stream=None, amplitudes=None, offset=None) -> ForwardRecord:
        import torch
        dev = torch.device("cuda", self.perlin.device)
        verts, inds, info, totals = _alloc_out(1, self.vert_cap, self.ind_cap, dev)
        tape = torch.empty((max(self.vert_cap, 1), 4), dtype=torch.int32, device=dev)
        out = _mesh_out(verts, inds, info, totals, self.vert_cap, self.ind_cap)
        with _FieldScope(self.perlin, amplitudes, offset):
            amps, off = self.perlin.field()
            rc = _lib.lib().uw_build_volume(self.perlin.ctx, *self.chunk, self.n, ctypes.byref(out),
                                            ctypes.c_void_p(tape.data_ptr()), _stream_ptr(stream))
            _lib.check(rc, self.perlin.ctx)
        rec = ForwardRecord(MeshBatch(verts, inds, info[:32], totals, None), tape, amps, off)
        self.last = rec
        return rec

    def backward(self, grad_verts, stream=None, record: ForwardRecord | None = None) -> Gradients:
        """dL/d(amplitudes) (1, octaves), dL/d(offset) (1, 3) and optionally dL/d(densities)
        (1, (n+1)^3, index x (n+1)^2 + y (n+1) + z) for dL/dverts laid out like the record's verts."""
        import torch
        rec = record if record is not None else self.last
        if rec is None:
            raise RuntimeError("backward() before any forward()")
        dev = torch.device("cuda", self.perlin.device)
        o = int(self.perlin.params["octaves"])
        gv = grad_verts.to(device=dev, dtype=torch.float32).contiguous()
        if gv.shape[0] < self.vert_cap:
            full = torch.zeros((self.vert_cap, 3), dtype=torch.float32, device=dev)
            full[: gv.shape[0]] = gv
            gv = full
        g = Gradients(torch.empty((1, o), dtype=torch.float64, device=dev),
                      torch.empty((1, 3), dtype=torch.float64, device=dev), None)
        gd = None
        if self.want_dens_grad:
            g.grad_dens = torch.zeros((1, (self.n + 1) ** 3), dtype=torch.float32, device=dev)
            gd = ctypes.c_void_p(g.grad_dens.data_ptr())
        with _FieldScope(self.perlin, rec.amplitudes, rec.offset):
            rc = _lib.lib().uw_volume_backward(self.perlin.ctx, *self.chunk, self.n,
                                               ctypes.c_void_p(rec.batch.info_raw.data_ptr()),
                                               ctypes.c_void_p(rec.tape.data_ptr()), ctypes.c_void_p(gv.data_ptr()),
                                               gd, ctypes.c_void_p(g.grad_amp.data_ptr()),
                                               ctypes.c_void_p(g.grad_off.data_ptr()), _stream_ptr(stream))
            _lib.check(rc, self.perlin.ctx)
        self.grads = g
        return g


def _autograd():
    import torch

    class ChunkMesh(torch.autograd.Function):
        @staticmethod
        def forward(ctx, amplitudes, offset, mesher: ChunkMesher):
            rec = mesher.forward(amplitudes=amplitudes.detach().cpu().numpy(), offset=offset.detach().cpu().numpy())
            n = int(rec.batch.totals()["n_verts"])
            ctx.mesher, ctx.record = mesher, rec
            ctx.amp_device, ctx.off_device = amplitudes.device, offset.device
            return rec.batch.verts[:n]

        @staticmethod
        def backward(ctx, grad_verts):
            g = ctx.mesher.backward(grad_verts, record=ctx.record)
            return (g.grad_amp.sum(0).to(ctx.amp_device), g.grad_off.sum(0).to(ctx.off_device), None)

    return ChunkMesh


ChunkMesh = None


def chunk_mesh(amplitudes, offset, mesher: ChunkMesher):
    """verts = f(amplitudes, offset) with the engine's adjoint as backward."""
    global ChunkMesh
    if ChunkMesh is None:
        ChunkMesh = _autograd()
    return ChunkMesh.apply(amplitudes, offset, mesher)


def plane_loss(verts, mesher: ChunkMesher, target: float, record: ForwardRecord | None = None):
    """sum over vertices of (v_z - chunk_origin_z - target)^2, the workload's loss; chunk origins
    are looked up from the forward's per-chunk vertex ranges (default: the latest forward)."""
    import torch
    rec = record if record is not None else mesher.last
    info = rec.batch.info()
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

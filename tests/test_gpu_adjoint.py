"""GPU adjoint (uw_adjoint_backward) against the FP64 oracle.

Two levels:
  * kernel arithmetic: given the engine's own densities and cube cases, the oracle's FP64 gradient
    (densities, octave amplitudes, world-space offset) and the kernel's agree to 1e-4 relative;
  * end to end: the kernel's amplitude and offset gradients against central finite differences
    (base step h = 1e-3, Richardson-extrapolated) of the FP64 loss with the same cube cases held
    fixed, to 1e-4 relative.
The loss is the workload's: sum over vertices of (v_z - chunk_origin_z - 0.25 H)^2, H = chunk height.
"""
import numpy as np
import pytest

from oracle import oracle as O
from underwaterworld_amd import params as PR

pytestmark = pytest.mark.gpu

CORNERS = [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)]


def _cases(d, n, iso):
    D = d.reshape(n + 1, n + 1, n + 1) < np.float32(iso)
    c = np.zeros((n, n, n), np.uint8)
    for i, (a, b, e) in enumerate(CORNERS):
        c |= (D[a:a + n, b:b + n, e:e + n].astype(np.uint8) << i)
    return c.ravel()


def _run(preset, coords, res, want_dens=True):
    import torch
    from underwaterworld_amd import Perlin, sample_density
    from underwaterworld_amd.adjoint import ChunkMesher, chunk_mesh, plane_loss
    p = PR.preset(preset)
    perlin = Perlin(p)
    m = ChunkMesher(perlin, coords, res, want_dens_grad=want_dens)
    amps = torch.tensor(p["amplitudes"], dtype=torch.float64, requires_grad=True)
    off = torch.zeros(3, dtype=torch.float64, requires_grad=True)
    verts = chunk_mesh(amps, off, m)
    target = 0.25 * p["chunk_extent"]
    loss = plane_loss(verts, m, target)
    loss.backward()
    torch.cuda.synchronize()
    dens = {}
    for n in sorted(set(res)):
        idx = [i for i, r in enumerate(res) if r == n]
        d, _ = sample_density(perlin, [coords[i] for i in idx], n)
        torch.cuda.synchronize()
        d = d.cpu().numpy()
        for j, i in enumerate(idx):
            dens[i] = d[j]
    return p, m, amps, off, float(loss), dens, target


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


COORDS = [(0, 0, -1), (3, 1, -1), (-5, 2, 0), (1, -4, -1), (6, 6, -2), (-2, -7, 1)]


def test_adjoint_kernel_matches_fp64_on_same_inputs():
    res = [24, 24, 12, 8, 24, 16]
    p, m, amps, off, loss, dens, target = _run("adjoint24", COORDS, res)
    ga = m.grad_amp.cpu().numpy()
    go = m.grad_off.cpu().numpy()
    gd = m.grad_dens.cpu().numpy()
    info = m.batch.info()
    checked = 0
    worst = {"amp": 0.0, "off": 0.0, "dens": 0.0}
    for i, (c, n) in enumerate(zip(COORDS, res)):
        if info[i]["n_verts"] == 0:
            continue
        cs = _cases(dens[i], n, p["iso"])
        r = O.adjoint_chunk(p, c, n, target, cases=cs, rho=dens[i].astype(np.float64))
        assert r["nverts"] == info[i]["n_verts"]
        S = (n + 1) ** 3
        worst["amp"] = max(worst["amp"], _rel(ga[i], r["g_amp"]))
        worst["off"] = max(worst["off"], _rel(go[i], r["g_off"]))
        worst["dens"] = max(worst["dens"], _rel(gd[i][:S], r["g_dens"]))
        checked += 1
    print(f"\nadjoint vs FP64 on identical densities/topology: {worst}")
    assert checked >= 4
    assert worst["dens"] <= 1e-4 and worst["amp"] <= 1e-4 and worst["off"] <= 1e-4
    # autograd: the amplitude/offset gradients are the per-chunk sums
    assert np.allclose(amps.grad.numpy(), ga[: len(COORDS)].sum(0), rtol=1e-12, atol=0)
    assert np.allclose(off.grad.numpy(), go[: len(COORDS)].sum(0), rtol=1e-12, atol=0)


def _central(f, h):
    return (f(h) - f(-h)) / (2 * h)


def test_adjoint_matches_finite_differences():
    coords, res = COORDS[:3], [24, 24, 12]
    p, m, amps, off, loss, dens, target = _run("adjoint24", coords, res, want_dens=False)
    ga = m.grad_amp.cpu().numpy()
    go = m.grad_off.cpu().numpy()
    info = m.batch.info()
    h = 1e-3
    worst = 0.0
    for i, (c, n) in enumerate(zip(coords, res)):
        if info[i]["n_verts"] == 0:
            continue
        cs = _cases(dens[i], n, p["iso"])

        def loss_at(amp=None, ax=None):
            def f(d):
                q = dict(p, amplitudes=list(p["amplitudes"]))
                if amp is not None:
                    q["amplitudes"][amp] += d
                if ax is not None:
                    o = [0.0, 0.0, 0.0]
                    o[ax] = d
                    q["offset"] = tuple(o)
                return O.adjoint_chunk(q, c, n, target, cases=cs, grads=False)["loss"]
            return f

        fd = [(4 * _central(loss_at(amp=j), h / 2) - _central(loss_at(amp=j), h)) / 3 for j in range(p["octaves"])]
        fo = [(4 * _central(loss_at(ax=k), h / 2) - _central(loss_at(ax=k), h)) / 3 for k in range(3)]
        g = np.concatenate([ga[i], go[i]])
        f = np.array(fd + fo)
        rel = float(np.max(np.abs(g - f) / np.maximum(np.abs(f), 1e-3 * np.abs(f).max())))
        worst = max(worst, rel)
    print(f"\nadjoint vs central differences (h=1e-3, Richardson): max relative error {worst:.3g}")
    assert worst <= 1e-4


def _grad_z(mesher, rec):
    import torch
    gv = torch.zeros((mesher.vert_cap, 3), dtype=torch.float32, device="cuda")
    nv = int(rec.batch.totals()["n_verts"])
    gv[:nv, 2] = 1.0
    gv[:nv, 0] = 0.5
    return gv


def test_two_outstanding_forwards_on_one_mesher():
    # two forwards with different fields on one mesher, backward of the first after the second:
    # each backward must use its own forward's tape and field (compared with separate meshers),
    # and neither call may change the context's own field
    import torch
    from underwaterworld_amd import Perlin, build_chunks_batched
    from underwaterworld_amd.adjoint import ChunkMesher
    p = PR.preset("adjoint24")
    coords, res = COORDS[:4], [24, 12, 12, 8]
    a1 = np.array(p["amplitudes"], np.float64)
    a2 = a1 * np.array([1.0, 0.7, 1.3, 0.9, 1.1, 0.8])
    o2 = np.array([0.3, -0.2, 0.1])
    shared = ChunkMesher(Perlin(p), coords, res, want_dens_grad=True)
    before = build_chunks_batched(shared.perlin, coords, res)
    r1 = shared.forward(amplitudes=a1)
    r2 = shared.forward(amplitudes=a2, offset=o2)
    g1 = shared.backward(_grad_z(shared, r1), record=r1)
    g2 = shared.backward(_grad_z(shared, r2), record=r2)
    after = build_chunks_batched(shared.perlin, coords, res)
    torch.cuda.synchronize()
    assert np.allclose(shared.perlin.field()[0], a1) and np.all(shared.perlin.field()[1] == 0)
    for want, got in ((before, after),):
        assert np.array_equal(want.totals(), got.totals())
    for amps, off, rec, g in ((a1, None, r1, g1), (a2, o2, r2, g2)):
        sep = ChunkMesher(Perlin(p), coords, res, want_dens_grad=True)
        sep.set_field(amps, off)
        rs = sep.forward()
        gs = sep.backward(_grad_z(sep, rs), record=rs)
        torch.cuda.synchronize()
        ti, si = rec.batch.info(), rs.batch.info()
        assert np.array_equal(ti["n_verts"], si["n_verts"]) and np.array_equal(ti["n_inds"], si["n_inds"])
        for i in range(len(coords)):
            assert np.array_equal(rec.batch.chunk_mesh(i, ti)[0], rs.batch.chunk_mesh(i, si)[0])
        assert np.allclose(g.grad_amp.cpu().numpy(), gs.grad_amp.cpu().numpy(), rtol=1e-12, atol=0)
        assert np.allclose(g.grad_off.cpu().numpy(), gs.grad_off.cpu().numpy(), rtol=1e-12, atol=0)
        assert np.allclose(g.grad_dens.cpu().numpy(), gs.grad_dens.cpu().numpy(), rtol=1e-5, atol=1e-7)
    assert not np.allclose(g1.grad_amp.cpu().numpy(), g2.grad_amp.cpu().numpy())


@pytest.mark.parametrize("copies", [1, 87, 400])
def test_backward_launch_forms(copies):
    # The backward cuts each chunk's vertices into pieces of 128 and gives every workgroup of its
    # persistent grid ceil(pieces / grid) consecutive pieces. 6 chunks: one piece per workgroup;
    # 6 x 87 chunks of 8^3: still about one; 6 x 400: several per workgroup, so workgroups walk
    # across chunk boundaries and leave one partial sum per chunk they touch. Every form must give
    # each chunk the gradients of the 6-chunk launch (f64 sums in another order: 1e-12).
    import torch
    from underwaterworld_amd import Perlin
    from underwaterworld_amd.adjoint import ChunkMesher
    p = PR.preset("adjoint24")
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    runs = []
    for cs in (COORDS, COORDS * copies):
        m = ChunkMesher(Perlin(p), cs, 8, want_dens_grad=True)
        rec = m.forward()
        g = m.backward(_grad_z(m, rec), record=rec)
        torch.cuda.synchronize()
        nv = rec.batch.info()["n_verts"].astype(np.int64)
        runs.append((g.grad_amp.cpu().numpy(), g.grad_off.cpu().numpy(), g.grad_dens.cpu().numpy(),
                     int(((nv + 127) // 128).sum())))
    (a0, o0, d0, _), (a1, o1, d1, pieces) = runs
    if copies == 400:  # more pieces than the grid has workgroups (at most 8 per CU)
        assert pieces > 8 * cus, (pieces, cus)
    assert np.any(a0 != 0)
    for k in range(len(COORDS) * copies):
        j = k % len(COORDS)
        assert np.allclose(a1[k], a0[j], rtol=1e-12, atol=1e-300), k
        assert np.allclose(o1[k], o0[j], rtol=1e-12, atol=1e-300), k
        assert np.allclose(d1[k], d0[j], rtol=1e-5, atol=1e-6), k


def test_capturing_stream_is_refused():
    # entry points whose launches bind context-owned scratch refuse a capturing stream (a later call
    # may reallocate that scratch under a captured graph); nothing is recorded into the graph
    import ctypes
    import torch
    from underwaterworld_amd import Perlin, _lib
    from underwaterworld_amd.engine import _alloc_out, _coords_tensor, _mesh_out, _res_tensor
    perlin = Perlin(PR.preset("chunks12"))
    ct = _coords_tensor(COORDS, torch.device("cuda", 0))
    rt = _res_tensor(12, len(COORDS), torch.device("cuda", 0))
    verts, inds, info, totals = _alloc_out(len(COORDS), 60000, 200000, torch.device("cuda", 0))
    out = _mesh_out(verts, inds, info, totals, 60000, 200000)
    L = _lib.lib()
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    rcs = {}
    with torch.cuda.graph(g, stream=s):
        sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        rcs["batched"] = L.uw_build_chunks_batched(perlin.ctx, ctypes.c_void_p(ct.data_ptr()),
                                                   ctypes.c_void_p(rt.data_ptr()), len(COORDS), ctypes.byref(out), sp)
        rcs["plain"] = L.uw_build_chunks_plain(perlin.ctx, ctypes.c_void_p(ct.data_ptr()), len(COORDS), 12,
                                               ctypes.byref(out), ctypes.c_void_p(0), sp)
        verts.zero_()  # the capture holds one ordinary node
    torch.cuda.synchronize()
    assert rcs == {"batched": _lib.UW_ERR_INVALID_ARG, "plain": _lib.UW_ERR_INVALID_ARG}, rcs
    assert b"capturing" in L.uw_last_error(perlin.ctx)
    # outside capture the same calls work
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.uw_build_chunks_batched(perlin.ctx, ctypes.c_void_p(ct.data_ptr()), ctypes.c_void_p(rt.data_ptr()),
                                     len(COORDS), ctypes.byref(out), sp) == _lib.UW_OK
    torch.cuda.synchronize()

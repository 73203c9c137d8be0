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

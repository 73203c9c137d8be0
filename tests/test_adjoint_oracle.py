"""Pins for the FP64 adjoint oracle: the analytic gradients of the plane loss against central finite
differences with the mesh topology frozen (the loss is piecewise smooth; a perturbation that
flips a cube case leaves the piece, so the cases of the unperturbed field are held fixed), and the
Perlin gradient against finite differences of the noise itself."""
import numpy as np
import pytest

from oracle import oracle as O
from underwaterworld_amd import params as PR


def _cases_fp64(p, coord, n):
    d = O.density_fp64(p, coord, n).reshape(n + 1, n + 1, n + 1)
    B = d < p["iso"]
    cases = np.zeros((n, n, n), np.uint8)
    corners = [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)]
    for i, (a, b, c) in enumerate(corners):
        cases |= (B[a:a + n, b:b + n, c:c + n].astype(np.uint8) << i)
    return cases.ravel()


def test_perlin_gradient_matches_finite_differences():
    P = O.perm(42)
    rng = np.random.default_rng(0)
    h = 1e-6
    for p in rng.uniform(-40, 40, size=(200, 3)):
        v, g = O.perlin3_grad(P, *p)
        assert abs(v - O.perlin3(P, p)[0]) < 1e-15
        if abs(v) >= 1.0:
            continue
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd = (O.perlin3(P, p + e)[0] - O.perlin3(P, p - e)[0]) / (2 * h)
            assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(g[k]))


def _central(f, h):
    return (f(h) - f(-h)) / (2 * h)


def _richardson(f, h):
    """Central difference at step h, Richardson-extrapolated with h/2: O(h^4) truncation."""
    return (4 * _central(f, h / 2) - _central(f, h)) / 3


def _loss_with(p, coord, n, target, cases, amp=None, off=None):
    def f(d):
        q = dict(p, amplitudes=list(p["amplitudes"]))
        if amp is not None:
            q["amplitudes"][amp] += d
        if off is not None:
            o = [0.0, 0.0, 0.0]
            o[off] = d
            q["offset"] = tuple(o)
        return O.adjoint_chunk(q, coord, n, target, cases, grads=False)["loss"]
    return f


@pytest.mark.parametrize("coord", [(0, 0, -1), (3, 1, -1)])
def test_amplitude_and_offset_gradients_match_central_differences(coord):
    """Central differences with base step h = 1e-3 (Richardson-extrapolated with h/2) agree with the
    analytic gradient to 1e-4 relative; the plain central difference converges to it at second
    order (ill-conditioned edges make its h = 1e-3 truncation error alone exceed 1e-4)."""
    p = PR.preset("adjoint24", resolution=8)
    n = 8
    target = 0.25 * p["chunk_extent"]
    cases = _cases_fp64(p, coord, n)
    base = O.adjoint_chunk(p, coord, n, target, cases=cases)
    assert base["nverts"] > 0
    h = 1e-3
    checks = [(base["g_amp"][j], _loss_with(p, coord, n, target, cases, amp=j)) for j in range(p["octaves"])]
    checks += [(base["g_off"][i], _loss_with(p, coord, n, target, cases, off=i)) for i in range(3)]
    scale = max(np.abs(base["g_amp"]).max(), np.abs(base["g_off"]).max())
    for g, f in checks:
        assert abs(_richardson(f, h) - g) <= 1e-4 * max(abs(g), 1e-3 * scale), (g, _richardson(f, h))
        e1, e2 = abs(_central(f, h) - g), abs(_central(f, h / 10) - g)
        assert e2 <= max(0.02 * e1, 1e-7 * max(abs(g), 1e-3 * scale)), (g, e1, e2)


def test_density_gradient_matches_central_differences():
    p = PR.preset("adjoint24", resolution=8)
    n, coord = 8, (0, 0, -1)
    target = 0.25 * p["chunk_extent"]
    step = float(np.float32(p["chunk_extent"]) / np.float32(n))
    rho = O.density_fp64(p, coord, n)
    cases = _cases_fp64(p, coord, n)
    loss, g = O.plane_loss_dens(n, step, p["iso"], rho, cases, target, grads=True)
    base = O.adjoint_chunk(p, coord, n, target, cases=cases)
    assert np.allclose(g, base["g_dens"], rtol=1e-12, atol=0)
    nz = np.nonzero(g)[0]
    assert len(nz) > 10
    h = 1e-7
    for s in nz[:: max(1, len(nz) // 40)]:
        r1, r2 = rho.copy(), rho.copy()
        r1[s] += h
        r2[s] -= h
        fd = (O.plane_loss_dens(n, step, p["iso"], r1, cases, target) -
              O.plane_loss_dens(n, step, p["iso"], r2, cases, target)) / (2 * h)
        assert abs(fd - g[s]) <= 1e-4 * abs(g[s]) + 1e-6, (s, fd, g[s])


def test_amplitude_gradient_includes_normalisation():
    # scaling every amplitude by the same factor leaves fbm unchanged: sum_j A_j dL/dA_j = 0
    p = PR.preset("adjoint24", resolution=8)
    r = O.adjoint_chunk(p, (0, 0, -1), 8, 3.0)
    assert abs(np.dot(p["amplitudes"], r["g_amp"])) <= 1e-9 * np.abs(r["g_amp"]).sum()

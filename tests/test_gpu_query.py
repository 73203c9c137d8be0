"""Batched density point queries (uw_density_at) against the oracle's iso_at
(perlin_util.rs:24-29), as the boids call it (boid.rs:131-143,324-331)."""
import numpy as np
import pytest

from oracle import oracle as O
from underwaterworld_amd import params as PR

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("preset", ["reference", "chunks12", "volume512"])
def test_point_queries_match_iso_at(preset):
    import torch
    from underwaterworld_amd import Perlin, density_at
    p = PR.preset(preset)
    g = np.random.default_rng(3)
    pts = np.concatenate([g.uniform(-40, 40, size=(1500, 3)),
                          g.integers(-20, 20, size=(100, 3)).astype(np.float64),     # lattice points
                          g.uniform(-3, 3, size=(400, 3))])
    got = density_at(Perlin(p), pts)
    torch.cuda.synchronize()
    want = O.iso_at(p, pts)
    diff = np.abs(got.cpu().numpy().astype(np.float64) - want)
    # 1e-6 absolute, or one f32 ulp of the result where |rho| is large: far above or below the
    # world's z range (|z| > 2 noise units) the z bias alone makes |rho| reach 20, where one ulp
    # is 1.9e-6.
    tol = np.maximum(1e-6, np.spacing(np.abs(want).astype(np.float32)).astype(np.float64))
    inside = np.abs(pts[:, 2]) <= 2.0
    print(f"\n[{preset}] max |rho_gpu - iso_at| = {diff.max():.3g} (|z| <= 2: {diff[inside].max():.3g}), "
          f"max diff/tol = {(diff / tol).max():.3g}")
    assert (diff <= tol).all()
    assert diff[inside].max() <= 1e-6


def test_point_queries_agree_with_grid_samples():
    import torch
    from underwaterworld_amd import Perlin, density_at, sample_density
    p = PR.preset("chunks12")
    perlin = Perlin(p)
    n, c = 12, (3, 1, -1)
    grid, _ = sample_density(perlin, [c], n)
    ax = [O.coords(p, c[k], n) for k in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    pq = density_at(perlin, pts)
    torch.cuda.synchronize()
    assert float(np.max(np.abs(pq.cpu().numpy() - grid.cpu().numpy()[0]))) <= 1e-6

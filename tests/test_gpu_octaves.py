"""GPU parity at octave counts that have no exact-count kernel (the generic density loop reads the
count from the field), at an even and an odd number of samples per z-line: densities against the
oracle, and the batched path bit-identical to the plain path chunk by chunk."""
import numpy as np
import pytest

from oracle import oracle as O
from parity import DENS_TOL
from test_gpu_batched import _assert_same, _plain_by_resolution, _sync
from underwaterworld_amd import params as PR

pytestmark = pytest.mark.gpu

COORDS = [(0, 0, -1), (3, 1, -1), (5, 5, -2), (0, 0, 0), (-2, 4, -1), (7, -6, 1)]


@pytest.mark.parametrize("octaves", [1, 2, 5, 7, 16])
def test_generic_octave_counts(octaves):
    from underwaterworld_amd import Perlin, build_chunks_batched, sample_density
    p = PR.preset("chunks12", octaves=octaves)
    perlin = Perlin(p)
    worst = 0.0
    for n in (7, 12):  # 8 and 13 samples per line
        dens, _ = sample_density(perlin, COORDS, n)
        _sync()
        d = dens.cpu().numpy()
        for k, c in enumerate(COORDS):
            worst = max(worst, float(np.max(np.abs(d[k].astype(np.float64) - O.density_ref(p, c, n)))))
        ref = _plain_by_resolution(perlin, COORDS, [n] * len(COORDS))
        b = build_chunks_batched(perlin, COORDS, [n] * len(COORDS))
        _sync()
        _assert_same(b, ref, len(COORDS))
    print(f"\n[{octaves} octaves] max|drho| = {worst:.3g}")
    assert worst <= DENS_TOL


@pytest.mark.parametrize("octaves", [3, 4, 6, 8])
def test_exact_octave_count_kernels(octaves):
    # the octave counts with exact-count specialisations of the density loops (the reference's 3,
    # the workloads' 4, 6 and 8): densities against the oracle, batched equal to plain
    from underwaterworld_amd import Perlin, build_chunks_batched, sample_density
    p = PR.preset("chunks12", octaves=octaves)
    perlin = Perlin(p, apply_tuning=False)  # the library's default shape accepts up to 32^3
    worst = 0.0
    for n in (7, 12, 24):
        dens, _ = sample_density(perlin, COORDS, n)
        _sync()
        d = dens.cpu().numpy()
        for k, c in enumerate(COORDS):
            worst = max(worst, float(np.max(np.abs(d[k].astype(np.float64) - O.density_ref(p, c, n)))))
        ref = _plain_by_resolution(perlin, COORDS, [n] * len(COORDS))
        b = build_chunks_batched(perlin, COORDS, [n] * len(COORDS))
        _sync()
        _assert_same(b, ref, len(COORDS))
    print(f"\n[{octaves} octaves, exact-count kernels] max|drho| = {worst:.3g}")
    assert worst <= DENS_TOL

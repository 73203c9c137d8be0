"""World streaming with the batched generation step on the GPU (uw_world_*), stepped along a sub
path beside the oracle's restatement of world.rs (oracle/world_ref.py) whose chunk builds come from
the oracle mesher: same render lists, counts and eviction; stored meshes equal the batched path."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import world_ref as R
from underwaterworld_amd import params as PR
from worldcam import frustum

pytestmark = pytest.mark.gpu


def _path(frames):
    pos = np.float32([5.0, 7.0, -3.0])
    for t in range(frames):
        ang = 0.25 * t
        bearing = np.float32([np.cos(ang), np.sin(ang), 0.05])
        yield pos.copy(), bearing
        pos = (pos + np.float32(9.0) * bearing).astype(np.float32)


@pytest.mark.parametrize("max_batch", [0, 60])
def test_world_matches_reference_planner(max_batch):
    from underwaterworld_amd import Perlin
    from underwaterworld_amd.world import World
    p = PR.preset("reference")
    perlin = Perlin(p)
    world = World(perlin, max_batch=max_batch)
    ref = R.WorldRef(max_batch=max_batch)
    cache = {}

    def not_blank_of(batch):
        out = []
        for pos in batch:
            if pos not in cache:
                cache[pos] = len(O.build_chunk(p, pos, 12)["inds"]) > 0
            out.append(cache[pos])
        return out

    total_built = 0
    for i, (pos, bearing) in enumerate(_path(14)):
        view, gen = frustum(pos, bearing, 55.0), frustum(pos, bearing, 90.0)
        built = world.update(pos, bearing, view, gen, sub_reset=(i == 0))
        batch = ref.update(pos, bearing, view.T.tolist(), gen.T.tolist(), i == 0, not_blank_of)
        assert built == len(batch), i
        total_built += built
        assert world.chunks_to_render() == ref.render, i
        assert world.generate_count() == len(ref.to_generate), i
        assert world.total_count() == len(ref.chunks), i
    assert total_built > 300
    print(f"\n[max_batch={max_batch}] built {total_built} chunks, {world.total_count()} stored, "
          f"{world.render_count()} to render")


def test_world_chunks_equal_batched_build():
    import torch
    from underwaterworld_amd import Perlin, build_chunks_batched
    from underwaterworld_amd.world import World
    p = PR.preset("reference")
    perlin = Perlin(p)
    world = World(perlin, vertex_format=1, index_format=1, collision=1)
    pos, bearing = np.float32([5.0, 7.0, -3.0]), np.float32([1.0, 0.0, 0.0])
    world.update(pos, bearing, frustum(pos, bearing, 55.0), frustum(pos, bearing, 90.0), sub_reset=True)
    render = world.chunks_to_render()
    assert len(render) > 10
    sample = render[:24]
    b = build_chunks_batched(perlin, sample, 12, colors=True, u16=True, collision=True)
    torch.cuda.synchronize()
    info = b.info()
    for i, c in enumerate(sample):
        wc = world.get_chunk(c)
        assert wc is not None and wc.not_blank()
        v, ix = b.chunk_mesh(i, info)
        cells, nrm = b.chunk_collision(i, info)
        assert np.array_equal(wc.verts, v) and np.array_equal(wc.inds, ix), c
        assert np.array_equal(wc.tri_cells, cells) and np.array_equal(wc.normals, nrm), c
        tv, tn = wc.tris_around((0.5, 0.5, 0.5), 3)
        ids = O.tris_around(cells, 12, (0.5, 0.5, 0.5), 3)
        assert len(tv) == len(ids)
    assert world.get_chunk((10 ** 6, 0, 0)) is None

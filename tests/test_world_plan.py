"""World streaming planner (World::update_nearby + GenPrio, world.rs:29-60,149-228): the engine's
host-side uw_world_plan against the oracle's pure-Python restatement (oracle/world_ref.py), on
camera paths through the world with random sets of already-built chunks. No GPU needed."""
import numpy as np
import pytest

from oracle import world_ref as R
from worldcam import column_major, frustum

from underwaterworld_amd import world as W


def _scenarios(count, seed):
    g = np.random.default_rng(seed)
    for _ in range(count):
        pos = np.array([g.uniform(-300, 300), g.uniform(-300, 300), g.uniform(-45, 45)], np.float32)
        ang = g.uniform(0, 2 * np.pi)
        bearing = np.array([np.cos(ang), np.sin(ang), g.uniform(-0.5, 0.5)], np.float32)
        yield pos, bearing, g


def _known(pos, g, frac):
    sc = R.sub_chunk(pos)
    known = {}
    for x in range(-5, 5):
        for y in range(-5, 5):
            for z in range(-2, 3):
                if g.uniform() < frac:
                    known[(sc[0] + x, sc[1] + y, z)] = bool(g.uniform() < 0.6)
    return known


@pytest.mark.parametrize("frac", [0.0, 0.3, 0.9])
def test_plan_matches_reference_restatement(frac):
    cfg = W.reference_config()
    total_cands = total_render = 0
    for pos, bearing, g in _scenarios(25, 17 + int(frac * 10)):
        view = frustum(pos, bearing, 55.0)
        gen = frustum(pos, bearing, 90.0)
        known = _known(pos, g, frac)
        r_ref, c_ref = R.update_nearby(pos, view.T.tolist(), gen.T.tolist(), known)
        r, c = W.plan(cfg, pos, view, gen, known)
        assert r == r_ref
        assert [x[0] for x in c] == [x[0] for x in c_ref]
        for got, (p, (d, z, iv, ig)) in zip(c, c_ref):
            assert got[1] == d and got[2] == z and got[3] == iv and got[4] == ig, p
        total_cands += len(c)
        total_render += len(r)
    assert total_cands > 100
    if frac > 0.5:
        assert total_render > 0


def test_plan_priority_rules():
    cfg = W.reference_config()
    pos, bearing = np.float32([8.0, 8.0, 0.0]), np.float32([1.0, 0.0, 0.0])
    view, gen = frustum(pos, bearing, 55.0), frustum(pos, bearing, 90.0)
    _, c = W.plan(cfg, pos, view, gen, {})
    rank = [(0 if iv else 2) + (0 if ig else 1) for _, _, _, iv, ig in c]
    assert rank == sorted(rank) and rank[0] == 0 and rank[-1] == 3
    # within a rank, dist^2 + z ascends
    for r in set(rank):
        keys = [np.float32(d * d + z) for (_, d, z, _, _), k in zip(c, rank) if k == r]
        assert keys == sorted(keys)
    # z range of the neighbourhood is clamped to MIN_Z..MAX_Z and the radius to 5 extents
    assert all(-2 <= p[2] <= 2 for p, *_ in c)
    assert all(d <= 80.0 for _, d, *_ in c)


def test_plan_overflow_and_bad_config():
    import ctypes
    from underwaterworld_amd import _lib
    cfg = W.reference_config()
    pos = np.float32([0, 0, 0])
    m = column_major(np.eye(4))
    nr, nc = ctypes.c_int32(0), ctypes.c_int32(0)
    cands = (W.UwWorldCandidate * 3)()
    rc = _lib.lib().uw_world_plan(ctypes.byref(cfg), pos.ctypes.data, m.ctypes.data, m.ctypes.data, None, None, 0,
                                  None, 0, ctypes.byref(nr), cands, 3, ctypes.byref(nc))
    assert rc == _lib.UW_ERR_OVERFLOW and nc.value > 3
    bad = W.reference_config(resolution=0)
    rc = _lib.lib().uw_world_plan(ctypes.byref(bad), pos.ctypes.data, m.ctypes.data, m.ctypes.data, None, None, 0,
                                  None, 0, ctypes.byref(nr), cands, 3, ctypes.byref(nc))
    assert rc == _lib.UW_ERR_INVALID_ARG

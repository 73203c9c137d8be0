"""Host logic of the launch-shape autotuner: budget-to-slab mapping and the per-resolution slab
refinement (its timing call replaced by a synthetic cost, so no device is needed)."""
from underwaterworld_amd import autotune as AT


def test_slabs_for_budget_fits_the_budget():
    for budget in AT.BUDGETS:
        s = AT.slabs_for_budget(budget)
        assert len(s) == 33 and s[0] == 1
        for n in range(1, 33):
            L = n + 1
            assert 1 <= s[n] <= n
            if s[n] == n:
                assert 4 * L ** 3 <= budget or 4 * L * L * (n + 1) <= budget
            elif s[n] > 1:  # planes of the slab plus the previous plane fit the budget
                assert 4 * L * L * (s[n] + 1) <= budget


def _coupled_cost(t):
    # optimum at (slab24, slab32) = (8, 6), reachable from (10, 5) only through (10, 6): while
    # slab32 != 6 the best slab24 is 10, so the first pass (24^3 first) can move only slab32 and
    # a second pass is needed to move slab24 (as when one size's slab decides the cell ring, and so
    # the occupancy, that the other's best value depends on)
    a, b = t["slab"][24], t["slab"][32]
    a_cost = 0.01 * abs(a - 8) if b == 6 else 0.01 * abs(a - 10)
    return 1.0 + a_cost + 0.02 * abs(b - 6) + (0.005 if b != 6 else 0.0)


def test_refinement_repeats_until_no_pass_improves(monkeypatch):
    calls = []

    def fake_time(perlin, coords, res, t, steps, stream, ref, log, label):
        calls.append(tuple(t["slab"]))
        return _coupled_cost(t)

    monkeypatch.setattr(AT, "_time_shape", fake_time)
    start = {"block": 256, "grid": 0, "max_res": 32, "slab": AT.slabs_for_budget(28 * 1024)}
    assert (start["slab"][24], start["slab"][32]) == (10, 5)
    res = [8, 12, 24, 32]
    ms, best, results = AT.refine_slabs(None, None, res, _coupled_cost(start), start, 1, None, [None],
                                        log=lambda *a: None)
    assert (best["slab"][24], best["slab"][32]) == (8, 6)
    assert ms == _coupled_cost(best)
    # one-slab resolutions (8, 12 fit whole) are never varied, and no shape is timed twice
    assert all(c[8] == start["slab"][8] and c[12] == start["slab"][12] for c in calls)
    assert len(calls) == len(set(calls))


def test_refinement_single_pass_when_nothing_improves(monkeypatch):
    monkeypatch.setattr(AT, "_time_shape", lambda *a, **k: 2.0)
    start = {"block": 256, "grid": 0, "max_res": 32, "slab": AT.slabs_for_budget(28 * 1024)}
    ms, best, results = AT.refine_slabs(None, None, [24, 32], 1.0, start, 1, None, [None], log=lambda *a: None)
    assert best is start and ms == 1.0 and len(results) > 0

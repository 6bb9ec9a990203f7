"""Host logic of the autotuner (no GPU): the fastest candidate wins, the
search spaces are well formed, and the cache returns a stored choice without
measuring."""
import pytest

from entity_amd.autotune import Candidate, candidates_for, signature, tune


def test_fastest_candidate_wins():
    cands = candidates_for(2)
    cost = {c: 10.0 + i for i, c in enumerate(cands)}
    fast = cands[5]
    cost[fast] = 1.0
    res = tune(lambda c: cost[c], cands)
    assert res.best == fast
    assert res.table()[0]["ms_per_step"] == 1.0
    assert len(res.ms_per_step) == len(cands)


@pytest.mark.parametrize("dim", [2, 3])
def test_search_space(dim):
    cands = candidates_for(dim)
    assert len(set(cands)) == len(cands)
    assert all(len(c.tile) == dim and c.particles_per_thread >= 1 for c in cands)
    assert any(c.sort_interval == 0 for c in cands) and any(c.sort_interval > 0 for c in cands)
    with pytest.raises(ValueError):
        candidates_for(1)


def test_candidates_fit_the_fused_window():
    # the built library's window (32 x 16 in 2D, 14 x 10 x 10 in 3D) keeps
    # 8x8 and 16x8, drops 8x16 and 16x16 (16 + 4 rows > 16)
    from entity_amd.autotune import fused_window
    assert fused_window(2) == (32, 16) and fused_window(3) == (14, 10, 10)
    tiles2 = {c.tile for c in candidates_for(2) if c.sort_interval}
    assert tiles2 == {(8, 8), (16, 8)}
    tiles3 = {c.tile for c in candidates_for(3) if c.sort_interval}
    assert tiles3 == {(4, 4, 4), (8, 4, 4), (4, 4, 2)}
    # a larger window admits the tall tiles
    assert (16, 16) in {c.tile for c in candidates_for(2, window=(32, 32))}


def test_cache_round_trip(tmp_path):
    from entity_amd.engine import SRParams
    key = signature(SRParams(dim=2, n=(64, 64)), 16, [2])
    path = tmp_path / "tune.json"
    calls = []
    cands = [Candidate((8, 8), 4), Candidate((16, 8), 2)]
    res = tune(lambda c: calls.append(c) or (1.0 if c.tile == (16, 8) else 2.0), cands,
               cache=path, key=key)
    assert res.best == Candidate((16, 8), 2) and len(calls) == 2
    again = tune(lambda c: calls.append(c) or 0.0, cands, cache=path, key=key)
    assert again.best == res.best and len(calls) == 2     # served from the cache
    with pytest.raises(ValueError):
        tune(lambda c: 0.0, [])

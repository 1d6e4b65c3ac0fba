"""Host logic of the launch-shape autotuner (no GPU: injected knobs and timer)."""
import json

from emsim_amd import autotune


def _fake(costs):
    state = {}

    def set_knobs(cfg):
        state.clear()
        state.update(cfg)

    def timer():
        return costs(state)
    return set_knobs, timer, state


def test_coordinate_descent_finds_separable_optimum():
    # cost minimised at row=4, col=2, shape=2 (separable)
    costs = lambda s: ((s["row_items_per_block"] - 4) ** 2 + (s["col_tiles_per_block"] - 2) ** 2
                       + (s["col_shape"] - 2) ** 2 + 1)
    set_knobs, timer, state = _fake(costs)
    best, table = autotune.tune(lambda: None, sync=lambda: None, set_knobs=set_knobs, timer=timer)
    assert best == {"row_items_per_block": 4, "col_tiles_per_block": 2, "col_shape": 2}
    assert state == best                       # the winner is left applied
    assert len(table) == sum(len(v) for v in autotune.KNOBS.values())


def test_default_rule_kept_when_nothing_is_faster():
    set_knobs, timer, _ = _fake(lambda s: 1.0 + sum(v > 0 for v in s.values()))
    best, _ = autotune.tune(lambda: None, sync=lambda: None, set_knobs=set_knobs, timer=timer)
    assert best == {"row_items_per_block": 0, "col_tiles_per_block": 0, "col_shape": 0}


def test_cache_roundtrip(tmp_path):
    calls = []
    set_knobs, timer, state = _fake(lambda s: calls.append(1) or
                                    abs(s["col_tiles_per_block"] - 8) + 1)
    path = tmp_path / "tuning.json"
    cfg = autotune.cached("dev/n512/b64", lambda: None, path=path, sync=lambda: None,
                          set_knobs=set_knobs, timer=timer)
    assert cfg["col_tiles_per_block"] == 8
    n_timed = len(calls)
    assert json.loads(path.read_text())["dev/n512/b64"] == cfg
    cfg2 = autotune.cached("dev/n512/b64", lambda: None, path=path, sync=lambda: None,
                           set_knobs=set_knobs, timer=timer)
    assert cfg2 == cfg and len(calls) == n_timed   # served from the cache, applied
    assert state == cfg

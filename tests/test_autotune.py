"""Host logic of the autotuner: the stored entry of the closest graph size is
the one applied."""
import json

from byo_amd import autotune


def test_lookup_closest_size(tmp_path):
    db = {"entries": [
        {"graph": "a", "n": 1 << 16, "m": 1 << 21, "knobs": {"pieces": 4}},
        {"graph": "b", "n": 1 << 22, "m": 1 << 27, "knobs": {"pieces": 16, "pr_lanes": 8}},
    ]}
    p = tmp_path / "t.json"
    p.write_text(json.dumps(db))
    assert autotune.lookup(1 << 22, 130_000_000, str(p)) == {"pieces": 16, "pr_lanes": 8}
    assert autotune.lookup(1 << 15, 1 << 20, str(p)) == {"pieces": 4}
    assert autotune.lookup(10, 0, str(p)) == {}
    assert autotune.lookup(1 << 20, 1 << 24, str(tmp_path / "missing.json")) == {}


def test_candidates_are_valid_knob_values():
    # the values the tuner tries lie in the ranges byo_set_tuning accepts
    ranges = {"coop_blocks_per_cu": (1, 4), "ppr_blocks_per_cu": (1, 8), "pieces": (1, 64), "ppr_pieces": (1, 16),
              "pr_lanes": (0, 32), "pr_grid": (256, 65536), "pr_relabel": (0, 1), "pr_xcd": (0, 1),
              "pull_dyn": (0, 64), "coop_grid_sqrt_div": (0, 1 << 20), "pr_xcd_vw": (0, 1024), "pr_pb": (0, 1)}
    for k, vals in autotune.CANDIDATES.items():
        lo, hi = ranges[k]
        assert all(lo <= v <= hi for v in vals), k
        assert autotune.WORKLOAD[k] in ("bfs", "ppr", "pr")

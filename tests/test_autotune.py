"""Host logic of the launch-shape autotuner: the coordinate descent finds
the optimum of a separable score, stays put on a flat one, and the tuning
table round-trips by workload signature."""
import os

from quicksilver_amd import autotune


def test_descent_finds_a_separable_optimum():
    target = {"waves_per_simd": 3, "waves_per_cu": 24, "refill_min": 4, "chunk": 32, "block_size": 128,
              "stack_entries": 2048, "sort_interval": 2}

    def score(s):
        return 100.0 - sum(abs(autotune.SPACE[k].index(s[k]) - autotune.SPACE[k].index(v))
                           for k, v in target.items())

    best, sc = autotune.coordinate_descent(score, dict(autotune.DEFAULT), rounds=3, log=lambda m: None)
    assert {k: best[k] for k in target} == target
    assert sc == 100.0


def test_flat_score_keeps_the_start():
    best, _ = autotune.coordinate_descent(lambda s: 1.0, dict(autotune.DEFAULT), log=lambda m: None)
    assert best == autotune.DEFAULT


def test_table_round_trip(tmp_path):
    path = os.path.join(tmp_path, "t.json")
    deck = "Simulation:\n   nx: 40\n   nParticles: 1000000\n"
    sig = autotune.signature("plain", [deck])
    assert sig == "plain:n40:p1000000"
    assert autotune.lookup(sig, path) is None
    shape = dict(autotune.DEFAULT, chunk=128)
    autotune.store(sig, shape, 1.5e9, path)
    assert autotune.lookup(sig, path) == shape
    assert autotune.signature("batch", [deck, deck]) == "batch:2:p2000000"

"""Batched form: several problems of uneven size and step count advance
together in shared launches; every problem must reproduce the reference
CPU path's tallies bit-exactly (flux to 1e-10), i.e. the same result alone
and in the mix.  The default form replays each step from captured HIP
graphs, so every cycle after the first checks a graph replay (with the
library's own eager launches and host synchronisations in between)."""
import pytest

import quicksilver_amd as q
from conftest import load_golden

MIX = ["tiny_cube", "small_cube_short", "nested_small", "spheres", "scatter_only", "no_fission",
       "nonflat", "all_escape", "no_collisions"]


def _check(name, g, got):
    assert len(got) == len(g["cycles"])
    for c, (ref, (bal, flux)) in enumerate(zip(g["cycles"], got)):
        for f in q.BALANCE_FIELDS:
            assert bal[f] == ref[f], f"{name} cycle {c} {f}: batch {bal[f]} reference {ref[f]}"
        rf = ref["flux_sum"]
        assert abs(flux - rf) <= 1e-10 * abs(rf), (name, c, flux, rf)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [{}, {"graphs": 0}, {"waves_per_simd": 2},
                                   {"waves_per_simd": 4, "stack_entries": 0},
                                   {"block_size": 64, "waves_per_cu": 3, "chunk": 8},
                                   {"sort_interval": 1}, {"sort_interval": 2, "graphs": 0}],
                         ids=["default_graphs", "eager_launches", "registers_2_waves",
                              "registers_4_waves_passes_only", "few_waves", "sorted_graphs",
                              "sorted_every_2nd_eager"])
def test_mixed_batch_matches_reference_per_problem(shape):
    print(f"[batch] mixed fixture decks, tuning {shape}", flush=True)
    golds = [load_golden(n) for n in MIX]
    probs = [q.Problem.from_file(g["deck_path"]) for g in golds]
    try:
        with q.Batch(probs) as b:
            if shape:
                b.set_tuning(**shape)
            out = b.run([len(g["cycles"]) for g in golds])
        for name, g, got in zip(MIX, golds, out):
            _check(name, g, got)
    finally:
        for p in probs:
            p.close()


@pytest.mark.gpu
def test_batch_then_alone_continues_identically():
    """A problem advanced two cycles in a batch and then alone gives the same
    tallies as the reference run."""
    print("[batch] handoff batch -> alone", flush=True)
    g = load_golden("small_cube_short")
    other = load_golden("nested_small")
    p = q.Problem.from_file(g["deck_path"])
    o = q.Problem.from_file(other["deck_path"])
    try:
        with q.Batch([p, o]) as b:
            first = [b.step()[0] for _ in range(2)]
        rest = [(bal, flux) for bal, flux, _ in p.run(len(g["cycles"]) - 2)]
        _check("small_cube_short", g, first + rest)
    finally:
        p.close()
        o.close()

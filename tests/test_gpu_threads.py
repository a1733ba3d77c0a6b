"""Problems built and released from several host threads at once.

Independent handles may be created concurrently (the workloads build the
mixed_batch problems on 8 threads), and released handles hand their
stream and pinned counter blocks to the next problem.  Each problem built
this way must still reproduce the reference CPU path's tallies
bit-exactly, run alone or in a batch.
"""
import pytest

import quicksilver_amd as q
from conftest import load_golden
from quicksilver_amd import workloads

DECKS = ["tiny_cube", "small_cube_short", "nested_small", "spheres", "scatter_only", "no_fission",
         "nonflat", "all_escape", "no_collisions", "all_absorb"]


def _check(name, g, got):
    assert len(got) == len(g["cycles"])
    for c, (ref, (bal, flux)) in enumerate(zip(g["cycles"], got)):
        for f in q.BALANCE_FIELDS:
            assert bal[f] == ref[f], f"{name} cycle {c} {f}: got {bal[f]} reference {ref[f]}"
        rf = ref["flux_sum"]
        assert abs(flux - rf) <= 1e-10 * abs(rf), (name, c, flux, rf)


@pytest.mark.gpu
def test_problems_built_on_threads_match_reference():
    print("[threads] 2 x 10 fixture decks built on 8 threads, run alone", flush=True)
    golds = [load_golden(n) for n in DECKS] * 2
    texts = [open(g["deck_path"]).read() for g in golds]
    probs = workloads.build_problems(texts, threads=8)
    try:
        for g, p in zip(golds, probs):
            out = p.run()
            _check(g["deck"], g, [(b, f) for b, f, _ in out])
            p.close()
    finally:
        for p in probs:
            p.close()


@pytest.mark.gpu
def test_reused_streams_and_counters_after_release():
    print("[threads] build, release, build again in a batch", flush=True)
    golds = [load_golden(n) for n in DECKS]
    texts = [open(g["deck_path"]).read() for g in golds]
    for p in workloads.build_problems(texts, threads=8):
        p.close()
    probs = workloads.build_problems(texts, threads=8)
    try:
        with q.Batch(probs) as b:
            out = b.run([len(g["cycles"]) for g in golds])
        for g, per in zip(golds, out):
            _check(g["deck"], g, per)
    finally:
        for p in probs:
            p.close()

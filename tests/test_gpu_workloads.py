"""The batched and served workloads at their full size against the reference.

tests/golden/mixed_batch.json holds, for each of the 64 decks of
``decks.mixed_batch()`` (8-39 cells per edge, 1.3e4-1.5e6 particles, 1-3
materials x 5-20 isotopes, 10 steps), the per-cycle tallies of the reference
CPU path run on that deck alone (oracle/_ref/qs_ref, main.cc:55-66: one
problem per run); tests/golden/serve_stream.json does the same for the first
32 requests of ``decks.serve_stream()`` (5-20 steps each).  Each deck's text
digest is stored beside its tallies, so a change in the generator shows as a
digest mismatch rather than as a parity failure.

Bar: every problem's 13 balance tallies bit-exact in every cycle, the scalar
flux sum within 1e-10 relative -- in the batch under graph replay, one at a
time, and served with staggered arrivals (each result the same alone and in
the mix).
"""
import hashlib

import pytest

import quicksilver_amd as q
from conftest import load_workload_set
from quicksilver_amd import decks, workloads

FLUX_RTOL = 1e-10


def _digest(text):
    return hashlib.sha256(text.encode()).hexdigest()[:16]


def _decks(name):
    fx = load_workload_set(name)
    if name == "mixed_batch":
        ds = decks.mixed_batch()
    else:
        ds = [r.deck for r in decks.serve_stream(count=len(fx["problems"]))]
    assert len(ds) == len(fx["problems"])
    for d, p in zip(ds, fx["problems"]):
        assert _digest(d.text()) == p["digest"], f"{name} deck {p['index']}: generator output changed"
    return ds, fx


def _check(label, ref_cycles, got):
    assert len(got) == len(ref_cycles), label
    for c, (ref, (bal, flux)) in enumerate(zip(ref_cycles, got)):
        for f in q.BALANCE_FIELDS:
            assert bal[f] == ref[f], f"{label} cycle {c} {f}: device {bal[f]} reference {ref[f]}"
        rf = ref["flux_sum"]
        assert abs(flux - rf) <= FLUX_RTOL * abs(rf), f"{label} cycle {c} flux {flux!r} vs {rf!r}"


def test_workload_fixtures_cover_the_generated_decks():
    """CPU check: the fixtures exist for every generated deck, with the
    deck's current digest, and pass the balance identities."""
    from quicksilver_amd.checks import balance_violations
    for name, count in (("mixed_batch", 64), ("serve_stream", 32)):
        ds, fx = _decks(name)
        assert len(ds) == count
        for d, p in zip(ds, fx["problems"]):
            assert len(p["cycles"]) == int(d.simulation["nSteps"])
            assert p["cycles"][0]["start"] == 0
            assert not balance_violations(p["cycles"]), (name, p["index"])


@pytest.mark.gpu
def test_mixed_batch_batched_matches_reference():
    print("[workloads] mixed_batch: 64 problems batched, graph replay", flush=True)
    ds, fx = _decks("mixed_batch")
    res = workloads.run_batched(ds)
    for i, p in enumerate(fx["problems"]):
        _check(f"mixed_batch[{i}] batched", p["cycles"], list(zip(res["balances"][i], res["fluxes"][i])))


@pytest.mark.gpu
def test_mixed_batch_one_at_a_time_matches_reference():
    print("[workloads] mixed_batch: 64 problems one at a time", flush=True)
    ds, fx = _decks("mixed_batch")
    res = workloads.run_one_at_a_time(ds)
    for i, p in enumerate(fx["problems"]):
        _check(f"mixed_batch[{i}] alone", p["cycles"], list(zip(res["balances"][i], res["fluxes"][i])))


@pytest.mark.gpu
def test_serve_stream_prefix_matches_reference():
    """The first 32 requests of the stream, submitted two per server step
    (so requests join while others run and leave as they finish)."""
    print("[workloads] serve_stream: 32 requests, staggered submits", flush=True)
    ds, fx = _decks("serve_stream")
    with q.Server(200 << 30) as s:
        ids = []
        nxt = 0
        for _ in range(1000):
            for _ in range(2):
                if nxt < len(ds):
                    ids.append(s.submit(ds[nxt].text()))
                    nxt += 1
            active, queued = s.step()
            if nxt == len(ds) and active == 0 and queued == 0:
                break
        assert len(ids) == len(ds)
        for i, rid in enumerate(ids):
            st, n = s.status(rid)
            assert st == q.Server.DONE, (i, st)
            _check(f"serve_stream[{i}]", fx["problems"][i]["cycles"], s.result(rid))

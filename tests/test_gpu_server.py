"""Serving loop: requests arriving while others run are admitted by HBM
budget and batched with them; every request's tallies equal the reference
CPU path's for its deck (bit-exact balance, flux to 1e-10) -- the same
result alone and in the mix."""
import pytest

import quicksilver_amd as q
from conftest import load_golden

ARRIVALS = [  # (step at which the request arrives, deck)
    (0, "nested_small"), (0, "tiny_cube"), (1, "spheres"), (2, "small_cube_short"),
    (2, "no_fission"), (5, "scatter_only"), (6, "nonflat")]


def _check(name, g, got):
    assert len(got) == len(g["cycles"]), name
    for c, (ref, (bal, flux)) in enumerate(zip(g["cycles"], got)):
        for f in q.BALANCE_FIELDS:
            assert bal[f] == ref[f], f"{name} cycle {c} {f}: served {bal[f]} reference {ref[f]}"
        assert abs(flux - ref["flux_sum"]) <= 1e-10 * abs(ref["flux_sum"]), (name, c)


def _serve(budget, max_active=0):
    golds = {n: load_golden(n) for _, n in ARRIVALS}
    with q.Server(budget, max_active) as s:
        ids = {}
        step = 0
        while True:
            for at, name in ARRIVALS:
                if at == step:
                    ids[name] = s.submit(open(golds[name]["deck_path"]).read())
            active, queued = s.step()
            step += 1
            if step > max(a for a, _ in ARRIVALS) and active == 0 and queued == 0:
                break
            assert step < 200
        for name, rid in ids.items():
            st, n = s.status(rid)
            assert st == q.Server.DONE, (name, st)
            _check(name, golds[name], s.result(rid))
        stats = s.stats()
    assert stats["completed"] == len(ARRIVALS)
    return stats


@pytest.mark.gpu
def test_served_requests_match_reference():
    print("[server] staggered arrivals, ample budget", flush=True)
    _serve(64 << 30)


@pytest.mark.gpu
def test_served_requests_match_reference_one_at_a_time():
    print("[server] staggered arrivals, max_active=1", flush=True)
    _serve(64 << 30, max_active=1)


@pytest.mark.gpu
def test_request_larger_than_budget_is_rejected():
    print("[server] over-budget request", flush=True)
    g = load_golden("tiny_cube")
    with q.Server(1 << 20) as s:
        rid = s.submit(open(g["deck_path"]).read())
        s.step()
        assert s.status(rid)[0] == q.Server.FAILED
        assert s.stats()["rejected"] == 1

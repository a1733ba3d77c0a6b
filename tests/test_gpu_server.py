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


@pytest.mark.gpu
def test_batch_slots_grow_and_are_reused():
    """More simultaneous requests than the first batch has slots (16): the
    batch grows while running requests keep their state; later arrivals
    take slots freed by finished requests.  Every result still equals the
    reference's."""
    print("[server] slot growth and reuse", flush=True)
    tiny, short = load_golden("tiny_cube"), load_golden("small_cube_short")
    texts = {n: open(g["deck_path"]).read() for n, g in (("tiny_cube", tiny), ("small_cube_short", short))}
    with q.Server(64 << 30) as s:
        ids = [("tiny_cube", s.submit(texts["tiny_cube"])) for _ in range(20)]
        s.step()
        ids += [("small_cube_short", s.submit(texts["small_cube_short"])) for _ in range(3)]
        for _ in range(3):
            s.step()
        ids += [("tiny_cube", s.submit(texts["tiny_cube"])) for _ in range(5)]
        for _ in range(50):
            active, queued = s.step()
            if active == 0 and queued == 0:
                break
        for name, rid in ids:
            assert s.status(rid)[0] == q.Server.DONE, name
            _check(name, tiny if name == "tiny_cube" else short, s.result(rid))
        assert s.stats()["completed"] == len(ids)

"""Failure handling on the device path: a problem whose buffers are sized too
small fails alone, with QSG_ERR_CAPACITY, and is refused any further cycle;
every other problem of the same batch or server keeps matching the reference
CPU path's tallies for its deck.  The reference itself only prints an
assertion and continues (qs_assert.hh:9-26), so the reports are this build's
own; the overflow is forced from the host (a small per-problem particle
capacity), and every device store past a capacity is skipped by a bounds
check in the kernels, so each kernel exits normally.  The numeric error
(QSG_ERR_NUMERIC) comes from a deck whose fission multiplicity can exceed
the reference's MAX_PRODUCTION_SIZE of 4 (CollisionEvent.cc:15, 100-104):
the kernel keeps 4 products, flags the cycle and finishes it."""
import pytest

import quicksilver_amd as q
from conftest import load_golden

CAPACITY, STATE, NUMERIC = 3, 4, 5


def _check(name, g, got, n=None):
    cycles = g["cycles"] if n is None else g["cycles"][:n]
    assert len(got) == len(cycles), name
    for c, (ref, (bal, flux)) in enumerate(zip(cycles, got)):
        for f in q.BALANCE_FIELDS:
            assert bal[f] == ref[f], f"{name} cycle {c} {f}: device {bal[f]} reference {ref[f]}"
        assert abs(flux - ref["flux_sum"]) <= 1e-10 * abs(ref["flux_sum"]), (name, c)


def _small(g):
    """A particle capacity below the deck's population target: the first
    cycle's population control (source split up to the target) overflows."""
    text = open(g["deck_path"]).read()
    with q.Problem(text) as p:
        inf = p.info()
    return text, int(inf.n_particles_target // 2), int(inf.source_per_cycle)


@pytest.mark.gpu
def test_plain_run_capacity_overflow_fails_problem():
    print("[failures] plain qsg_run with a too-small capacity", flush=True)
    g = load_golden("small_cube_short")
    text, cap, src = _small(g)
    assert src < cap
    with q.Problem(text, particle_capacity=cap) as p:
        assert p.info().particle_capacity == cap
        with pytest.raises(q.QsgError) as e:
            p.run()
        assert e.value.code == CAPACITY
        code, msg = p.error()
        assert code == CAPACITY and "population-control buffer overflow" in msg
        with pytest.raises(q.QsgError) as e2:
            p.run(1)
        assert e2.value.code == STATE
        with pytest.raises(q.QsgError) as e3:
            p.cycle_init()
        assert e3.value.code == STATE
    # a problem that does not fit at construction-time checks stays usable
    with q.Problem(text) as p:
        assert p.error()[0] == 0
        _check("small_cube_short", g, [(b, f) for b, f, _ in p.run()])


@pytest.mark.gpu
def test_single_step_capacity_overflow_fails_problem():
    print("[failures] three-phase calls with a too-small capacity", flush=True)
    g = load_golden("small_cube_short")
    text, cap, _ = _small(g)
    with q.Problem(text, particle_capacity=cap) as p:
        with pytest.raises(q.QsgError) as e:
            p.cycle_init()
        assert e.value.code == CAPACITY
        assert p.error()[0] == CAPACITY
        with pytest.raises(q.QsgError) as e2:
            p.cycle_tracking()
        assert e2.value.code == STATE


@pytest.mark.gpu
@pytest.mark.parametrize("graphs", [1, 0], ids=["graphs", "eager"])
def test_batch_isolates_failed_problem(graphs):
    print(f"[failures] batch with one undersized problem, graphs={graphs}", flush=True)
    good = [load_golden(n) for n in ("nested_small", "tiny_cube", "spheres")]
    bad = load_golden("small_cube_short")
    text, cap, _ = _small(bad)
    probs = [q.Problem.from_file(g["deck_path"]) for g in good[:2]]
    probs.insert(1, q.Problem(text, particle_capacity=cap))
    probs.append(q.Problem.from_file(good[2]["deck_path"]))
    names = ["nested_small", "bad", "tiny_cube", "spheres"]
    golds = [good[0], None, good[1], good[2]]
    try:
        with q.Batch(probs) as b:
            b.set_tuning(graphs=graphs)
            steps = [len(g["cycles"]) if g else 0 for g in golds]
            out = [[] for _ in probs]
            first = b.step(strict=False)
            assert first[1] is None
            assert b.problem_error(1) == CAPACITY
            assert [b.problem_error(i) for i in (0, 2, 3)] == [0, 0, 0]
            for i in (0, 2, 3):
                out[i].append(first[i])
            with pytest.raises(q.QsgError) as e:
                b.set_active(1, True)
            assert e.value.code == STATE
            for k in range(1, max(steps)):
                for i, s in enumerate(steps):
                    if k == s and i != 1:
                        b.set_active(i, False)
                res = b.step()
                for i, s in enumerate(steps):
                    if i != 1 and k < s:
                        out[i].append(res[i])
        assert probs[1].error()[0] == CAPACITY
        with pytest.raises(q.QsgError) as e:
            probs[1].run(1)
        assert e.value.code == STATE
        for i in (0, 2, 3):
            _check(names[i], golds[i], out[i])
    finally:
        for p in probs:
            p.close()


@pytest.mark.gpu
def test_server_isolates_failed_request_and_cancels():
    print("[failures] server: undersized request, cancelled request, others exact", flush=True)
    gold = {n: load_golden(n) for n in ("nested_small", "tiny_cube", "spheres", "small_cube_short")}
    text = {n: open(g["deck_path"]).read() for n, g in gold.items()}
    bad_text, cap, _ = _small(gold["small_cube_short"])
    with q.Server(64 << 30) as s:
        ids = {"nested_small": s.submit(text["nested_small"]),
               "bad": s.submit(bad_text, particle_capacity=cap),
               "tiny_cube": s.submit(text["tiny_cube"])}
        s.step()
        assert s.status(ids["bad"])[0] == q.Server.FAILED
        with pytest.raises(q.QsgError) as e:
            s.result(ids["bad"])
        assert e.value.code == CAPACITY
        ids["spheres"] = s.submit(text["spheres"])
        s.step()
        # cancel a running request between steps: its slot and memory are
        # freed, the cycles it finished stay readable
        assert s.status(ids["nested_small"]) == (q.Server.RUNNING, 2)
        s.cancel(ids["nested_small"])
        assert s.status(ids["nested_small"]) == (q.Server.CANCELLED, 2)
        _check("nested_small", gold["nested_small"], s.result(ids["nested_small"]), n=2)
        queued = s.submit(text["tiny_cube"])
        s.cancel(queued)
        assert s.status(queued)[0] == q.Server.CANCELLED
        for _ in range(50):
            active, pending = s.step()
            if active == 0 and pending == 0:
                break
        for name in ("tiny_cube", "spheres"):
            assert s.status(ids[name])[0] == q.Server.DONE, name
            _check(name, gold[name], s.result(ids[name]))
        # a running request can also be released (cancelled and forgotten)
        rid = s.submit(text["tiny_cube"])
        s.step()
        s.release(rid)
        with pytest.raises(q.QsgError):
            s.status(rid)
        assert s.stats()["bytes_in_use"] == 0


def _too_many_products(g):
    """nuBar 4.5: (int)(nuBar + r) is 5 for half the fissions
    (NuclearData.cc:71), more than MAX_PRODUCTION_SIZE."""
    text = open(g["deck_path"]).read()
    assert text.count("nuBar: 2.0") == 1
    return text.replace("nuBar: 2.0", "nuBar: 4.5")


@pytest.mark.gpu
def test_plain_run_too_many_fission_products_is_numeric_error():
    print("[failures] plain qsg_run, fission multiplicity above 4", flush=True)
    g = load_golden("small_cube_short")
    with q.Problem(_too_many_products(g)) as p:
        with pytest.raises(q.QsgError) as e:
            p.run(1)
        assert e.value.code == NUMERIC
        code, msg = p.error()
        assert code == NUMERIC and "more than 4" in msg
        with pytest.raises(q.QsgError) as e2:
            p.run(1)
        assert e2.value.code == STATE


@pytest.mark.gpu
def test_batch_isolates_numeric_error():
    print("[failures] batch with one problem over the production limit", flush=True)
    good = load_golden("tiny_cube")
    bad = load_golden("small_cube_short")
    probs = [q.Problem.from_file(good["deck_path"]), q.Problem(_too_many_products(bad))]
    try:
        with q.Batch(probs) as b:
            out = []
            first = b.step(strict=False)
            assert first[1] is None
            assert b.problem_error(1) == NUMERIC and b.problem_error(0) == 0
            out.append(first[0])
            for _ in range(1, len(good["cycles"])):
                out.append(b.step()[0])
        _check("tiny_cube", good, out)
        assert probs[1].error()[0] == NUMERIC
    finally:
        for p in probs:
            p.close()

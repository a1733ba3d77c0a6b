"""Device path vs the reference CPU path on the same decks.

Fixtures: tests/golden/<deck>.json, written by tests/golden/make_golden.py
from the unmodified reference (oracle/_ref/qs_ref).  Bar: the 13 balance
tallies of every cycle are bit-exact; the per-cycle scalar-flux sum agrees
to 1e-10 relative.
"""
import numpy as np
import pytest

import quicksilver_amd as q
from conftest import golden_names, load_golden

FLUX_RTOL = 1e-10

DECKS = [n for n in golden_names()]


def _run(name):
    g = load_golden(name)
    with q.Problem.from_file(g["deck_path"]) as p:
        out = p.run()
    return g, out


@pytest.mark.gpu
@pytest.mark.parametrize("name", DECKS)
def test_balance_and_flux_match_reference(name):
    print(f"[parity] deck={name}", flush=True)
    g, out = _run(name)
    assert len(out) == len(g["cycles"])
    for c, (ref, (bal, flux, _)) in enumerate(zip(g["cycles"], out)):
        for f in q.BALANCE_FIELDS:
            assert bal[f] == ref[f], f"cycle {c} {f}: device {bal[f]} reference {ref[f]}"
        rf = ref["flux_sum"]
        if rf == 0.0:
            assert flux == 0.0
        else:
            assert abs(flux - rf) <= FLUX_RTOL * abs(rf), f"cycle {c} flux {flux!r} vs {rf!r}"


def _check_against(g, out, label):
    assert len(out) == len(g["cycles"])
    for c, (ref, (bal, flux, _)) in enumerate(zip(g["cycles"], out)):
        for f in q.BALANCE_FIELDS:
            assert bal[f] == ref[f], f"{label} cycle {c} {f}: device {bal[f]} reference {ref[f]}"
        rf = ref["flux_sum"]
        assert abs(flux - rf) <= FLUX_RTOL * abs(rf), f"{label} cycle {c} flux {flux!r} vs {rf!r}"


# Launch shapes that change which wave tracks which particle, when
# secondaries are tracked (same launch from a wave stack, or a later pass)
# and how much of a stack is usable; none may change a tally.
SHAPES = {
    "passes_only": {"stack_entries": 0},
    "small_stack": {"stack_entries": 512, "block_size": 64, "waves_per_cu": 2},
    "big_stack_fine_refill": {"stack_entries": 4096, "refill_min": 1, "chunk": 16},
    "coarse_chunks": {"block_size": 128, "waves_per_cu": 32, "refill_min": 32, "chunk": 1024},
    "registers_2_waves": {"waves_per_simd": 2},
    "registers_4_waves": {"waves_per_simd": 4, "waves_per_cu": 16},
    "sorted_every_cycle": {"sort_interval": 1},
    "sorted_every_2nd_passes_only": {"sort_interval": 2, "stack_entries": 0},
}


@pytest.mark.gpu
@pytest.mark.parametrize("shape", sorted(SHAPES))
@pytest.mark.parametrize("name", ["small_cube_short", "nested_small", "spheres"])
def test_launch_shape_does_not_change_tallies(name, shape):
    print(f"[parity] deck={name} shape={shape}", flush=True)
    g = load_golden(name)
    with q.Problem.from_file(g["deck_path"]) as p:
        p.set_tuning(**SHAPES[shape])
        out = p.run()
    _check_against(g, out, shape)


@pytest.mark.gpu
def test_census_staging_flushed_within_a_pass():
    """A wave stages its census particles (256 entries after its stack) and
    appends them to the census buffer when fewer than 64 entries are left,
    or when its queue is drained.  With 256 waves (64-thread workgroups, one
    wave per CU), checker_small's 1.5e5 census particles per cycle give each
    wave about 580, so every wave flushes within the pass as well as at its
    end; the next cycle tracks what those flushes wrote.  Every cycle must
    match the reference's tallies."""
    print("[parity] deck=checker_small shape=one wave per CU (census flushed within the pass)", flush=True)
    g = load_golden("checker_small")
    assert min(c["census"] for c in g["cycles"]) > 256 * 256  # > 256 per wave
    with q.Problem.from_file(g["deck_path"]) as p:
        p.set_tuning(block_size=64, waves_per_cu=1)
        out = p.run()
    _check_against(g, out, "one wave per CU")


@pytest.mark.gpu
def test_reused_device_blocks_do_not_change_tallies():
    """A released problem's device buffers are kept and handed to the next
    problem of a similar size, with their old contents.  Problems built
    one after another on reused blocks (a deck, another deck, the first
    again, and the first beside a second copy of itself) must each match
    the reference's tallies."""
    print("[parity] decks built on reused device blocks", flush=True)
    gs = {n: load_golden(n) for n in ("nested_small", "small_cube_short")}
    for name in ("nested_small", "small_cube_short", "nested_small"):
        with q.Problem.from_file(gs[name]["deck_path"]) as p:
            out = p.run()
        _check_against(gs[name], out, f"{name} on reused blocks")
    with q.Problem.from_file(gs["nested_small"]["deck_path"]) as a, \
            q.Problem.from_file(gs["nested_small"]["deck_path"]) as b:
        out_a, out_b = a.run(), b.run()
    _check_against(gs["nested_small"], out_a, "first of two")
    _check_against(gs["nested_small"], out_b, "second of two")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tiny_cube", "nested_small"])
def test_flux_per_bin_and_census_particles(name):
    """Per-(cell, group) flux of the last cycle and every census particle
    (matched by its 64-bit seed) against the reference's dumps: one material
    on 5^3 cells, and three nested materials on 12^3 cells."""
    import os
    from conftest import GOLDEN
    print(f"[parity] deck={name}: flux bins and census of the last cycle", flush=True)
    g = load_golden(name)
    if name == "tiny_cube":
        ref_flux = np.load(os.path.join(GOLDEN, "tiny_cube_flux.npy"))
    else:
        ref_flux = np.load(os.path.join(GOLDEN, name + "_flux.npz"))["flux"]
    cen = np.load(os.path.join(GOLDEN, name + "_census.npz"))
    with q.Problem.from_file(g["deck_path"]) as p:
        n = p.info().n_steps
        p.run(n - 1)
        p.cycle_init()
        p.cycle_tracking()
        flux = p.scalar_flux()
        census = p.census()
        bal, _ = p.cycle_finalize()
    assert flux.shape == ref_flux.shape
    assert (ref_flux != 0).sum() > 2000  # the sourced energies fill the top groups
    scale = np.abs(ref_flux).max()
    np.testing.assert_allclose(flux, ref_flux, rtol=1e-10, atol=1e-12 * scale)
    assert bal["census"] == len(cen["seed"]) == len(census)
    order_d = np.argsort(census.seed)
    order_r = np.argsort(cen["seed"])
    np.testing.assert_array_equal(census.seed[order_d], cen["seed"][order_r])
    np.testing.assert_array_equal(census.cell[order_d], cen["cell"][order_r])
    np.testing.assert_allclose(census.state()[order_d], cen["state"][order_r], rtol=1e-11, atol=1e-9)


@pytest.mark.gpu
def test_device_fp64_builtins_round_like_ieee():
    """sqrt, /, * and u64->f64 are correctly rounded on the device (bit-equal
    to numpy); log/sin/cos stay within 2 ulp of the host libm."""
    rng = np.random.default_rng(7)
    a = rng.random(200000) * 10.0 ** rng.integers(-20, 20, 200000)
    b = rng.random(200000) * 10.0 ** rng.integers(-20, 20, 200000) + 1e-300
    np.testing.assert_array_equal(q.math_probe(0, a), np.sqrt(a))
    np.testing.assert_array_equal(q.math_probe(1, a, b), a / b)
    np.testing.assert_array_equal(q.math_probe(5, a, b), a * b)
    u = rng.integers(0, 2**63, 200000, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    np.testing.assert_array_equal(q.math_probe(6, u.view(np.float64)), u.astype(np.float64))
    r = rng.random(200000)
    for op, fn, x in [(2, np.log, r), (3, np.sin, r * 6.2831853), (4, np.cos, r * 6.2831853)]:
        d = q.math_probe(op, x)
        h = fn(x)
        ulp = np.spacing(np.abs(h))
        assert np.max(np.abs(d - h) / ulp) <= 2.0


@pytest.mark.gpu
def test_device_sincos_equals_sin_and_cos():
    """The device's sincos gives bit for bit the device's sin and cos, on the
    azimuth angles updateTrajectory draws (2 * 3.14159265 * rngSample,
    CollisionEvent.cc:33-34) and on a wider range, so the rotation may take
    both from one call."""
    rng = np.random.default_rng(11)
    r = rng.random(1 << 20)
    for x in (2 * 3.14159265 * r, (r - 0.5) * 1e4):
        np.testing.assert_array_equal(q.math_probe(7, x), q.math_probe(3, x))
        np.testing.assert_array_equal(q.math_probe(8, x), q.math_probe(4, x))



@pytest.mark.gpu
@pytest.mark.parametrize("name,shape", [("nested_small", "default"), ("coral2_p2", "default"),
                                        ("all_escape", "default"), ("nested_small", "passes_only"),
                                        ("spheres", "passes_only")])
def test_run_matches_the_three_phase_calls(name, shape):
    """qsg_run enqueues each cycle before it waits for the previous one (the
    set-up reads the census count on the device); a cycle whose first pass
    leaves secondaries holds the kernels enqueued after it, and the host
    finishes it with further passes (every cycle with fission does so when
    the wave stacks are off).  The three per-cycle entry points synchronise
    after each phase.  Both must give the same tallies bit for bit, the same
    stats counts, and flux sums within the flux tolerance (the flux bins are
    FP64 atomic sums, whose last bits depend on the order the device
    happens to add them in)."""
    print(f"[parity] qsg_run vs cycle_init/tracking/finalize, deck={name} shape={shape}", flush=True)
    g = load_golden(name)
    n = len(g["cycles"])
    knobs = SHAPES.get(shape, {})
    with q.Problem.from_file(g["deck_path"]) as p:
        p.set_tuning(**knobs)
        fast = p.run(n)
        st_fast = p.stats()
    steps = []
    with q.Problem.from_file(g["deck_path"]) as p:
        p.set_tuning(**knobs)
        for _ in range(n):
            p.cycle_init()
            p.cycle_tracking()
            steps.append(p.cycle_finalize())
        st_step = p.stats()
    if shape == "passes_only":
        assert st_fast["track_launches"] > st_fast["cycles"], "no cycle needed a second pass"
    for c, ((b1, f1, _), (b2, f2)) in enumerate(zip(fast, steps)):
        assert b1 == b2, f"cycle {c}: {b1} vs {b2}"
        assert abs(f1 - f2) <= FLUX_RTOL * abs(f2), f"cycle {c} flux {f1!r} vs {f2!r}"
    for k in ("segments", "collisions", "facet_crossings", "particle_loads", "cycles", "track_launches"):
        assert st_fast[k] == st_step[k], k
    _check_against(g, [(b, f, 0.0) for b, f, _ in fast], "qsg_run")


@pytest.mark.gpu
@pytest.mark.parametrize("name,shape", [("nested_small", "default"), ("spheres", "sorted_every_cycle"),
                                        ("coral2_p1", "default"), ("nonflat", "passes_only")])
def test_runs_in_pieces_and_mixed_with_single_steps(name, shape):
    """Cycles split over several qsg_run calls, with single-step cycles in
    between, continue from one another: each run starts from the host's
    census count and leaves its last cycle closed (its tail does no set-up),
    and a single step afterwards starts from the same state."""
    print(f"[parity] runs in pieces, deck={name} shape={shape}", flush=True)
    g = load_golden(name)
    n = len(g["cycles"])
    out = []
    with q.Problem.from_file(g["deck_path"]) as p:
        p.set_tuning(**SHAPES.get(shape, {}))
        pieces = [1, 2, 1]
        done = 0
        while done < n:
            for k in pieces:
                if done >= n:
                    break
                k = min(k, n - done)
                out += p.run(k)
                done += k
                if done < n:
                    p.cycle_init()
                    p.cycle_tracking()
                    bal, flux = p.cycle_finalize()
                    out.append((bal, flux, 0.0))
                    done += 1
        assert p.stats()["cycles"] == n
        fluence = p.fluence() if name.startswith("coral2") else None
    _check_against(g, out, f"pieces/{shape}")
    if fluence is not None:
        # the cumulative fluence matches one uninterrupted run's
        with q.Problem.from_file(g["deck_path"]) as p:
            p.set_tuning(**SHAPES.get(shape, {}))
            p.run(n)
            whole = p.fluence()
        np.testing.assert_allclose(fluence, whole, rtol=1e-10, atol=0.0)


@pytest.mark.gpu
def test_run_stats_are_consistent():
    """qsg_run's device timings: every cycle's tracking time is positive and
    covers its kernel time; the time between cycles counts as set-up and the
    last cycle's copy and finalize as finalize."""
    g = load_golden("nested_small")
    n = len(g["cycles"])
    with q.Problem.from_file(g["deck_path"]) as p:
        out = p.run(n)
        st = p.stats()
    assert st["cycles"] == n and st["track_launches"] >= n
    assert all(t > 0.0 for _, _, t in out)
    assert st["init_seconds"] > 0.0 and st["finalize_seconds"] > 0.0
    assert st["track_kernel_seconds"] <= st["tracking_seconds"] * (1 + 1e-6)
    assert abs(sum(t for _, _, t in out) - st["tracking_seconds"]) <= 1e-9 + 1e-6 * st["tracking_seconds"]
    assert st["segments"] == sum(c["num_segments"] for c in g["cycles"])

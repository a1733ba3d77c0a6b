"""The single large problem at one GPU's full share: 80^3 cells (12.3 M
facets), two materials in a seeded checkerboard of 8^3 blocks, 20 isotopes x
3 reactions with non-flat polynomial cross sections, 4e7 particles, 20
steps, population control active.  No CPU run of this size finishes in
seconds, so the checks are the size-independent ones:

* every cycle satisfies the balance identities (quicksilver_amd.checks,
  pinned on every reference fixture in test_invariants.py);
* the tallies do not depend on how the work is executed: the first cycles
  run with secondaries through extra passes instead of wave stacks, with a
  different register budget, with the queue sorted by cell, and inside a
  batch, give bit-identical balance tallies and flux within 1e-10;
* the same generator at 40^3 (checker_small) is bit-exact against the
  reference (test_gpu_parity.py).
"""
import numpy as np
import pytest

import quicksilver_amd as q
from quicksilver_amd import checks, decks

FLUX_RTOL = 1e-10


@pytest.fixture(scope="module")
def deck_text():
    return decks.checkerboard().text()


@pytest.mark.gpu
def test_checkerboard_80_full_run_balance_identities(deck_text):
    print("[large] checkerboard 80^3, 4e7 particles, 20 steps", flush=True)
    with q.Problem(deck_text) as p:
        info = p.info()
        assert (info.nx, info.n_particles_target, info.n_steps) == (80, 40000000, 20)
        out = p.run()
        st = p.stats()
    bals = [b for b, _, _ in out]
    fluxes = np.array([f for _, f, _ in out])
    print(f"[large] {st['segments']} segments, {st['segments'] / st['track_kernel_seconds']:.3e} segments/s "
          f"of tracking-kernel time", flush=True)
    assert checks.balance_violations(bals, info.source_per_cycle) == []
    assert np.all(np.isfinite(fluxes)) and np.all(fluxes > 0)
    assert bals[-1]["census"] > 0


@pytest.mark.gpu
def test_checkerboard_80_execution_order_does_not_change_tallies(deck_text):
    print("[large] execution-order invariance, 2 cycles", flush=True)
    runs = {}
    for label, shape in [("default", {}), ("passes_2_waves", {"stack_entries": 0, "waves_per_simd": 2}),
                         ("small_grid", {"block_size": 128, "waves_per_cu": 4, "chunk": 256}),
                         ("sorted", {"sort_interval": 1})]:
        with q.Problem(deck_text) as p:
            if shape:
                p.set_tuning(**shape)
            runs[label] = [(b, f) for b, f, _ in p.run(2)]
    other = q.Problem(decks.small_cube(n_particles=20000, n_steps=3).text())
    p = q.Problem(deck_text)
    try:
        with q.Batch([p, other]) as b:
            runs["batched"] = b.run([2, 2])[0]
    finally:
        p.close()
        other.close()
    ref = runs["default"]
    for label, got in runs.items():
        for c, ((rb, rf), (gb, gf)) in enumerate(zip(ref, got)):
            assert gb == rb, f"{label} cycle {c}"
            assert abs(gf - rf) <= FLUX_RTOL * abs(rf), (label, c)

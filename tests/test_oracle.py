"""The C restatement (oracle/qs_oracle.c) pinned against the reference.

Every balance tally of every cycle must equal the reference CPU path's
(tests/golden/*.json, produced by the unmodified reference) and the scalar
flux sum must agree to 1e-10 relative.  The RNG pieces are additionally
checked against closed forms that a plausible slip would break.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden

LIB = os.path.join(ROOT, "oracle", "_build", "libqsoracle.so")
FIELDS = ["start", "source", "rr", "split", "absorb", "scatter", "fission", "produce",
          "collision", "escape", "census", "num_segments", "end"]

# decks small enough for a single-threaded C run in seconds
CPU_DECKS = ["tiny_cube", "small_cube_short", "small_cube", "nested_small", "all_absorb",
             "all_escape", "no_collisions", "scatter_only", "no_fission", "nonflat", "spheres"]


@pytest.fixture(scope="module")
def oracle():
    if not os.path.exists(LIB):
        pytest.fail("oracle/_build/libqsoracle.so missing: run make -C oracle restatement")
    L = ctypes.CDLL(LIB)
    L.qso_run_deck.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.qso_hash_state.restype = ctypes.c_uint64
    L.qso_hash_state.argtypes = [ctypes.c_uint64]
    L.qso_rng_sample.restype = ctypes.c_double
    L.qso_rng_sample.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    L.qso_rng_spawn.restype = ctypes.c_uint64
    L.qso_rng_spawn.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    return L


def run_oracle(L, deck_path, cycles):
    bal = (ctypes.c_uint64 * (13 * cycles))()
    flux = (ctypes.c_double * cycles)()
    rc = L.qso_run_deck(deck_path.encode(), cycles, bal, flux, None)
    assert rc == 0, rc
    return [dict(zip(FIELDS, bal[13 * c:13 * c + 13])) for c in range(cycles)], list(flux)


@pytest.mark.parametrize("name", CPU_DECKS)
def test_oracle_matches_reference(oracle, name):
    g = load_golden(name)
    bals, fluxes = run_oracle(oracle, g["deck_path"], len(g["cycles"]))
    for c, (ref, bal, fl) in enumerate(zip(g["cycles"], bals, fluxes)):
        for f in FIELDS:
            assert bal[f] == ref[f], f"{name} cycle {c} {f}: oracle {bal[f]} reference {ref[f]}"
        rf = ref["flux_sum"]
        assert abs(fl - rf) <= 1e-10 * abs(rf), (name, c, fl, rf)


def test_lcg_closed_form(oracle):
    """s <- 2862933555777941757 s + 3037000493 (mod 2^64); r = s * 5.42e-20
    (MC_RNG_State.hh:26-31)."""
    rng = np.random.default_rng(1)
    for s0 in [0, 1, 2**63 + 12345, 2**64 - 1] + [int(x) for x in rng.integers(0, 2**63, 50)]:
        s = ctypes.c_uint64(s0)
        r = oracle.qso_rng_sample(ctypes.byref(s))
        expect = (2862933555777941757 * s0 + 3037000493) % 2**64
        assert s.value == expect
        assert r == 5.4210108624275222e-20 * float(expect)
    # the state 0 maps to the increment, so the stream never sticks at 0
    s = ctypes.c_uint64(0)
    oracle.qso_rng_sample(ctypes.byref(s))
    assert s.value == 3037000493


def test_spawn_hashes_parent_then_advances_it(oracle):
    s = ctypes.c_uint64(123456789)
    h = oracle.qso_hash_state(123456789)
    child = oracle.qso_rng_spawn(ctypes.byref(s))
    assert child == h
    assert s.value == (2862933555777941757 * 123456789 + 3037000493) % 2**64
    # psdes mixes both halves: flipping one input bit changes ~half the output bits
    flips = [bin(oracle.qso_hash_state(1 << k) ^ oracle.qso_hash_state(0)).count("1") for k in range(64)]
    assert 20 < np.mean(flips) < 44
    assert len({oracle.qso_hash_state(k) for k in range(4096)}) == 4096

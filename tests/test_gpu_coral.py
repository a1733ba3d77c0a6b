"""CORAL2 benchmark decks (the reference's Examples/CORAL2_Benchmark,
one-rank size, 100 cycles): the device run reproduces the reference's
tallies bit-exactly, its four CORAL2 correctness verdicts
(CoralBenchmark.cc), and its per-cell fluence to 1e-10 -- alone and inside a
batch."""
import os

import numpy as np
import pytest

import quicksilver_amd as q
from conftest import GOLDEN, load_golden

FLUX_RTOL = 1e-10
NAMES = ["coral2_p1", "coral2_p2"]


def _check_cycles(name, g, out):
    assert len(out) == len(g["cycles"]), name
    for c, (ref, got) in enumerate(zip(g["cycles"], out)):
        bal, flux = got[0], got[1]
        for f in q.BALANCE_FIELDS:
            assert bal[f] == ref[f], f"{name} cycle {c} {f}: device {bal[f]} reference {ref[f]}"
        assert abs(flux - ref["flux_sum"]) <= FLUX_RTOL * abs(ref["flux_sum"]), (name, c)


def _check_coral(name, g, report, fluence):
    ref_fl = np.load(os.path.join(GOLDEN, name + "_fluence.npy"))
    assert report.verdicts() == g["coral_verdicts"]
    assert report.gains == report.losses
    assert fluence.shape == ref_fl.shape
    assert np.max(np.abs(fluence - ref_fl)) <= FLUX_RTOL * np.max(np.abs(ref_fl))


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_coral_verdicts_and_fluence_match_reference(name):
    print(f"[coral] deck={name}", flush=True)
    g = load_golden(name)
    with q.Problem.from_file(g["deck_path"]) as p:
        out = p.run()
        report = p.coral_check()
        fluence = p.fluence()
    _check_cycles(name, g, out)
    _check_coral(name, g, report, fluence)


@pytest.mark.gpu
def test_coral_checks_inside_a_batch():
    print("[coral] batched with another problem", flush=True)
    g = load_golden("coral2_p2")
    other = load_golden("tiny_cube")
    p = q.Problem.from_file(g["deck_path"])
    o = q.Problem.from_file(other["deck_path"])
    try:
        with q.Batch([p, o]) as b:
            out = b.run([len(g["cycles"]), len(other["cycles"])])
        _check_cycles("coral2_p2", g, out[0])
        _check_coral("coral2_p2", g, p.coral_check(), p.fluence())
    finally:
        p.close()
        o.close()

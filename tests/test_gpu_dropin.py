"""The drop-in cycleTracking(MonteCarlo*) inside the reference's own driver.

oracle/_ref/qs_ref_gpu is the reference CPU build (unmodified sources,
oracle/Makefile) whose driver loop calls quicksilver_amd/shim's
cycleTracking(MonteCarlo*) (main.cc:31,138) instead of the reference's
tracking loop: the reference still runs initMC, its own source, population
control and low-weight roulette on the host, cycleFinalize, the printout and
the CORAL2 checks; the shim moves the processing vaults to the device,
tracks them there, and returns the census particles, balance counters and
per-(cell, group) flux into the reference's own structures.

Bar: the reference's printed tallies (QSREF lines) match the fixtures the
plain reference wrote, bit-exact for the 13 balance counters, the flux sum
within 1e-10 relative; for tiny_cube also every flux bin (1e-10) and every
census particle (matched by seed: cell exact, state within 1e-11); for the
CORAL2 Problem 2 deck also the four verdict lines and the per-cell fluence.
"""
import os
import subprocess

import numpy as np
import pytest

import quicksilver_amd as q
from conftest import GOLDEN, ROOT, load_golden

QS_REF_GPU = os.path.join(ROOT, "oracle", "_ref", "qs_ref_gpu")
FLUX_RTOL = 1e-10


def _run(deck_path, tmp_path, env_extra=None):
    if not os.path.exists(QS_REF_GPU):
        pytest.skip("oracle/_ref/qs_ref_gpu is built only where the reference sources are present")
    env = dict(os.environ, OMP_NUM_THREADS="8", **(env_extra or {}))
    out = subprocess.run([QS_REF_GPU, "-i", deck_path], capture_output=True, text=True, env=env,
                         cwd=str(tmp_path), timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    cycles, verdicts = [], []
    for line in out.stdout.splitlines():
        if "PASS::" in line or "FAIL::" in line:
            verdicts.append(line.strip())
        tok = line.split()
        if tok and tok[0] == "QSREF":
            cycles.append(({f: int(t) for f, t in zip(q.BALANCE_FIELDS, tok[2:15])}, float(tok[15])))
    return cycles, verdicts


def _check(name, g, cycles):
    assert len(cycles) == len(g["cycles"]), name
    for c, (ref, (bal, flux)) in enumerate(zip(g["cycles"], cycles)):
        for f in q.BALANCE_FIELDS:
            assert bal[f] == ref[f], f"{name} cycle {c} {f}: drop-in {bal[f]} reference {ref[f]}"
        assert abs(flux - ref["flux_sum"]) <= FLUX_RTOL * abs(ref["flux_sum"]), (name, c, flux)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["small_cube_short", "nested_small", "spheres", "no_fission"])
def test_reference_driver_with_device_tracking(name, tmp_path):
    print(f"[dropin] qs_ref_gpu deck={name}", flush=True)
    g = load_golden(name)
    cycles, _ = _run(g["deck_path"], tmp_path)
    _check(name, g, cycles)


@pytest.mark.gpu
def test_reference_driver_flux_bins_and_census(tmp_path):
    print("[dropin] qs_ref_gpu deck=tiny_cube, flux bins and census", flush=True)
    g = load_golden("tiny_cube")
    flux_path, census_path = tmp_path / "flux.bin", tmp_path / "census.txt"
    cycles, _ = _run(g["deck_path"], tmp_path, {"QSREF_FLUX_DUMP": str(flux_path),
                                                 "QSREF_CENSUS_DUMP": str(census_path)})
    _check("tiny_cube", g, cycles)
    ref_flux = np.load(os.path.join(GOLDEN, "tiny_cube_flux.npy"))
    flux = np.fromfile(flux_path, dtype=np.float64).reshape(-1, *ref_flux.shape)[-1]
    assert np.all(np.abs(flux - ref_flux) <= FLUX_RTOL * np.abs(ref_flux)), "per-bin flux differs"
    ref = np.load(os.path.join(GOLDEN, "tiny_cube_census.npz"))
    cols = np.loadtxt(census_path, dtype=str, ndmin=2)
    seed = cols[:, 0].astype(np.uint64)
    state = cols[:, 1:11].astype(np.float64)
    cell = cols[:, 11].astype(np.int32)
    assert len(seed) == len(ref["seed"])
    a, b = np.argsort(seed), np.argsort(ref["seed"])
    assert np.array_equal(seed[a], ref["seed"][b])
    assert np.array_equal(cell[a], ref["cell"][b])
    scale = np.maximum(np.abs(ref["state"][b]), 1e-300)
    assert np.all(np.abs(state[a] - ref["state"][b]) <= 1e-11 * scale + 1e-300)


@pytest.mark.gpu
def test_reference_driver_coral2_verdicts_and_fluence(tmp_path):
    print("[dropin] qs_ref_gpu deck=coral2_p2 (100 cycles), verdicts and fluence", flush=True)
    g = load_golden("coral2_p2")
    fl_path = tmp_path / "fluence.bin"
    cycles, verdicts = _run(g["deck_path"], tmp_path, {"QSREF_FLUENCE_DUMP": str(fl_path)})
    _check("coral2_p2", g, cycles)
    assert verdicts == g["coral_verdicts"]
    ref_fl = np.load(os.path.join(GOLDEN, "coral2_p2_fluence.npy"))
    fl = np.fromfile(fl_path, dtype=np.float64)
    assert fl.shape == ref_fl.shape
    assert np.max(np.abs(fl - ref_fl)) <= FLUX_RTOL * np.max(np.abs(ref_fl))

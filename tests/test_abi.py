"""C-ABI library surface (no GPU needed)."""
import ctypes
import os
import re

import pytest

import quicksilver_amd as q
from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "qsg.h")).read()
    return sorted(set(re.findall(r"\b(qsg_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = q.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert sorted(q.EXPORTS) == names


def test_create_without_gpu_fails_loudly(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    deck = tmp_path / "d.inp"
    from quicksilver_amd import decks
    decks.small_cube(n_particles=1000, n_steps=1).write(str(deck))
    with pytest.raises(q.QsgError) as e:
        q.Problem.from_file(str(deck))
    assert e.value.code in (2,)


def test_bad_deck_is_an_argument_error(tmp_path):
    deck = tmp_path / "bad.inp"
    deck.write_text("Geometry:\n   material: m\n   shape: cone\n")
    with pytest.raises(q.QsgError) as e:
        q.Problem.from_file(str(deck))
    assert e.value.code == 1
    assert "shape" in str(e.value)

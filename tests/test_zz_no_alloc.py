"""qsg_nearest_facet_probe_noalloc (no device allocation; rays staged
through one fixed device array per GPU) against qsg_nearest_facet_probe (a
per-call allocation), bit for bit: distance, facet and nudged position.

Two calls in a row on the same problem, the second with fewer and different
rays, so a result left in the shared array by the first call would show in
the second; then the same on a second problem, whose mesh differs.  Ray
families are those of the facet-probe parity test (random interior points,
rays at face diagonals, edges, corners and centres, grazing and
axis-parallel directions, starts on faces and outside the cell)."""
import numpy as np
import pytest

import quicksilver_amd as q
from test_facet_probe import deck_text, make_rays

ARG = 1


def _same(a, b):
    da, fa, ma = a
    db, fb, mb = b
    assert np.array_equal(da.view(np.uint64), db.view(np.uint64))
    assert np.array_equal(fa, fb)
    assert np.array_equal(ma.view(np.uint64), mb.view(np.uint64))


@pytest.mark.gpu
def test_noalloc_probe_matches_allocating_probe():
    print("[no-alloc] facet probe: two calls per problem, second smaller", flush=True)
    for name, (n1, n2) in (("anisotropic", (6000, 700)), ("far", (3000, 400))):
        with q.Problem(deck_text(name)) as p:
            for n, seed in ((n1, 7), (n2, 11)):
                xyz, d, cell, _ = make_rays(name, n_per_family=n, seed=seed)
                assert 0 < len(cell) <= q.NOALLOC_MAX_RAYS, len(cell)
                print(f"[no-alloc] mesh={name} rays={len(cell)}", flush=True)
                ref = p.nearest_facet(xyz, d, cell)
                got = p.nearest_facet(xyz, d, cell, noalloc=True)
                _same(got, ref)


@pytest.mark.gpu
def test_noalloc_probe_rejects_more_rays_than_its_array():
    print("[no-alloc] facet probe: too many rays, and zero rays", flush=True)
    with q.Problem(deck_text("anisotropic")) as p:
        n = q.NOALLOC_MAX_RAYS + 1
        xyz = np.full((n, 3), 0.5)
        d = np.tile([1.0, 0.0, 0.0], (n, 1))
        cell = np.zeros(n, np.int32)
        with pytest.raises(q.QsgError) as e:
            p.nearest_facet(xyz, d, cell, noalloc=True)
        assert e.value.code == ARG
        dist, facet, moved = p.nearest_facet(xyz[:0], d[:0], cell[:0], noalloc=True)
        assert dist.size == 0 and facet.size == 0 and moved.size == 0

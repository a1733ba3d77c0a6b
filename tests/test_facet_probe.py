"""Nearest-facet search, ray by ray.

The tracking kernels find the nearest facet with a shortcut: one probe of
each exit face names the one or two triangles that can contain the exit
point, and only those get the reference's exact test (all 12 exit-face
facets when the probe cannot decide).  Full runs exercise it billions of
times but rarely on the edge cases, so here the device search
(qsg_nearest_facet_probe, the same device function) is compared bit for bit
with the oracle's full 24-facet scan (qso_nearest_facet, a restatement of
MCT_Nearest_Facet, MCT.cc:87-137 / 487-621) on rays aimed at the places
where the shortcut decides: face diagonals, centres, edges and corners,
offsets straddling the probe's margin, grazing and axis-parallel
directions, and starting points on or just outside the cell's faces.

The oracle's search is pinned, beyond the full-run fixtures, by the
reference's own MCT_Nearest_Facet on the same rays (oracle/_ref/qs_ref_mct,
the unmodified reference objects, bit for bit), by the closed form of a ray
leaving a box (distance and face), and by the invariant that each facet
index stands for one triangle of its face.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "oracle", "_build", "libqsoracle.so")
QS_REF_MCT = os.path.join(ROOT, "oracle", "_ref", "qs_ref_mct")

_DECK = """Simulation:
   dt: 1e-08
   boundaryCondition: reflect
   lx: {lx!r}
   ly: {ly!r}
   lz: {lz!r}
   nParticles: 1000
   nSteps: 1
   nx: {nx}
   ny: {ny}
   nz: {nz}
   seed: 1
   xDom: 1
   yDom: 1
   zDom: 1
   eMax: 20.0
   eMin: 1e-09
   nGroups: 16
   lowWeightCutoff: 0.001

Geometry:
   material: m
   shape: brick
   xMax: {lx!r}
   xMin: 0.0
   yMax: {ly!r}
   yMin: 0.0
   zMax: {lz!r}
   zMin: 0.0

Material:
   name: m
   mass: 1000.0
   nIsotopes: 1
   nReactions: 3
   sourceRate: 10000000000.0
   totalCrossSection: 1.0
   absorptionCrossSection: flat
   fissionCrossSection: flat
   scatteringCrossSection: flat
   absorptionCrossSectionRatio: 0.3
   fissionCrossSectionRatio: 0.1
   scatteringCrossSectionRatio: 0.6

CrossSection:
   name: flat
   A: 0.0
   B: 0.0
   C: 0.0
   D: 0.0
   E: 1.0
   nuBar: 2.0
"""

# (lx, ly, lz, nx, ny, nz): unequal cell sides, cells far from the origin
# (the reference's plane tolerance scales with |r|^2), and tiny cells
# (where the probe's absolute margin floor applies)
MESHES = {
    "anisotropic": (3.0, 5.0, 7.0, 3, 4, 5),
    "far": (1.0e4, 2.0e4, 3.0e3, 4, 4, 3),
    "tiny": (1.0e-3, 1.0e-3, 2.0e-3, 2, 2, 2),
}


def deck_text(name):
    lx, ly, lz, nx, ny, nz = MESHES[name]
    return _DECK.format(lx=lx, ly=ly, lz=lz, nx=nx, ny=ny, nz=nz)


@pytest.fixture(scope="module")
def oracle():
    if not os.path.exists(LIB):
        pytest.fail("oracle/_build/libqsoracle.so missing: run make -C oracle restatement")
    L = ctypes.CDLL(LIB)
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int32)
    L.qso_nearest_facet_rays.argtypes = [ctypes.c_char_p, ctypes.c_uint64, D, D, I, D, I, D]
    return L


def oracle_rays(L, deck_path, xyz, d, cell):
    n = cell.size
    xyz = np.ascontiguousarray(xyz, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    cell = np.ascontiguousarray(cell, np.int32)
    dist = np.empty(n, np.float64)
    facet = np.empty(n, np.int32)
    moved = np.empty((n, 3), np.float64)
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int32)
    rc = L.qso_nearest_facet_rays(deck_path.encode(), n, xyz.ctypes.data_as(D), d.ctypes.data_as(D),
                                  cell.ctypes.data_as(I), dist.ctypes.data_as(D), facet.ctypes.data_as(I),
                                  moved.ctypes.data_as(D))
    assert rc == 0, rc
    return dist, facet, moved


def reference_rays(deck_path, xyz, d, cell, tmp_path):
    """The reference's MCT_Nearest_Facet on the rays (oracle/_ref/qs_ref_mct)."""
    import subprocess
    n = cell.size
    rays = tmp_path / "rays.bin"
    out = tmp_path / "out.bin"
    with open(rays, "wb") as f:
        f.write(np.uint64(n).tobytes())
        f.write(np.ascontiguousarray(np.concatenate([xyz, d], axis=1), np.float64).tobytes())
        f.write(np.ascontiguousarray(cell, np.int32).tobytes())
    r = subprocess.run([QS_REF_MCT, "-i", deck_path, str(rays), str(out)], capture_output=True, text=True,
                       cwd=str(tmp_path), timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    raw = open(out, "rb").read()
    dist = np.frombuffer(raw[:8 * n], np.float64)
    facet = np.frombuffer(raw[8 * n:12 * n], np.int32)
    moved = np.frombuffer(raw[12 * n:], np.float64).reshape(n, 3)
    return dist, facet, moved


def _geometry(name):
    lx, ly, lz, nx, ny, nz = MESHES[name]
    return np.array([lx / nx, ly / ny, lz / nz]), (nx, ny, nz)


def _cells(rng, dims, n):
    nx, ny, nz = dims
    c = rng.integers(0, nx * ny * nz, n)
    ijk = np.stack([c % nx, (c // nx) % ny, c // (nx * ny)], axis=1)
    return c.astype(np.int32), ijk


def _normalise(v):
    return v / np.sqrt((v * v).sum(axis=1))[:, None]


def _interior(rng, lo, h, n):
    return lo + h * rng.uniform(0.02, 0.98, (n, 3))


def make_rays(name, n_per_family=4000, seed=7):
    """Families of rays on the mesh `name`: (xyz, direction, cell, family)."""
    h, dims = _geometry(name)
    rng = np.random.default_rng(seed)
    out = []

    def add(fam, xyz, d, cell):
        keep = np.isfinite(d).all(axis=1) & ((d * d).sum(axis=1) > 0)
        out.append((xyz[keep], d[keep], cell[keep], np.full(keep.sum(), fam)))

    n = n_per_family
    # random interior points, isotropic directions
    c, ijk = _cells(rng, dims, n)
    lo = ijk * h
    add("random", _interior(rng, lo, h, n), _normalise(rng.normal(size=(n, 3))), c)

    # rays aimed at points of an exit face: diagonals, edges, corners and the
    # centre, displaced by offsets around the probe's margin and the
    # reference's 1e-9 tolerances
    def face_targets(kind):
        c, ijk = _cells(rng, dims, n)
        lo = ijk * h
        ax = rng.integers(0, 3, n)
        up = rng.integers(0, 2, n)
        u_ax, v_ax = (ax + 1) % 3, (ax + 2) % 3
        rows = np.arange(n)
        hu, hv = 0.5 * h[u_ax], 0.5 * h[v_ax]
        cu, cv = lo[rows, u_ax] + hu, lo[rows, v_ax] + hv
        margin = np.maximum(1e-6 * np.minimum(hu, hv), 1e-7)
        rel = np.array([0.0, 1e-15, 1e-12, 1e-9, 1e-7, 1e-6, 1e-5, 1e-3])
        offs = np.concatenate([rel, -rel])
        scale = np.maximum(hu, hv)
        d1 = offs[rng.integers(0, offs.size, n)] * scale
        d2 = offs[rng.integers(0, offs.size, n)] * scale
        # offsets straddling the margin itself
        near_m = rng.random(n) < 0.3
        k = np.array([1 - 1e-9, 1.0, 1 + 1e-9, 1 - 1e-3, 1 + 1e-3, 0.5, 2.0])
        d1 = np.where(near_m, margin * k[rng.integers(0, k.size, n)] * rng.choice([-1, 1], n), d1)
        t = rng.uniform(-1, 1, n)
        s = rng.choice([-1.0, 1.0], n)
        if kind == "diagonal":
            pu, pv = cu + t * hu + d2, cv + s * t * hv + d1
        elif kind == "edge":
            pu, pv = cu + s * hu + d1, cv + t * hv + d2
        elif kind == "corner":
            pu, pv = cu + s * hu + d1, cv + rng.choice([-1.0, 1.0], n) * hv + d2
        else:  # centre
            pu, pv = cu + d1, cv + d2
        target = np.empty((n, 3))
        target[rows, ax] = lo[rows, ax] + up * h[ax]
        target[rows, u_ax] = pu
        target[rows, v_ax] = pv
        origin = _interior(rng, lo, h, n)
        return origin, _normalise(target - origin), c

    for kind in ("diagonal", "edge", "corner", "centre"):
        add(kind, *face_targets(kind))

    # grazing and axis-parallel directions
    c, ijk = _cells(rng, dims, n)
    lo = ijk * h
    d = _normalise(rng.normal(size=(n, 3)))
    ax = rng.integers(0, 3, n)
    eps = np.array([0.0, -0.0, 1e-300, 1e-20, 1e-12, 1e-9, 1e-6, 1e-3])
    d[np.arange(n), ax] = eps[rng.integers(0, eps.size, n)] * rng.choice([-1, 1], n)
    d = _normalise(d)
    axis_dirs = np.array([[1.0, 0, 0], [-1.0, 0, 0], [0, 1.0, 0], [0, -1.0, 0], [0, 0, 1.0], [0, 0, -1.0]])
    pick = rng.random(n) < 0.25
    d[pick] = axis_dirs[rng.integers(0, 6, pick.sum())]
    add("grazing", _interior(rng, lo, h, n), d, c)

    # cell centre: axis-parallel rays hit a face centre exactly
    c, ijk = _cells(rng, dims, n)
    lo = ijk * h
    d = axis_dirs[rng.integers(0, 6, n)].copy()
    tilt = rng.random(n) < 0.5
    d[tilt] = _normalise(d[tilt] + rng.normal(size=(tilt.sum(), 3)) * 1e-8)
    add("from_centre", lo + 0.5 * h, d, c)

    # starting on a face (as after a crossing), on an edge or corner of the
    # cell, or just outside it (rounding of the previous segment)
    c, ijk = _cells(rng, dims, n)
    lo = ijk * h
    xyz = _interior(rng, lo, h, n)
    rows = np.arange(n)
    n_fix = rng.integers(1, 4, n)
    for r in range(3):
        sel = n_fix > r
        axr = (rng.integers(0, 3, n) + r) % 3
        side = rng.integers(0, 2, n)
        val = lo[rows, axr] + side * h[axr]
        out_by = np.array([0.0, 0.0, 1e-15, 1e-13, 1e-11, 1e-9])[rng.integers(0, 6, n)] * h[axr]
        val = val + np.where(side == 1, out_by, -out_by)
        xyz[rows[sel], axr[sel]] = val[sel]
    add("on_boundary", xyz, _normalise(rng.normal(size=(n, 3))), c)

    # well outside, heading away: no facet ahead, the nudge-and-retry path
    m = max(n // 20, 50)
    c, ijk = _cells(rng, dims, m)
    lo = ijk * h
    ax = rng.integers(0, 3, m)
    xyz = _interior(rng, lo, h, m)
    rows = np.arange(m)
    xyz[rows, ax] = lo[rows, ax] + h[ax] * (1 + 10.0 ** rng.uniform(-8, -3, m))
    d = _normalise(rng.normal(size=(m, 3)))
    d[rows, ax] = np.abs(d[rows, ax])
    add("outside", xyz, _normalise(d), c)

    xyz = np.concatenate([o[0] for o in out])
    d = np.concatenate([o[1] for o in out])
    cell = np.concatenate([o[2] for o in out]).astype(np.int32)
    fam = np.concatenate([o[3] for o in out])
    return xyz, d, cell, fam


def _write_deck(tmp_path, name):
    path = tmp_path / f"{name}.inp"
    path.write_text(deck_text(name))
    return str(path)


# ------------------------------------------------------------- oracle pins ----

@pytest.mark.parametrize("name", sorted(MESHES))
def test_oracle_matches_reference_mct(oracle, tmp_path, name):
    """Every ray family, bit for bit: the restatement's search against the
    reference's own MCT_Nearest_Facet (distance, facet, nudged position)."""
    if not os.path.exists(QS_REF_MCT):
        pytest.skip("oracle/_ref/qs_ref_mct is built only where the reference sources are present")
    xyz, d, cell, fam = make_rays(name, n_per_family=2000, seed=5)
    deck = _write_deck(tmp_path, name)
    dist_o, facet_o, moved_o = oracle_rays(oracle, deck, xyz, d, cell)
    dist_r, facet_r, moved_r = reference_rays(deck, xyz, d, cell, tmp_path)
    bad = (dist_o.view(np.uint64) != dist_r.view(np.uint64)) | (facet_o != facet_r) | \
        (moved_o.view(np.uint64) != moved_r.view(np.uint64)).any(axis=1)
    assert not bad.any(), {f: int((fam[bad] == f).sum()) for f in np.unique(fam[bad])}
    assert (fam == "outside").sum() > 0 and (moved_r != xyz).any(), "the nudge path is exercised"


@pytest.mark.parametrize("name", sorted(MESHES))
def test_oracle_matches_ray_box_exit(oracle, tmp_path, name):
    """For a point inside the cell, the nearest facet is where the ray leaves
    the box: distance min over axes of (bound - x) / d, on that face.  Facet
    f lies on face f // 4 in the order +x -x +y -y +z -z."""
    h, dims = _geometry(name)
    xyz, d, cell, fam = make_rays(name, n_per_family=3000)
    keep = np.isin(fam, ["random", "diagonal", "edge", "corner", "centre", "grazing", "from_centre"])
    xyz, d, cell = xyz[keep], d[keep], cell[keep]
    dist, facet, moved = oracle_rays(oracle, _write_deck(tmp_path, name), xyz, d, cell)
    assert (moved == xyz).all(), "no nudge for points inside the cell"
    nx, ny, nz = dims
    ijk = np.stack([cell % nx, (cell // nx) % ny, cell // (nx * ny)], axis=1)
    lo = ijk * h
    hi = lo + h
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(d > 0, (hi - xyz) / d, np.where(d < 0, (lo - xyz) / d, np.inf))
    order = np.sort(t, axis=1)
    best = np.argmin(t, axis=1)
    face = 2 * best + (d[np.arange(d.shape[0]), best] < 0)
    scale = h.max()
    np.testing.assert_allclose(dist, order[:, 0], rtol=1e-9, atol=1e-9 * scale)
    clear = (order[:, 1] - order[:, 0]) > 1e-6 * scale
    assert clear.sum() > 0.5 * clear.size
    assert (facet[clear] // 4 == face[clear]).all()


@pytest.mark.parametrize("name", sorted(MESHES))
def test_oracle_facet_is_one_triangle_of_its_face(oracle, tmp_path, name):
    """Each face is split by its diagonals into four triangles, one per
    edge.  Away from the diagonals the exit point names the edge; every exit
    through facet f must name the same edge, and the four facets of a face
    four different edges."""
    h, dims = _geometry(name)
    xyz, d, cell, fam = make_rays(name, n_per_family=3000, seed=11)
    keep = fam == "random"
    xyz, d, cell = xyz[keep], d[keep], cell[keep]
    dist, facet, _ = oracle_rays(oracle, _write_deck(tmp_path, name), xyz, d, cell)
    nx, ny, nz = dims
    ijk = np.stack([cell % nx, (cell // nx) % ny, cell // (nx * ny)], axis=1)
    centre = ijk * h + 0.5 * h
    p = xyz + dist[:, None] * d
    ax = facet // 8
    rows = np.arange(ax.size)
    u = (p[rows, (ax + 1) % 3] - centre[rows, (ax + 1) % 3]) / (0.5 * h[(ax + 1) % 3])
    v = (p[rows, (ax + 2) % 3] - centre[rows, (ax + 2) % 3]) / (0.5 * h[(ax + 2) % 3])
    off_diag = np.abs(np.abs(u) - np.abs(v)) > 1e-6
    edge = np.where(np.abs(u) > np.abs(v), np.where(u > 0, 1, 3), np.where(v > 0, 2, 0))
    mapping = {}
    for f, e in zip(facet[off_diag], edge[off_diag]):
        mapping.setdefault(int(f), set()).add(int(e))
    assert len(mapping) == 24, sorted(mapping)
    assert all(len(s) == 1 for s in mapping.values()), mapping
    for face in range(6):
        edges = [next(iter(mapping[4 * face + k])) for k in range(4)]
        assert sorted(edges) == [0, 1, 2, 3], (face, edges)


# ------------------------------------------------------------ device parity ----

@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MESHES))
def test_device_nearest_facet_matches_oracle(oracle, tmp_path, name):
    import quicksilver_amd as q
    print(f"[facet-probe] mesh={name}", flush=True)
    xyz, d, cell, fam = make_rays(name)
    deck = _write_deck(tmp_path, name)
    dist_o, facet_o, moved_o = oracle_rays(oracle, deck, xyz, d, cell)
    if os.path.exists(QS_REF_MCT):
        # the reference's own search, where its build travelled with the tree
        dist_r, facet_r, moved_r = reference_rays(deck, xyz, d, cell, tmp_path)
        assert np.array_equal(dist_r.view(np.uint64), dist_o.view(np.uint64))
        assert np.array_equal(facet_r, facet_o) and np.array_equal(moved_r.view(np.uint64), moved_o.view(np.uint64))
    with q.Problem(deck_text(name)) as p:
        dist_g, facet_g, moved_g = p.nearest_facet(xyz, d, cell)
    bad = (dist_g.view(np.uint64) != dist_o.view(np.uint64)) | (facet_g != facet_o) | \
        (moved_g.view(np.uint64) != moved_o.view(np.uint64)).any(axis=1)
    if bad.any():
        idx = np.flatnonzero(bad)[:5]
        detail = [(fam[i], int(cell[i]), xyz[i].tolist(), d[i].tolist(), float(dist_g[i]), int(facet_g[i]),
                   float(dist_o[i]), int(facet_o[i])) for i in idx]
        fams = {f: int((fam[bad] == f).sum()) for f in np.unique(fam[bad])}
        pytest.fail(f"{bad.sum()} of {bad.size} rays differ (by family {fams}); first: {detail}")

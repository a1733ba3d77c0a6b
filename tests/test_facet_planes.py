"""The arithmetic the tracking kernels use for a facet's plane, checked on the
CPU against the reference's own formula.

MC_General_Plane (MC_Facet_Geometry.hh:18-39) forms the plane of facet
nodes r0, r1, r2 as the cross product (r1 - r0) x (r2 - r0), then divides by
sqrt(A*A + B*B + C*C).  On this mesh each facet is two corners that share an
edge of a face, then the face centre (the corner tuples of
GlobalFccGrid.cc:47-69 and the node triples of MC_Domain.cc:42-47, written
out below), so only the component along the face's axis is nonzero, and one
of its two products is a signed zero.  The kernels therefore form that
component as +-(E * H): E the cell's extent along the shared edge, H the face
centre's offset from that edge.  These tests take random cells, evaluate the
reference's full cross product in IEEE double (numpy, no contraction), and
require the kernels' form to give the same bits; they also require that
sqrt(fl(c * c)) is |c|, which lets the kernels skip the square root, and
that the two other components are signed zeros whose signs depend only on
the facet, which lets a reflection form the plane without reading it.
"""
import numpy as np
import pytest

# corner tuples (x, y, z, face-centre kind): GlobalFccGrid.cc:47-69
CORNER = [(0, 0, 0, 0), (1, 0, 0, 0), (0, 1, 0, 0), (1, 1, 0, 0), (0, 0, 1, 0), (1, 0, 1, 0), (0, 1, 1, 0),
          (1, 1, 1, 0), (1, 0, 0, 1), (0, 0, 0, 1), (0, 1, 0, 2), (0, 0, 0, 2), (0, 0, 1, 3), (0, 0, 0, 3)]
# facet node triples: MC_Domain.cc:42-47
TRI = [(1, 3, 8), (3, 7, 8), (7, 5, 8), (5, 1, 8), (0, 4, 9), (4, 6, 9), (6, 2, 9), (2, 0, 9),
       (3, 2, 10), (2, 6, 10), (6, 7, 10), (7, 3, 10), (0, 1, 11), (1, 5, 11), (5, 4, 11), (4, 0, 11),
       (4, 5, 12), (5, 7, 12), (7, 6, 12), (6, 4, 12), (0, 2, 13), (2, 3, 13), (3, 1, 13), (1, 0, 13)]


def node_codes(f):
    """Per node, per axis: 0 -> i*d, 1 -> (i+1)*d, 2 -> i*d + d/2."""
    out = []
    for c in TRI[f]:
        x, y, z, k = CORNER[c]
        out.append((2 if k in (2, 3) else x, 2 if k in (1, 3) else y, 2 if k in (1, 2) else z))
    return out


def cell_values(rng, n):
    """Random cells of random meshes: per axis (v0, v1, vh) as the kernels
    form them (i*d, (i+1)*d, i*d + d/2)."""
    vals = []
    for _ in range(3):
        d = np.exp(rng.uniform(np.log(1e-4), np.log(1e4), n))
        i = rng.integers(0, 5000, n).astype(np.float64)
        v0 = i * d
        v1 = (i + 1) * d
        vh = v0 + d / 2.0
        vals.append((v0, v1, vh))
    return vals


def pick(vals, axis, code):
    return vals[axis][code]


@pytest.mark.parametrize("f", range(24))
def test_plane_component_is_edge_extent_times_centre_offset(f):
    rng = np.random.default_rng(100 + f)
    vals = cell_values(rng, 20000)
    codes = node_codes(f)
    ax = f >> 3
    p, q = (ax + 1) % 3, (ax + 2) % 3
    # the reference's component along ax: (p1 - p0)(q2 - q0) - (q1 - q0)(p2 - p0)
    P = [pick(vals, p, codes[m][p]) for m in range(3)]
    Q = [pick(vals, q, codes[m][q]) for m in range(3)]
    with np.errstate(all="raise"):
        ref = (P[1] - P[0]) * (Q[2] - Q[0]) - (Q[1] - Q[0]) * (P[2] - P[0])
    # the kernels' form
    p0, p1, q0, q1 = codes[0][p], codes[1][p], codes[0][q], codes[1][q]
    assert (p0 == p1) != (q0 == q1), "the corners share exactly one coordinate"
    if p0 == p1:
        E = vals[q][1] - vals[q][0]
        H = vals[p][2] - vals[p][p0]
        neg = q1 > q0
    else:
        E = vals[p][1] - vals[p][0]
        H = vals[q][2] - vals[q][q0]
        neg = p1 < p0
    c = E * H
    if neg:
        c = -c
    assert np.array_equal(c.view(np.uint64), ref.view(np.uint64))
    assert np.all(c != 0.0)
    # the face axis holds one value for the three nodes: the face is a plane
    assert len({codes[m][ax] for m in range(3)}) == 1
    # sqrt(fl(c * c)) is |c|
    assert np.array_equal(np.sqrt(c * c), np.abs(c))


def test_square_root_of_square_is_magnitude():
    """sqrt(fl(c*c)) == |c| for doubles whose square is a normal number
    (binary floating point, rounding to nearest), over mantissas spread
    through every binade edge and a wide exponent range."""
    rng = np.random.default_rng(5)
    m = np.concatenate([rng.random(2_000_000) + 1.0,
                        1.0 + rng.random(500_000) * 1e-12,
                        2.0 - rng.random(500_000) * 1e-12])
    e = rng.integers(-480, 480, m.size)
    c = np.ldexp(m, e) * np.where(rng.random(m.size) < 0.5, -1.0, 1.0)
    assert np.array_equal(np.sqrt(c * c), np.abs(c))


def reference_plane_components(r0, r1, r2):
    """MC_General_Plane's A, B, C before normalisation (MC_Facet_Geometry.hh:18-39)."""
    A = ((r1[1] - r0[1]) * (r2[2] - r0[2])) - ((r1[2] - r0[2]) * (r2[1] - r0[1]))
    B = ((r1[2] - r0[2]) * (r2[0] - r0[0])) - ((r1[0] - r0[0]) * (r2[2] - r0[2]))
    C = ((r1[0] - r0[0]) * (r2[1] - r0[1])) - ((r1[1] - r0[1]) * (r2[0] - r0[0]))
    return A, B, C


@pytest.mark.parametrize("f", range(24))
def test_zero_plane_components_have_signs_fixed_by_the_facet(f):
    """The two plane components off the face axis are signed zeros whose
    signs depend only on the facet: the same on random cells as on the
    representative nodes 1 < 2 < 3 the kernels' table is built from (a
    reflection reads them; MCT.cc:401-429)."""
    codes = node_codes(f)
    ax = f >> 3
    rep = [[(1.0, 3.0, 2.0)[codes[m][a]] for a in range(3)] for m in range(3)]
    rep_comp = reference_plane_components(*[np.array(rep[m]) for m in range(3)])
    want = [np.signbit(rep_comp[a]) for a in range(3)]
    rng = np.random.default_rng(300 + f)
    vals = cell_values(rng, 20000)
    nodes = [[pick(vals, a, codes[m][a]) for a in range(3)] for m in range(3)]
    comp = reference_plane_components(*nodes)
    for a in range(3):
        if a == ax:
            assert np.all(comp[a] != 0.0)
            continue
        assert np.all(comp[a] == 0.0)
        assert np.all(np.signbit(comp[a]) == want[a]), (f, a)
    # normalising by a positive factor keeps each sign
    inv = 1.0 / np.sqrt(comp[0] * comp[0] + comp[1] * comp[1] + comp[2] * comp[2])
    for a in range(3):
        if a != ax:
            assert np.all(np.signbit(comp[a] * inv) == want[a])

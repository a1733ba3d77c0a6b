"""The tracking kernels' closed form for the reaction a collision selects,
checked on the CPU against the reference's scan.

CollisionEvent (CollisionEvent.cc:59-82) subtracts each isotope's reaction
cross sections in turn from cur = total * random and selects the reaction at
the first partial sum below zero.  The isotopes of a material carry the
same macroscopic cross sections m_0 .. m_(R-1), so the kernels find the
crossing as isotope k = floor(cur / S), remainder cur - k S compared with the
running sums of m, and accept that answer only when the partial sums on
both sides of the crossing are farther from zero than a bound on the scan's
rounding; other lanes run the scan.  These tests run the scan in IEEE
double and require every accepted answer to equal it, on random inputs and
on inputs placed within a few roundings of each crossing.
"""
from fractions import Fraction

import numpy as np
import pytest


def scan(cur, m, n_iso):
    """The reference's loop: the reaction index, or -1 when no sum falls
    below zero."""
    c = cur
    for _ in range(n_iso):
        for r, mr in enumerate(m):
            c = c - mr
            if c < 0:
                return r
    return -1


def closed_form(cur, m, n_iso):
    """The kernels' closed form (qs_kernels.hip, CollisionEvent): the
    reaction index, or -1 where the lane falls back to the scan."""
    m0, m1, m2 = (list(m) + [0.0, 0.0])[:3]
    M1 = m0 + m1
    S = M1 + m2
    if not S > 0.0:
        return -1
    kq = float(np.floor(cur / S))
    if not kq < float(n_iso):
        return -1
    rem = float(Fraction(cur) - Fraction(kq) * Fraction(S))  # fma(-kq, S, cur)
    B = (12.0 * n_iso + 16.0) * 2.0 ** -53 * (cur + S)
    r = 0 if rem < m0 else (1 if rem < M1 else 2)
    below = 0.0 if r == 0 else (m0 if r == 1 else M1)
    at = m0 if r == 0 else (M1 if r == 1 else S)
    if r < len(m) and rem - below >= B and rem - at <= -B:
        return r
    return -1


def cases(rng, n, near):
    out = []
    for _ in range(n):
        R = int(rng.integers(1, 4))
        n_iso = int(rng.integers(1, 21))
        scale = 10.0 ** rng.uniform(-6, 6)
        m = [float(v) for v in scale * rng.uniform(0.0, 1.0, R) ** 2]
        if rng.random() < 0.1:
            m[int(rng.integers(0, R))] = 1e-20  # the reference's floor value
        total = 0.0
        for _ in range(n_iso):
            for mr in m:
                total += mr
        if near:
            # just around the exact crossing of isotope k, reaction r
            k = int(rng.integers(0, n_iso))
            r = int(rng.integers(0, R))
            exact = Fraction(k) * sum(Fraction(v) for v in m) + sum(Fraction(v) for v in m[: r + 1])
            cur = float(exact) * (1.0 + float(rng.integers(-64, 65)) * 2.0 ** -53)
            cur = max(cur, 0.0)
        else:
            cur = total * float(rng.random())
        out.append((cur, m, n_iso))
    return out


@pytest.mark.parametrize("near", [False, True])
def test_closed_form_agrees_with_the_scan(near):
    rng = np.random.default_rng(11 if near else 12)
    accepted = 0
    cs = cases(rng, 20000, near)
    for cur, m, n_iso in cs:
        got = closed_form(cur, m, n_iso)
        if got >= 0:
            accepted += 1
            assert got == scan(cur, m, n_iso), (cur, m, n_iso)
    # random draws almost never need the scan; the near-crossing draws are
    # built to need it often
    if near:
        assert 0 < accepted < len(cs)
    else:
        assert accepted > 0.999 * len(cs)


def test_a_smaller_bound_would_be_wrong():
    """The bound is needed: accepting every crossing (B = 0) disagrees with
    the scan on some near-crossing inputs."""
    rng = np.random.default_rng(13)
    wrong = 0
    for cur, m, n_iso in cases(rng, 20000, True):
        m0, m1, m2 = (list(m) + [0.0, 0.0])[:3]
        M1 = m0 + m1
        S = M1 + m2
        kq = float(np.floor(cur / S))
        if not kq < n_iso:
            continue
        rem = float(Fraction(cur) - Fraction(kq) * Fraction(S))
        r = 0 if rem < m0 else (1 if rem < M1 else 2)
        if r < len(m) and r != scan(cur, m, n_iso):
            wrong += 1
    assert wrong > 0

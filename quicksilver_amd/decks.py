"""Seeded generators for Quicksilver input decks (the ``qs`` deck format).

The deck format is the reference's: a ``Block:`` line followed by indented
``key: value`` lines (Parameters.cc:280-339, parseUtils.cc:17-111).  Blocks are
``Simulation``, ``Geometry`` (ordered; the last geometry that contains a cell
centre wins, MC_Domain.cc:440-449), ``Material`` and ``CrossSection``.

Workloads (all FP64, 230 energy groups from 1e-9 to 20 MeV, dt = 1e-8 s,
1 cm cells, reflecting boundaries, one spatial domain):

* ``small_cube``   10^3 cells, one material / one isotope, flat cross sections
                   (total 1.0, scatter 0.6, absorb 0.3, fission 0.1,
                   nuBar 2), 1e5 particles, 10 steps.
* ``nested_boxes`` 40^3 cells, three materials in nested boxes, 10 isotopes x
                   3 reactions each, seeded degree-2/3 polynomial cross
                   sections, 1e6 particles, 10 steps.
* ``checkerboard`` 80^3 cells, two materials in a seeded checkerboard of 8^3
                   blocks, 20 isotopes x 3 reactions, 4e7 particles, 20 steps.
* ``mixed_batch``  64 problems: 1e4..2e6 particles (log-uniform), edge
                   8..48 cells (log-uniform, capped so every cell sources
                   particles), 1-3 nested-box materials with 5-20 isotopes,
                   10 steps.
* ``serve_stream`` Poisson arrivals of problems drawn like ``mixed_batch``
                   with 5..20 steps.

Reference behaviour that these decks inherit (stated, not changed): the source
samples energies uniformly in [eMin, eMax] (MC_SourceNow.cc:104-105); all
isotopes of a material share one polynomial per reaction type
(NuclearData.hh:147-148); ``nParticles`` is the population target and 10% of
it is sourced per step (MC_SourceNow.cc:60-61).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

N_GROUPS = 230
E_MIN = 1e-9
E_MAX = 20.0


def _fmt(v) -> str:
    if isinstance(v, float):
        return repr(v)
    return str(v)


@dataclass
class Deck:
    """An input deck; ``text()`` renders it in the reference format."""
    simulation: Dict[str, object]
    geometries: List[Dict[str, object]] = field(default_factory=list)
    materials: List[Dict[str, object]] = field(default_factory=list)
    cross_sections: List[Dict[str, object]] = field(default_factory=list)
    name: str = "deck"

    def text(self) -> str:
        out = ["Simulation:"]
        out += [f"   {k}: {_fmt(v)}" for k, v in self.simulation.items()]
        out.append("")
        for g in self.geometries:
            out.append("Geometry:")
            out += [f"   {k}: {_fmt(v)}" for k, v in g.items()]
            out.append("")
        for m in self.materials:
            out.append("Material:")
            out += [f"   {k}: {_fmt(v)}" for k, v in m.items()]
            out.append("")
        for c in self.cross_sections:
            out.append("CrossSection:")
            for k, v in c.items():
                if k == "D" and self.name == "nonflat" and c.get("name") == "absorb":
                    out.append("   D: -0.5243")  # overridden by the next line
                out.append(f"   {k}: {_fmt(v)}")
            out.append("")
        return "\n".join(out) + "\n"

    def write(self, path: str) -> str:
        with open(path, "w") as f:
            f.write(self.text())
        return path

    @property
    def n_cells(self) -> int:
        s = self.simulation
        return int(s["nx"]) * int(s["ny"]) * int(s["nz"])


def _simulation(n: int, n_particles: int, n_steps: int, bc: str = "reflect", **extra):
    sim = {
        "dt": 1e-08,
        "boundaryCondition": bc,
        "lx": float(n), "ly": float(n), "lz": float(n),
        "nParticles": int(n_particles),
        "nSteps": int(n_steps),
        "nx": int(n), "ny": int(n), "nz": int(n),
        "seed": 1,
        "xDom": 1, "yDom": 1, "zDom": 1,
        "eMax": E_MAX,
        "eMin": E_MIN,
        "nGroups": N_GROUPS,
        "lowWeightCutoff": 0.001,
    }
    sim.update(extra)
    return sim


def _brick(material: str, lo: Sequence[float], hi: Sequence[float]):
    return {"material": material, "shape": "brick",
            "xMax": float(hi[0]), "xMin": float(lo[0]),
            "yMax": float(hi[1]), "yMin": float(lo[1]),
            "zMax": float(hi[2]), "zMin": float(lo[2])}


def small_cube(n_particles: int = 100000, n_steps: int = 10, n: int = 10) -> Deck:
    """10^3-cell cube, one flat-XS material (the CPU-runnable workload)."""
    sim = _simulation(n, n_particles, n_steps)
    mat = {"name": "sourceMaterial", "mass": 1000.0, "nIsotopes": 1, "nReactions": 3,
           "sourceRate": 1e10, "totalCrossSection": 1.0,
           "absorptionCrossSection": "flat", "fissionCrossSection": "flat",
           "scatteringCrossSection": "flat",
           "absorptionCrossSectionRatio": 0.3, "fissionCrossSectionRatio": 0.1,
           "scatteringCrossSectionRatio": 0.6}
    xs = {"name": "flat", "A": 0.0, "B": 0.0, "C": 0.0, "D": 0.0, "E": 1.0, "nuBar": 2.0}
    return Deck(sim, [_brick("sourceMaterial", (0, 0, 0), (n, n, n))], [mat], [xs], "small_cube")


def _flat_material(name, total, absorb, fission, scatter, source_rate, n_iso=10, n_react=9):
    return {"name": name, "nIsotopes": n_iso, "nReactions": n_react, "sourceRate": source_rate,
            "totalCrossSection": total,
            "absorptionCrossSection": "flat", "fissionCrossSection": "flat",
            "scatteringCrossSection": "flat",
            "absorptionCrossSectionRatio": absorb, "fissionCrossSectionRatio": fission,
            "scatteringCrossSectionRatio": scatter}


_FLAT_24 = {"name": "flat", "A": 0.0, "B": 0.0, "C": 0.0, "D": 0.0, "E": 1.0, "nuBar": 2.4}


def limit_case(kind: str, n_particles: int = 9999, n_steps: int = 5) -> Deck:
    """Small restatements of the reference's scenario decks (Examples/):
    ``all_absorb`` (every source particle absorbed at its first collision),
    ``all_escape`` (no collisions, escape boundary), ``no_collisions``
    (streaming only), ``scatter_only`` (octant boundary, scatter only),
    ``no_fission`` (octant boundary, scatter + absorb)."""
    if kind in ("all_absorb", "all_escape"):
        absorb = kind == "all_absorb"
        sim = _simulation(10, n_particles, n_steps, bc="reflect" if absorb else "escape", dt=1e6)
        xs = 1e10 if absorb else 1e-20
        ratios = (1.0, 0.0, 0.0) if absorb else (0.0, 0.0, 1.0)
        mats = [_flat_material("boxMaterial", xs, *ratios, 0.0),
                _flat_material("sourceMaterial", xs, *ratios, 1e-2)]
        geo = [_brick("boxMaterial", (0, 0, 0), (10, 10, 10)),
               _brick("sourceMaterial", (0, 0, 0), (1, 1, 1))]
        return Deck(sim, geo, mats, [dict(_FLAT_24)], kind)
    if kind == "no_collisions":
        sim = _simulation(10, n_particles, n_steps, lx=100.0, ly=100.0, lz=100.0)
        mats = [_flat_material("boxMaterial", 1e-80, 1.0, 0.0, 1.0, 0.0),
                _flat_material("sourceMaterial", 1e-80, 1.0, 1.0, 1.0, 1e10)]
        geo = [_brick("boxMaterial", (0, 0, 0), (100, 100, 100)),
               _brick("sourceMaterial", (0, 0, 0), (10, 10, 10))]
        return Deck(sim, geo, mats, [dict(_FLAT_24)], kind)
    if kind == "scatter_only":
        sim = _simulation(10, n_particles, n_steps, bc="octant", loadBalance=1,
                          lx=100.0, ly=100.0, lz=100.0)
        mats = [_flat_material("sourceMaterial", 0.1, 0.0, 0.0, 1.0, 1e10)]
        return Deck(sim, [_brick("sourceMaterial", (0, 0, 0), (100, 100, 100))], mats,
                    [dict(_FLAT_24)], kind)
    if kind == "no_fission":
        sim = _simulation(10, n_particles, n_steps, bc="octant", lx=100.0, ly=100.0, lz=100.0)
        mats = [_flat_material("sourceMaterial", 0.198, 0.494949495, 0.0, 0.505050505, 1e10)]
        return Deck(sim, [_brick("sourceMaterial", (0, 0, 0), (100, 100, 100))], mats,
                    [dict(_FLAT_24)], kind)
    raise ValueError(kind)


def nonflat(n_particles: int = 20000, n_steps: int = 4, dt: float = 1e-10) -> Deck:
    """The reference's NonFlatXC scenario (Examples/NonFlatXC/NonFlatXC.inp):
    9 reactions, polynomial absorb/fission/scatter tables, an unused second
    material, and a CrossSection block with a duplicated key (last wins).
    The scenario is strongly supercritical above 1 MeV (fission multiplies
    the population many times within a 1e-8 s step), so the default step is
    1e-10 s to keep the work per step bounded."""
    sim = _simulation(10, n_particles, n_steps, lx=100.0, ly=100.0, lz=100.0, eMin=1e-08, dt=dt)
    mats = [{"name": "sourceMaterial", "mass": 1000.0, "nIsotopes": 10, "nReactions": 9,
             "sourceRate": 1e10, "totalCrossSection": 6.0,
             "absorptionCrossSection": "absorb", "fissionCrossSection": "fission",
             "scatteringCrossSection": "scatter",
             "absorptionCrossSectionRatio": 6e-3, "fissionCrossSectionRatio": 1.0,
             "scatteringCrossSectionRatio": 5.0},
            _flat_material("flatMaterial", 1.0, 1.0, 1.0, 1.0, 1e10, n_iso=20)]
    xs = [dict(_FLAT_24),
          {"name": "absorb", "A": 0.0, "B": 0.0, "C": 0.0, "D": -0.8446, "E": -2.22},
          {"name": "fission", "A": 0.0, "B": 0.0, "C": 0.0, "D": -0.342, "E": 0.0, "nuBar": 2.4},
          {"name": "scatter", "A": 0.0, "B": 0.0, "C": 0.0, "D": 0.0, "E": 0.7}]
    deck = Deck(sim, [_brick("sourceMaterial", (0, 0, 0), (100, 100, 100))], mats, xs, "nonflat")
    return deck


def spheres(n_particles: int = 20000, n_steps: int = 4, n: int = 12) -> Deck:
    """Sphere geometries (Parameters.cc:451-456, MC_Domain.cc:416-422): a
    reflecting cube of flat material with two overlapping spheres of
    different materials (the later sphere wins where they overlap)."""
    sim = _simulation(n, n_particles, n_steps)
    mats = [_flat_material("background", 0.5, 0.3, 0.02, 0.68, 1e10, n_iso=2, n_react=3),
            _flat_material("shellA", 2.0, 0.6, 0.02, 0.38, 1e10, n_iso=3, n_react=4),
            _flat_material("shellB", 1.0, 0.5, 0.04, 0.46, 1e10, n_iso=4, n_react=5)]
    geo = [_brick("background", (0, 0, 0), (n, n, n)),
           {"material": "shellA", "shape": "sphere", "radius": n / 3.0,
            "xCenter": n / 2.0, "yCenter": n / 2.0, "zCenter": n / 2.0},
           {"material": "shellB", "shape": "sphere", "radius": n / 4.0,
            "xCenter": n / 3.0, "yCenter": n / 2.0, "zCenter": n / 2.0}]
    xs = [{"name": "flat", "A": 0.0, "B": 0.0, "C": 0.0, "D": 0.0, "E": 1.0, "nuBar": 2.0}]
    return Deck(sim, geo, mats, xs, "spheres")


# --- seeded polynomial cross sections -------------------------------------

def _group_mid_log10() -> np.ndarray:
    # NuclearData.cc:105-119 energy grid; group midpoints as in NuclearData.cc:21-24
    lo, hi = math.log(E_MIN), math.log(E_MAX)
    delta = (hi - lo) / (N_GROUPS + 1.0)
    e = np.empty(N_GROUPS + 1)
    e[0] = E_MIN
    e[N_GROUPS] = E_MAX
    for i in range(1, N_GROUPS):
        e[i] = math.exp(lo + delta * i)
    mid = (e[:-1] + e[1:]) / 2.0
    norm_group = int(np.argmax(e[1:] >= 1.0))
    return np.log10(mid), norm_group


_LOGE, _NORM_G = _group_mid_log10()


def _poly(c, x):
    a, b, cc, d, e = c
    return e + x * (d + x * (cc + x * (b + x * a)))


def _shape(c) -> np.ndarray:
    v = 10.0 ** _poly(c, _LOGE)
    return v / v[_NORM_G]


def _material_xs(rng: np.random.Generator, name: str, n_iso: int, lo: float = 0.1, hi: float = 10.0,
                 k_max: float = 0.0):
    """Draw one material's three reaction polynomials, ratios, total and nuBar,
    rejecting draws whose per-group macroscopic total leaves [lo, hi] and,
    when k_max > 0, draws whose particle multiplication per absorbing
    collision, nuBar * sigma_f / (sigma_a + sigma_f), exceeds k_max in any
    group (so a population cannot grow without bound within a time step)."""
    for _ in range(10000):
        coeffs = {}
        for kind in ("scatter", "absorb", "fission"):
            cubic = rng.random() < 0.5
            coeffs[kind] = (0.0,
                            float(rng.uniform(-0.002, 0.002)) if cubic else 0.0,
                            float(rng.uniform(-0.02, 0.02)),
                            float(rng.uniform(-0.2, 0.2)),
                            float(rng.uniform(0.0, 1.0)))
        total = float(math.exp(rng.uniform(math.log(0.3), math.log(3.0))))
        ratios = {k: float(rng.uniform(0.1, 1.0)) for k in ("scatter", "absorb", "fission")}
        nubar = float(rng.uniform(1.5, 2.5))
        wsum = sum(ratios.values())
        tot = sum(total * ratios[k] / wsum * _shape(coeffs[k]) for k in coeffs)
        if k_max > 0:
            sa = ratios["absorb"] * _shape(coeffs["absorb"])
            sf = ratios["fission"] * _shape(coeffs["fission"])
            if np.max(nubar * sf / np.maximum(sa + sf, 1e-300)) > k_max:
                continue
        if tot.min() >= lo and tot.max() <= hi:
            break
    else:  # pragma: no cover
        raise RuntimeError("cross-section rejection sampling failed")
    xs = []
    for kind, c in coeffs.items():
        d = {"name": f"{name}_{kind}", "A": c[0], "B": c[1], "C": c[2], "D": c[3], "E": c[4]}
        if kind == "fission":
            d["nuBar"] = nubar
        xs.append(d)
    mat = {"name": name, "mass": 1000.0, "nIsotopes": int(n_iso), "nReactions": 3,
           "sourceRate": 1e10, "totalCrossSection": total,
           "absorptionCrossSection": f"{name}_absorb",
           "fissionCrossSection": f"{name}_fission",
           "scatteringCrossSection": f"{name}_scatter",
           "absorptionCrossSectionRatio": ratios["absorb"],
           "fissionCrossSectionRatio": ratios["fission"],
           "scatteringCrossSectionRatio": ratios["scatter"]}
    return mat, xs


def nested_boxes(n_particles: int = 1000000, n_steps: int = 10, seed: int = 2,
                 n: int = 40, n_iso: int = 10) -> Deck:
    """40^3 mesh, materials A (whole box) / B ([n/4,3n/4]^3) / C ([3n/8,5n/8]^3)."""
    rng = np.random.default_rng(seed)
    sim = _simulation(n, n_particles, n_steps)
    mats, xss = [], []
    for name in ("matA", "matB", "matC"):
        m, x = _material_xs(rng, name, n_iso)
        mats.append(m)
        xss += x
    geo = [_brick("matA", (0, 0, 0), (n, n, n)),
           _brick("matB", (n / 4, n / 4, n / 4), (3 * n / 4, 3 * n / 4, 3 * n / 4)),
           _brick("matC", (3 * n / 8, 3 * n / 8, 3 * n / 8), (5 * n / 8, 5 * n / 8, 5 * n / 8))]
    return Deck(sim, geo, mats, xss, "nested_boxes")


def checkerboard(n_particles: int = 40000000, n_steps: int = 20, seed: int = 3,
                 n: int = 80, blocks: int = 8, n_iso: int = 20) -> Deck:
    """80^3 mesh; background matA plus matB on a seeded coin-flip subset of the
    8^3 blocks (one Geometry brick per chosen block)."""
    rng = np.random.default_rng(seed)
    sim = _simulation(n, n_particles, n_steps)
    mats, xss = [], []
    for name in ("matA", "matB"):
        m, x = _material_xs(rng, name, n_iso)
        mats.append(m)
        xss += x
    geo = [_brick("matA", (0, 0, 0), (n, n, n))]
    b = n / blocks
    for k in range(blocks):
        for j in range(blocks):
            for i in range(blocks):
                if rng.random() < 0.5:
                    # keep the brick strictly inside the block so cell centres
                    # on the block faces are unambiguous
                    geo.append(_brick("matB", (i * b + 0.25, j * b + 0.25, k * b + 0.25),
                                      ((i + 1) * b - 0.25, (j + 1) * b - 0.25, (k + 1) * b - 0.25)))
    return Deck(sim, geo, mats, xss, "checkerboard")


def _random_problem(rng: np.random.Generator, n_steps: int, tag: str) -> Deck:
    # The reference sources (int)(cell weight / source weight) particles per
    # cell (MC_SourceNow.cc:72-76), so a mesh with more cells than
    # nParticles/10 sources nothing.  The edge is therefore drawn
    # log-uniformly between 8 and the largest edge (<= 48) whose cells each
    # source at least 1.5 particles per step.
    n_particles = int(round(math.exp(rng.uniform(math.log(1e4), math.log(2e6)))))
    n_max = max(8.0, min(48.0, (n_particles / 15.0) ** (1.0 / 3.0)))
    n = int(math.floor(math.exp(rng.uniform(math.log(8), math.log(n_max + 1e-9)))))
    n = max(8, min(n, int(n_max)))
    n_mat = int(rng.integers(1, 4))
    sim = _simulation(n, n_particles, n_steps)
    mats, xss, geo = [], [], []
    names = ["matA", "matB", "matC"][:n_mat]
    for idx, name in enumerate(names):
        m, x = _material_xs(rng, name, int(rng.integers(5, 21)), k_max=1.0)
        mats.append(m)
        xss += x
        lo = n * idx / (2.0 * n_mat + 1)
        geo.append(_brick(name, (lo, lo, lo), (n - lo, n - lo, n - lo)))
    return Deck(sim, geo, mats, xss, tag)


def mixed_batch(count: int = 64, seed: int = 4, n_steps: int = 10) -> List[Deck]:
    rng = np.random.default_rng(seed)
    return [_random_problem(rng, n_steps, f"mixed_{i:02d}") for i in range(count)]


@dataclass
class Request:
    arrival_s: float
    deck: Deck


def serve_stream(count: int = 256, seed: int = 5, rate_per_s: float = 1.0) -> List[Request]:
    """Poisson arrivals (exponential gaps at ``rate_per_s``) of problems drawn
    like ``mixed_batch`` with 5..20 steps."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate_per_s))
        steps = int(rng.integers(5, 21))
        out.append(Request(t, _random_problem(rng, steps, f"req_{i:03d}")))
    return out


def scaled(deck: Deck, n_particles: Optional[int] = None, n_steps: Optional[int] = None) -> Deck:
    """Copy of ``deck`` with a different population target / step count."""
    sim = dict(deck.simulation)
    if n_particles is not None:
        sim["nParticles"] = int(n_particles)
    if n_steps is not None:
        sim["nSteps"] = int(n_steps)
    return Deck(sim, list(deck.geometries), list(deck.materials), list(deck.cross_sections), deck.name)


def write_all(outdir: str) -> List[str]:  # pragma: no cover - convenience
    os.makedirs(outdir, exist_ok=True)
    paths = [small_cube().write(os.path.join(outdir, "small_cube.inp")),
             nested_boxes().write(os.path.join(outdir, "nested_boxes.inp")),
             checkerboard().write(os.path.join(outdir, "checkerboard.inp"))]
    for d in mixed_batch():
        paths.append(d.write(os.path.join(outdir, d.name + ".inp")))
    return paths

"""MI355X (gfx950) particle-tracking path of Quicksilver, behind a C ABI.

The package is a thin ctypes view of ``libqsgpu.so`` (include/qsg.h).  The
names mirror the reference driver: ``cycle_init`` / ``cycle_tracking`` /
``cycle_finalize`` (main.cc:96-324), ``run`` for a whole deck.  There is no
CPU fallback: if the library is missing or no GPU is visible, calls raise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqsgpu.so")
QS_EXE = os.path.join(_HERE, "qs")

QSG_OK = 0
ERRORS = {1: "ARG", 2: "DEVICE", 3: "CAPACITY", 4: "STATE", 5: "NUMERIC"}

BALANCE_FIELDS = ["start", "source", "rr", "split", "absorb", "scatter", "fission", "produce",
                  "collision", "escape", "census", "num_segments", "end"]

#: rays per qsg_nearest_facet_probe_noalloc call (QSG_PROBE_NOALLOC_MAX_RAYS)
NOALLOC_MAX_RAYS = 65536

#: every entry point declared in include/qsg.h
EXPORTS = ["qsg_last_error", "qsg_create", "qsg_create_from_deck", "qsg_create_from_deck_opts",
           "qsg_problem_error", "qsg_destroy", "qsg_get_info",
           "qsg_set_tuning", "qsg_get_tuning", "qsg_cycle_init", "qsg_cycle_begin_host", "qsg_cycle_tracking",
           "qsg_cycle_finalize", "qsg_run", "qsg_get_scalar_flux", "qsg_set_particles",
           "qsg_get_census_count", "qsg_get_census", "qsg_get_processing_count",
           "qsg_get_processing", "qsg_math_probe", "qsg_get_stats", "qsg_reset_stats",
           "qsg_batch_create", "qsg_batch_destroy", "qsg_batch_step", "qsg_batch_problem_error",
           "qsg_batch_set_active",
           "qsg_batch_set_tuning", "qsg_batch_get_stats", "qsg_server_create", "qsg_server_destroy",
           "qsg_server_submit", "qsg_server_submit_opts", "qsg_server_step", "qsg_server_status",
           "qsg_server_result", "qsg_server_cancel", "qsg_server_release", "qsg_server_set_tuning",
           "qsg_server_get_stats",
           "qsg_coral_check", "qsg_get_fluence", "qsg_get_energy_spectrum", "qsg_nearest_facet_probe",
           "qsg_nearest_facet_probe_noalloc"]


class QsgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"qsg error {ERRORS.get(code, code)}: {msg}")
        self.code = code


class Balance(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in BALANCE_FIELDS]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f in BALANCE_FIELDS}


class Info(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("n_groups", ctypes.c_int32), ("n_materials", ctypes.c_int32),
                ("n_steps", ctypes.c_int32), ("n_cells", ctypes.c_uint64),
                ("n_particles_target", ctypes.c_uint64), ("source_per_cycle", ctypes.c_uint64),
                ("particle_capacity", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("dt", ctypes.c_double), ("source_weight", ctypes.c_double),
                ("cycle", ctypes.c_int32), ("lds_tables", ctypes.c_int32)]


class CoralReport(ctypes.Structure):
    """qsg_coral_report: the reference's CORAL2 correctness tests."""
    _fields_ = [("benchmark", ctypes.c_int32), ("ratio_pass", ctypes.c_int32), ("event_pass", ctypes.c_int32),
                ("missing_pass", ctypes.c_int32), ("fluence_pass", ctypes.c_int32),
                ("ratio_error", ctypes.c_double), ("ratio_tolerance", ctypes.c_double),
                ("event_ratio", ctypes.c_double), ("gains", ctypes.c_uint64), ("losses", ctypes.c_uint64),
                ("fluence_max_percent_diff", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}

    def verdicts(self) -> List[str]:
        """The PASS/FAIL lines the reference prints (CoralBenchmark.cc)."""
        tol = f"{self.ratio_tolerance * 100.0:g}"
        return [
            f"PASS:: Absorption / Fission / Scatter Ratios maintained with {tol}% tolerance" if self.ratio_pass
            else f"FAIL:: Absorption / Fission / Scatter Ratios NOT maintained with {tol}% tolerance",
            "PASS:: Collision to Facet Crossing Ratio maintained even balanced within 1% tolerance"
            if self.event_pass else "FAIL:: Collision to Facet Crossing Ratio balanced NOT maintained within 1% tolerance",
            "PASS:: No Particles Lost During Run" if self.missing_pass
            else "FAIL:: Particles Were Lost During Run, test for done should have failed",
            "PASS:: Fluence is homogenous across cells with 6% tolerance" if self.fluence_pass
            else "FAIL:: Fluence not homogenous across cells within 6% tolerance",
        ]


class Options(ctypes.Structure):
    """qsg_options: per-problem sizing (0 = the library's default)."""
    _fields_ = [("particle_capacity", ctypes.c_uint64)]


class Stats(ctypes.Structure):
    _fields_ = [("track_kernel_seconds", ctypes.c_double), ("tracking_seconds", ctypes.c_double),
                ("init_seconds", ctypes.c_double), ("finalize_seconds", ctypes.c_double),
                ("track_launches", ctypes.c_uint64), ("segments", ctypes.c_uint64),
                ("collisions", ctypes.c_uint64), ("facet_crossings", ctypes.c_uint64),
                ("particle_loads", ctypes.c_uint64), ("cycles", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("submitted", "admitted", "completed", "rejected", "steps",
                                               "segments", "bytes_in_use", "active", "queued")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_int32), ("waves_per_cu", ctypes.c_int32),
                ("refill_min", ctypes.c_int32), ("chunk", ctypes.c_int32),
                ("stack_entries", ctypes.c_int32), ("sort_interval", ctypes.c_int32),
                ("waves_per_simd", ctypes.c_int32), ("graphs", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


# Launch shape used when none is set (the library's own default).
DEFAULT_TUNING = {"block_size": 256, "waves_per_cu": 12, "refill_min": 1, "chunk": 32,
                  "stack_entries": 512, "sort_interval": 0, "waves_per_simd": 3,
                  "graphs": 1}


def _tuning(kw: dict) -> Tuning:
    unknown = set(kw) - set(DEFAULT_TUNING)
    if unknown:
        raise TypeError(f"unknown tuning knobs: {sorted(unknown)}")
    return Tuning(**{**DEFAULT_TUNING, **kw})


_P = ctypes.POINTER
_D = _P(ctypes.c_double)


class ParticlesC(ctypes.Structure):
    _fields_ = [("x", _D), ("y", _D), ("z", _D), ("vx", _D), ("vy", _D), ("vz", _D),
                ("energy", _D), ("weight", _D), ("time_to_census", _D),
                ("num_mean_free_paths", _D), ("seed", _P(ctypes.c_uint64)),
                ("cell", _P(ctypes.c_int32))]


PARTICLE_FLOAT_FIELDS = ["x", "y", "z", "vx", "vy", "vz", "energy", "weight", "time_to_census",
                         "num_mean_free_paths"]


@dataclass
class ParticleSet:
    """Host copy of particle state (structure of arrays)."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    energy: np.ndarray
    weight: np.ndarray
    time_to_census: np.ndarray
    num_mean_free_paths: np.ndarray
    seed: np.ndarray
    cell: np.ndarray

    @classmethod
    def empty(cls, n: int) -> "ParticleSet":
        f = {k: np.zeros(n, dtype=np.float64) for k in PARTICLE_FLOAT_FIELDS}
        return cls(**f, seed=np.zeros(n, dtype=np.uint64), cell=np.zeros(n, dtype=np.int32))

    def __len__(self):
        return len(self.seed)

    def c_view(self) -> ParticlesC:
        for k in PARTICLE_FLOAT_FIELDS:
            a = getattr(self, k)
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                setattr(self, k, np.ascontiguousarray(a, dtype=np.float64))
        self.seed = np.ascontiguousarray(self.seed, dtype=np.uint64)
        self.cell = np.ascontiguousarray(self.cell, dtype=np.int32)
        args = {k: getattr(self, k).ctypes.data_as(_D) for k in PARTICLE_FLOAT_FIELDS}
        return ParticlesC(**args, seed=self.seed.ctypes.data_as(_P(ctypes.c_uint64)),
                          cell=self.cell.ctypes.data_as(_P(ctypes.c_int32)))

    def state(self) -> np.ndarray:
        """[n, 10] float64: x y z vx vy vz E w tc nmfp."""
        return np.stack([getattr(self, k) for k in PARTICLE_FLOAT_FIELDS], axis=1)


_lib = None


def lib() -> ctypes.CDLL:
    """Load libqsgpu.so (built by ``make -C quicksilver_amd``); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    i, c, vp = ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p
    argv_t = _P(c)
    L.qsg_last_error.restype = c
    L.qsg_create.argtypes = [i, argv_t, i, _P(vp)]
    L.qsg_create_from_deck.argtypes = [c, i, argv_t, i, _P(vp)]
    L.qsg_create_from_deck_opts.argtypes = [c, i, argv_t, i, _P(Options), _P(vp)]
    L.qsg_problem_error.argtypes = [vp, _P(i), c, ctypes.c_size_t]
    L.qsg_destroy.argtypes = [vp]
    L.qsg_get_info.argtypes = [vp, _P(Info)]
    L.qsg_set_tuning.argtypes = [vp, _P(Tuning)]
    L.qsg_get_tuning.argtypes = [vp, _P(Tuning)]
    L.qsg_cycle_init.argtypes = [vp]
    L.qsg_cycle_begin_host.argtypes = [vp, _P(ParticlesC), ctypes.c_uint64]
    L.qsg_cycle_tracking.argtypes = [vp]
    L.qsg_cycle_finalize.argtypes = [vp, _P(Balance), _D]
    L.qsg_run.argtypes = [vp, i, _P(Balance), _D, _D]
    L.qsg_get_scalar_flux.argtypes = [vp, _D, ctypes.c_uint64]
    L.qsg_set_particles.argtypes = [vp, _P(ParticlesC), ctypes.c_uint64]
    L.qsg_get_census_count.argtypes = [vp, _P(ctypes.c_uint64)]
    L.qsg_get_census.argtypes = [vp, _P(ParticlesC), ctypes.c_uint64]
    L.qsg_get_processing_count.argtypes = [vp, _P(ctypes.c_uint64)]
    L.qsg_get_processing.argtypes = [vp, _P(ParticlesC), ctypes.c_uint64]
    L.qsg_math_probe.argtypes = [i, _D, _D, _D, ctypes.c_uint64, i]
    _I32 = _P(ctypes.c_int32)
    L.qsg_nearest_facet_probe.argtypes = [vp, ctypes.c_uint64, _D, _D, _I32, _D, _I32, _D]
    L.qsg_nearest_facet_probe_noalloc.argtypes = [vp, ctypes.c_uint64, _D, _D, _I32, _D, _I32, _D]
    L.qsg_get_energy_spectrum.argtypes = [vp, _D, _P(ctypes.c_uint64), ctypes.c_uint64]
    L.qsg_get_fluence.argtypes = [vp, _D, ctypes.c_uint64]
    L.qsg_get_stats.argtypes = [vp, _P(Stats)]
    L.qsg_reset_stats.argtypes = [vp]
    L.qsg_batch_create.argtypes = [_P(vp), i, _P(vp)]
    L.qsg_batch_destroy.argtypes = [vp]
    L.qsg_batch_step.argtypes = [vp, _P(Balance), _D]
    L.qsg_batch_problem_error.argtypes = [vp, i, _P(i)]
    L.qsg_batch_set_active.argtypes = [vp, i, i]
    L.qsg_batch_set_tuning.argtypes = [vp, _P(Tuning)]
    L.qsg_batch_get_stats.argtypes = [vp, _P(Stats)]
    u64 = ctypes.c_uint64
    L.qsg_server_create.argtypes = [i, u64, i, _P(vp)]
    L.qsg_server_destroy.argtypes = [vp]
    L.qsg_server_submit.argtypes = [vp, c, _P(u64)]
    L.qsg_server_submit_opts.argtypes = [vp, c, _P(Options), _P(u64)]
    L.qsg_server_cancel.argtypes = [vp, u64]
    L.qsg_server_step.argtypes = [vp, _P(i), _P(i)]
    L.qsg_server_status.argtypes = [vp, u64, _P(i), _P(i)]
    L.qsg_server_result.argtypes = [vp, u64, _P(Balance), _D, i]
    L.qsg_server_release.argtypes = [vp, u64]
    L.qsg_server_set_tuning.argtypes = [vp, _P(Tuning)]
    L.qsg_server_get_stats.argtypes = [vp, _P(ServerStats)]
    _lib = L
    return L


def _check(rc: int):
    if rc != QSG_OK:
        raise QsgError(rc, lib().qsg_last_error().decode())


def _argv(args: Sequence[str]):
    arr = (ctypes.c_char_p * (len(args) + 1))()
    for k, a in enumerate(args):
        arr[k] = a.encode()
    arr[len(args)] = None
    return arr


class Problem:
    """One transport problem resident on a GPU (a qsg_problem handle)."""

    def __init__(self, deck_text: Optional[str] = None, args: Sequence[str] = (), device: int = 0,
                 particle_capacity: int = 0):
        """particle_capacity: entries of each particle buffer (0: the
        library's default, 2 x nParticles + one cycle's source + 65536)."""
        L = lib()
        h = ctypes.c_void_p()
        argv = ["qs", *args]
        if deck_text is None:
            if particle_capacity:
                raise ValueError("particle_capacity needs the deck text")
            _check(L.qsg_create(len(argv), _argv(argv), device, ctypes.byref(h)))
        else:
            opts = Options(particle_capacity=int(particle_capacity))
            _check(L.qsg_create_from_deck_opts(deck_text.encode(), len(argv), _argv(argv), device,
                                               ctypes.byref(opts), ctypes.byref(h)))
        self._h = h

    @classmethod
    def from_file(cls, path: str, device: int = 0, extra_args: Sequence[str] = ()):
        return cls(None, ["-i", path, *extra_args], device)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().qsg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- introspection / tuning
    def info(self) -> Info:
        inf = Info()
        _check(lib().qsg_get_info(self._h, ctypes.byref(inf)))
        return inf

    def tuning(self) -> Tuning:
        t = Tuning()
        _check(lib().qsg_get_tuning(self._h, ctypes.byref(t)))
        return t

    def set_tuning(self, **knobs):
        """Launch shape; knobs not given take DEFAULT_TUNING's values."""
        t = _tuning(knobs)
        _check(lib().qsg_set_tuning(self._h, ctypes.byref(t)))

    def stats(self) -> dict:
        st = Stats()
        _check(lib().qsg_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def reset_stats(self):
        _check(lib().qsg_reset_stats(self._h))

    def error(self):
        """(code, message) the problem failed with; code 0 if it did not."""
        code = ctypes.c_int()
        msg = ctypes.create_string_buffer(1024)
        _check(lib().qsg_problem_error(self._h, ctypes.byref(code), msg, len(msg)))
        return code.value, msg.value.decode()

    # ---- per-cycle phases (main.cc:55-66)
    def cycle_init(self):
        _check(lib().qsg_cycle_init(self._h))

    def cycle_begin_host(self, ps: "ParticleSet"):
        """In place of cycle_init: ``ps`` (from the caller's own source and
        population control) becomes this cycle's processing population."""
        v = ps.c_view()
        _check(lib().qsg_cycle_begin_host(self._h, ctypes.byref(v), len(ps)))

    def cycle_tracking(self):
        _check(lib().qsg_cycle_tracking(self._h))

    def cycle_finalize(self):
        b = Balance()
        fs = ctypes.c_double()
        _check(lib().qsg_cycle_finalize(self._h, ctypes.byref(b), ctypes.byref(fs)))
        return b.as_dict(), fs.value

    def run(self, n_cycles: Optional[int] = None):
        """Runs n cycles (default: the deck's nSteps); returns a list of
        (balance dict, flux sum, tracking seconds)."""
        n = self.info().n_steps if n_cycles is None else n_cycles
        bal = (Balance * max(n, 1))()
        flux = (ctypes.c_double * max(n, 1))()
        secs = (ctypes.c_double * max(n, 1))()
        _check(lib().qsg_run(self._h, n, bal, flux, secs))
        return [(bal[k].as_dict(), flux[k], secs[k]) for k in range(n)]

    def scalar_flux(self) -> np.ndarray:
        inf = self.info()
        out = np.zeros((inf.n_cells, inf.n_groups), dtype=np.float64)
        _check(lib().qsg_get_scalar_flux(self._h, out.ctypes.data_as(_D), out.size))
        return out

    def energy_spectrum(self):
        """(group energies [G+1], census counts [G+1]) accumulated over the
        cycles run so far (runs that set energySpectrum)."""
        n = self.info().n_groups + 1
        e = np.zeros(n, dtype=np.float64)
        c = np.zeros(n, dtype=np.uint64)
        _check(lib().qsg_get_energy_spectrum(self._h, e.ctypes.data_as(_D), c.ctypes.data_as(_P(ctypes.c_uint64)), n))
        return e, c

    def fluence(self) -> np.ndarray:
        """Cumulative per-cell fluence (decks with coralBenchmark set)."""
        out = np.zeros(self.info().n_cells, dtype=np.float64)
        _check(lib().qsg_get_fluence(self._h, out.ctypes.data_as(_D), out.size))
        return out

    def nearest_facet(self, xyz: np.ndarray, direction: np.ndarray, cell: np.ndarray, noalloc: bool = False):
        """The tracking kernels' nearest-facet search on given rays (see
        qsg_nearest_facet_probe): returns (distance, facet, position after
        any nudge).  noalloc: through qsg_nearest_facet_probe_noalloc (no
        device allocation; at most NOALLOC_MAX_RAYS rays)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        direction = np.ascontiguousarray(direction, dtype=np.float64).reshape(-1, 3)
        cell = np.ascontiguousarray(cell, dtype=np.int32).reshape(-1)
        n = cell.size
        if xyz.shape[0] != n or direction.shape[0] != n:
            raise ValueError("xyz, direction and cell must describe the same number of rays")
        dist = np.empty(n, np.float64)
        facet = np.empty(n, np.int32)
        moved = np.empty((n, 3), np.float64)
        i32 = ctypes.POINTER(ctypes.c_int32)
        fn = lib().qsg_nearest_facet_probe_noalloc if noalloc else lib().qsg_nearest_facet_probe
        _check(fn(self._h, n, xyz.ctypes.data_as(_D), direction.ctypes.data_as(_D), cell.ctypes.data_as(i32),
                  dist.ctypes.data_as(_D), facet.ctypes.data_as(i32), moved.ctypes.data_as(_D)))
        return dist, facet, moved

    def coral_check(self) -> CoralReport:
        r = CoralReport()
        _check(lib().qsg_coral_check(self._h, ctypes.byref(r)))
        return r

    # ---- particle state
    def set_particles(self, ps: ParticleSet):
        v = ps.c_view()
        _check(lib().qsg_set_particles(self._h, ctypes.byref(v), len(ps)))

    def census(self) -> ParticleSet:
        n = ctypes.c_uint64()
        _check(lib().qsg_get_census_count(self._h, ctypes.byref(n)))
        ps = ParticleSet.empty(n.value)
        v = ps.c_view()
        _check(lib().qsg_get_census(self._h, ctypes.byref(v), n.value))
        return ps

    def processing(self) -> ParticleSet:
        n = ctypes.c_uint64()
        _check(lib().qsg_get_processing_count(self._h, ctypes.byref(n)))
        ps = ParticleSet.empty(n.value)
        v = ps.c_view()
        _check(lib().qsg_get_processing(self._h, ctypes.byref(v), n.value))
        return ps


class Batch:
    """Many problems advancing one cycle per ``step()`` in shared launches
    (qsg_batch).  The problems stay usable afterwards."""

    def __init__(self, problems: Sequence[Problem]):
        self.problems = list(problems)
        arr = (ctypes.c_void_p * len(self.problems))(*[p._h.value for p in self.problems])
        h = ctypes.c_void_p()
        _check(lib().qsg_batch_create(arr, len(self.problems), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().qsg_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, strict: bool = True):
        """One cycle of every active problem: list of (balance dict, flux sum)
        (zeros for inactive problems).  A problem whose cycle raised a device
        error flag is taken out of the batch (``problem_error``) and gets
        None; the others advance.  strict: raise QsgError after such a step
        (its other results are then lost to the caller)."""
        n = len(self.problems)
        bal = (Balance * n)()
        flux = (ctypes.c_double * n)()
        rc = lib().qsg_batch_step(self._h, bal, flux)
        if rc != QSG_OK and (strict or rc not in (3, 5)):
            _check(rc)
        out = [(bal[k].as_dict(), flux[k]) for k in range(n)]
        if rc != QSG_OK:
            for k in range(n):
                if self.problem_error(k):
                    out[k] = None
        return out

    def problem_error(self, index: int) -> int:
        """The error code slot ``index``'s problem was taken out with (0: none)."""
        code = ctypes.c_int()
        _check(lib().qsg_batch_problem_error(self._h, index, ctypes.byref(code)))
        return code.value

    def set_active(self, index: int, active: bool):
        _check(lib().qsg_batch_set_active(self._h, index, 1 if active else 0))

    def set_tuning(self, **knobs):
        """Launch shape; knobs not given take DEFAULT_TUNING's values."""
        t = _tuning(knobs)
        _check(lib().qsg_batch_set_tuning(self._h, ctypes.byref(t)))

    def stats(self) -> dict:
        st = Stats()
        _check(lib().qsg_batch_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def run(self, steps: Sequence[int]):
        """Runs every problem for its own number of cycles (steps[i]),
        deactivating each when it is done.  Returns per-problem lists of
        (balance, flux sum)."""
        out = [[] for _ in self.problems]
        for k in range(max(steps)):
            for i, s in enumerate(steps):
                if k == s:
                    self.set_active(i, False)
            res = self.step()
            for i, s in enumerate(steps):
                if k < s:
                    out[i].append(res[i])
        return out


class Server:
    """Continuous-batching serving loop (qsg_server): requests are admitted by
    HBM budget, stay resident, and advance together one cycle per step()."""

    QUEUED, RUNNING, DONE, FAILED, CANCELLED = 0, 1, 2, 3, 4

    def __init__(self, hbm_budget_bytes: int, max_active: int = 0, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().qsg_server_create(device, int(hbm_budget_bytes), max_active, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().qsg_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, deck_text: str, particle_capacity: int = 0) -> int:
        """Queues a request; particle_capacity sizes its particle buffers
        (0: the library's default)."""
        rid = ctypes.c_uint64()
        opts = Options(particle_capacity=int(particle_capacity))
        _check(lib().qsg_server_submit_opts(self._h, deck_text.encode(), ctypes.byref(opts), ctypes.byref(rid)))
        return rid.value

    def step(self):
        a, qd = ctypes.c_int(), ctypes.c_int()
        _check(lib().qsg_server_step(self._h, ctypes.byref(a), ctypes.byref(qd)))
        return a.value, qd.value

    def status(self, rid: int):
        st, n = ctypes.c_int(), ctypes.c_int()
        _check(lib().qsg_server_status(self._h, rid, ctypes.byref(st), ctypes.byref(n)))
        return st.value, n.value

    def result(self, rid: int):
        _, n = self.status(rid)
        bal = (Balance * max(n, 1))()
        flux = (ctypes.c_double * max(n, 1))()
        _check(lib().qsg_server_result(self._h, rid, bal, flux, n))
        return [(bal[k].as_dict(), flux[k]) for k in range(n)]

    def cancel(self, rid: int):
        """Stops a queued or running request; its completed cycles stay readable."""
        _check(lib().qsg_server_cancel(self._h, rid))

    def release(self, rid: int):
        _check(lib().qsg_server_release(self._h, rid))

    def set_tuning(self, **knobs):
        """Launch shape; knobs not given take DEFAULT_TUNING's values."""
        t = _tuning(knobs)
        _check(lib().qsg_server_set_tuning(self._h, ctypes.byref(t)))

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().qsg_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()


def math_probe(op: int, a: np.ndarray, b: Optional[np.ndarray] = None, device: int = 0) -> np.ndarray:
    """Evaluates one FP64 builtin on the device (see qsg_math_probe)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(a if b is None else b, dtype=np.float64)
    out = np.zeros_like(a)
    _check(lib().qsg_math_probe(op, a.ctypes.data_as(_D), b.ctypes.data_as(_D), out.ctypes.data_as(_D),
                                a.size, device))
    return out

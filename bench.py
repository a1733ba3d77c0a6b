"""Benchmark of the per-cycle particle-tracking path on one MI355X.

The headline (``value``) is the batched form on the mixed_batch workload:
64 independent transport problems (quicksilver_amd.decks.mixed_batch(): mesh
edges 8-39 cells, log-uniform and capped so every cell sources particles;
1.3e4-1.5e6 particles; 1-3 nested-box materials x 5-20 isotopes x 3
reactions; seeded polynomial cross sections over 230 groups; FP64), all
advanced together one cycle per step, each step replayed from captured HIP
graphs.  One "step" is one time step (cycle) of all 64 problems.  W untimed
warm-up steps run first, then K steps are timed; the same problems run one
at a time (each W untimed cycles, then K timed) give the baseline, so both
sides time identical work (the tallies are bit-identical).  Problem
construction and release are outside both timed regions and reported as
end_to_end.

Also in the line:
  roofline          the batch tracking kernel (k_batch_track), the kernel that
                    dominates a step: bound by VALU issue (SQ counters in
                    profiles/k_batch_track_mixed_batch_sq.md), with its
                    measured HBM traffic (profiles/..._traffic.md) and the
                    reference data model's bytes per segment beside it
  serving           Poisson stream (decks.serve_stream) at 0.75 x the batched
                    capacity: aggregate segments/s and p50/p99 latency, vs the
                    same stream one request at a time
  plain_nested_boxes  one 40^3 problem (1e6 particles), the plain method
  large_checkerboard  one 80^3 problem (4e7 particles, 20 steps)
  cpu_baseline      the reference CPU path (oracle/_ref/qs_ref, OpenMP) on a
                    bounded sample of the same 64 decks, timed on this host
  cpu_reference_small_cube  the reference on the 10^3 deck (1e5 particles,
                    10 steps) with all usable cores and with one thread

Reference data model bytes per segment (kept beside the roofline):
  B_seg = 8 (sigma_t) + 768 (24 planes x 32 B) + 336 (14 nodes x 24 B)
        + 144 (~12 facets x 3 node ids x 4 B) + 16 (flux read+write)
        + 40 f_fc (adjacency) + f_col (16 nIso + 8 nIso nReact) (collision scan)
        + 272 L/S (136 B particle record loaded + stored per history load)
  with f_fc, f_col and L/S taken from the run's own bit-exact tallies.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "segments/sec on 1 MI355X, batched vs one-at-a-time; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
# VALU issue: each of the 4 SIMDs of a CU issues a wave64 VALU instruction
# every 4 clocks (16 lanes per clock), one instruction per CU per clock:
# 256 CUs x 2.4 GHz.  The FP64 vector peak (78.6 TFLOP/s = 256 x 2.4e9 x 64
# lanes x 2) is the same rate counted in FMA flops.
CUS = 256
PEAK_CLOCK_HZ = 2.4e9
VALU_PEAK_INSTR_S = CUS * PEAK_CLOCK_HZ
QS_REF = os.path.join(ROOT, "oracle", "_ref", "qs_ref")
PROFILES = os.path.join(ROOT, "profiles")
NESTED_DECK = os.path.join(ROOT, "tests", "golden", "decks", "nested_boxes.inp")
SMALL_CUBE_DECK = os.path.join(ROOT, "tests", "golden", "decks", "small_cube.inp")


def progress(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------ profiles ----
def profile_values(name: str) -> dict:
    """``key: value`` lines of a committed profile summary (profiles/<name>),
    as floats; {} when the file is absent."""
    out = {}
    try:
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                m = re.match(r"^([a-z_0-9]+):\s*([-+0-9.eE]+)\s*$", line.strip())
                if m:
                    out[m.group(1)] = float(m.group(2))
    except OSError:
        pass
    return out


def data_model_bytes(stats: dict, n_iso: float, n_react: float) -> float:
    S = max(stats["segments"], 1)
    f_fc = stats["facet_crossings"] / S
    f_col = stats["collisions"] / S
    loads = stats["particle_loads"] / S
    return (8 + 768 + 336 + 144 + 16 + 40 * f_fc + f_col * (16 * n_iso + 8 * n_iso * n_react)
            + 272 * loads)


def roofline(kernel: str, workload: str, segments: float, launches: float, kernel_s: float,
             n_iso: float, n_react: float, stats: dict) -> dict:
    """The tracking kernel against what bounds it.  VALU issue: VALU
    instructions per segment (SQ_INSTS_VALU / segments of the profiled
    launches, profiles/<kernel>_<workload>_sq.md) x this run's segments per
    launch / this run's average launch time, against 256 CUs x 2.4 GHz.
    HBM: measured bytes per segment (FETCH_SIZE / WRITE_SIZE,
    profiles/<kernel>_<workload>_traffic.md) on the same basis, and the
    reference data model's bytes per segment."""
    sq = profile_values(f"{kernel}_{workload}_sq.md")
    tr = profile_values(f"{kernel}_{workload}_traffic.md")
    launches = max(launches, 1)
    seg_launch = segments / launches
    avg_s = kernel_s / launches
    out = {"bound": "valu", "kernel": kernel, "unit": "VALU instr/s (wave64)", "peak": VALU_PEAK_INSTR_S,
           "segments_per_launch": seg_launch, "avg_launch_ms": avg_s * 1e3}
    vps = sq.get("valu_insts_per_segment")
    if vps and avg_s > 0:
        out["achieved"] = vps * seg_launch / avg_s
        out["frac"] = out["achieved"] / VALU_PEAK_INSTR_S
        out["valu_insts_per_segment"] = vps
    else:
        out["achieved"] = out["frac"] = None
    for k in ("active_lanes_per_valu_inst", "lane_utilisation", "valu_busy", "waves_per_simd_resident",
              "wait_any_frac"):
        if k in sq:
            out[k] = sq[k]
    if "lane_utilisation" in sq and out.get("frac") is not None:
        out["useful_lane_frac"] = out["frac"] * sq["lane_utilisation"]
    out["counters"] = f"profiles/{kernel}_{workload}_sq.md"
    tb = tr.get("traffic_bytes_per_segment")
    bseg = data_model_bytes(stats, n_iso, n_react)
    out["traffic"] = tb * seg_launch if tb else None
    out["traffic_bytes_per_segment"] = tb
    out["traffic_gbs"] = tb * seg_launch / avg_s / 1e9 if tb and avg_s > 0 else None
    out["traffic_frac"] = out["traffic_gbs"] / HBM_PEAK_GBS if out["traffic_gbs"] else None
    out["traffic_source"] = (f"profiles/{kernel}_{workload}_traffic.md: rocprofv3 FETCH_SIZE and WRITE_SIZE, "
                             "scaled by the factors a 2 GiB copy gives, per segment, x this run's segments "
                             "per launch")
    out["hbm_data_model"] = {
        "bytes_per_segment": bseg,
        "achieved_gbs": bseg * seg_launch / avg_s / 1e9 if avg_s > 0 else None,
        "frac_of_hbm_peak": bseg * seg_launch / avg_s / 1e9 / HBM_PEAK_GBS if avg_s > 0 else None,
        "note": "the reference's data model (24 planes, 14 nodes, node ids and a 136 B particle record per "
                "segment); this kernel reads far fewer bytes (planes from L2, nodes recomputed, a particle "
                "record once per history), so this figure can exceed the HBM peak and is not the bound",
    }
    return out


# -------------------------------------------------------------- host CPU ----
def host_cpus() -> dict:
    info = {"nproc": os.cpu_count()}
    try:
        info["usable"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["usable"] = os.cpu_count()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("Socket(s)", "Core(s) per socket", "Thread(s) per core", "Model name"):
                info[k] = v.strip()
        if "Socket(s)" in info and "Core(s) per socket" in info:
            info["physical_cores"] = int(info["Socket(s)"]) * int(info["Core(s) per socket"])
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def reference_threads(cpus: dict) -> int:
    """All physical cores this process may use: the usable CPUs, at most the
    physical cores, at most OMP_NUM_THREADS when the host caps it."""
    n = cpus.get("usable") or 1
    if cpus.get("physical_cores"):
        n = min(n, cpus["physical_cores"])
    if os.environ.get("OMP_NUM_THREADS"):
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    return max(1, n)


def run_reference(deck_text: str, threads: int, n_steps: int = 0):
    """The reference CPU path (oracle/_ref/qs_ref, OpenMP) on one deck
    (n_steps > 0: that many steps, the reference's -N): per-cycle balances,
    total segments and tracking seconds (its FOM numerator and denominator,
    MC_Fast_Timer.cc:97-104)."""
    fields = ["start", "source", "rr", "split", "absorb", "scatter", "fission", "produce", "collision",
              "escape", "census", "num_segments", "end"]
    with tempfile.TemporaryDirectory() as tmp:
        deck = os.path.join(tmp, "deck.inp")
        with open(deck, "w") as f:
            f.write(deck_text)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        cmd = [QS_REF, "-i", deck] + (["-N", str(n_steps)] if n_steps > 0 else [])
        out = subprocess.run(cmd, capture_output=True, text=True, env=env, cwd=tmp, timeout=900,
                             check=True).stdout
    bals, seg, secs = [], 0, 0.0
    for line in out.splitlines():
        tok = line.split()
        if tok and tok[0] == "QSREF":
            bals.append(dict(zip(fields, (int(t) for t in tok[2:15]))))
        elif tok and tok[0] == "QSREF_FOM":
            seg, secs = int(tok[1]), float(tok[2])
    return bals, seg, secs


def cpu_mixed_sample(ds, gpu_balances, warmup: int, threads: int, budget_segments: float):
    """The reference on the decks of the mix in order, each taken if its
    segments (estimated from the GPU's timed cycles of the same deck) still
    fit the budget; each run for the GPU's warm-up + timed steps (-N), one
    deck after another.  segments / sum of tracking times.  The reference's
    tallies of the timed cycles are compared with the GPU's."""
    seg, secs, used, same, compared = 0, 0.0, [], True, 0
    for i, d in enumerate(ds):
        timed_gpu = gpu_balances[i]
        steps = len(timed_gpu)
        est = sum(b["num_segments"] for b in timed_gpu) * (warmup + steps) / max(steps, 1)
        if seg + est > budget_segments:
            continue
        bals, s, t = run_reference(d.text(), threads, warmup + steps)
        timed = bals[warmup:warmup + steps]
        if len(timed) == steps:
            same = same and timed == timed_gpu
            compared += 1
        seg += s
        secs += t
        used.append(i)
    if not used:
        return None
    return {"value": seg / secs if secs > 0 else None, "unit": "segments/s", "cores": threads,
            "kind": "reference",
            "sample": f"{len(used)} of the {len(ds)} mixed_batch decks (indices {used}: taken in order while "
                      f"their segments fit {budget_segments:.2g}), each for the GPU's {warmup} warm-up + timed "
                      "steps, one after another "
                      f"with {threads} OpenMP threads; {seg} segments / {secs:.3f} s of the reference's "
                      "tracking time (its FOM; excludes source and population control)",
            "segments": seg, "decks_compared_with_gpu": compared, "tallies_identical_to_gpu": same}


def cpu_small_cube(threads: int):
    with open(SMALL_CUBE_DECK) as f:
        text = f.read()
    out = {}
    for label, n in (("all_cores", threads), ("one_thread", 1)):
        _, seg, secs = run_reference(text, n)
        out[label] = {"threads": n, "segments": seg, "tracking_seconds": secs,
                      "segments_per_s": seg / secs if secs > 0 else None}
    return out


# ------------------------------------------------------------- workloads ----
def mixed_batch(steps: int, warmup: int):
    from quicksilver_amd import autotune, decks, workloads
    ds = decks.mixed_batch(count=64)
    texts = [d.text() for d in ds]
    batch_tune = autotune.lookup(autotune.signature("batch", texts)) or dict(autotune.DEFAULT)
    seq_tune = autotune.lookup(autotune.signature("sequence", texts))
    progress(f"mixed_batch: 64 problems batched, {warmup} + {steps} steps")
    batched = workloads.run_batched(ds, batch_tune, cycles=steps, warmup=warmup)
    progress(f"mixed_batch: batched {batched['segments_per_s']:.4g} segments/s; one at a time")
    alone = {}
    shapes = [("tuned_for_the_sequence", seq_tune), ("batch_shape", batch_tune)] if seq_tune else \
        [("batch_shape", batch_tune)]
    for label, shape in shapes:
        r = workloads.run_one_at_a_time(ds, shape, cycles=steps, warmup=warmup)
        if r["balances"] != batched["balances"] or r["fluxes"] and any(
                abs(a - b) > 1e-10 * abs(b) for fa, fb in zip(r["fluxes"], batched["fluxes"])
                for a, b in zip(fa, fb)):
            raise SystemExit(f"one-at-a-time ({label}) tallies differ from the batched tallies")
        alone[label] = {"segments_per_s": r["segments_per_s"], "solve_s": r["solve_s"],
                        "end_to_end_segments_per_s": r["end_to_end_segments_per_s"],
                        "kernel_segments_per_s": r["segments"] / r["kernel_seconds"] if r["kernel_seconds"] else None,
                        "tuning": shape}
        progress(f"mixed_batch: one at a time ({label}) {r['segments_per_s']:.4g} segments/s")
    best = max(alone.values(), key=lambda v: v["segments_per_s"])
    return ds, batched, alone, best, batch_tune


def serving(n_requests: int, capacity_problems_per_s: float, tune):
    from quicksilver_amd import decks, workloads
    rate = 0.75 * capacity_problems_per_s
    reqs = decks.serve_stream(count=n_requests, rate_per_s=rate)
    progress(f"serve_stream: {n_requests} requests at {rate:.3g}/s, continuous batching")
    served = workloads.serve_stream(reqs, batched=True, tuning=tune)
    progress("serve_stream: one at a time")
    base = workloads.serve_stream(reqs, batched=False)
    keys = ("segments_per_s", "p50_s", "p99_s", "mean_s", "max_s", "wall_s")
    return {
        "workload": f"serve_stream: {n_requests} requests drawn like mixed_batch with 5-20 steps, Poisson "
                    f"arrivals at {rate:.3g}/s (0.75 x the batched capacity in problems/s), HBM budget 200 GiB",
        "continuous_batching": {k: served[k] for k in keys},
        "one_at_a_time": {k: base[k] for k in keys},
        "throughput_ratio": served["segments_per_s"] / base["segments_per_s"],
        "p50_ratio": base["p50_s"] / served["p50_s"],
        "p99_ratio": base["p99_s"] / served["p99_s"],
    }


def plain(deck_path: str, steps: int, warmup: int):
    import quicksilver_amd as q
    from quicksilver_amd import autotune
    text = open(deck_path).read()
    shape = autotune.lookup(autotune.signature("plain", [text]))
    with q.Problem(text) as p:
        if shape:
            p.set_tuning(**shape)
        if warmup:
            p.run(warmup)
        p.reset_stats()
        t0 = time.perf_counter()
        out = p.run(steps)
        wall = time.perf_counter() - t0
        st = p.stats()
        tune = p.tuning().as_dict()
    seg = sum(b["num_segments"] for b, _, _ in out)
    res = {"segments_per_s": seg / wall, "ms_per_step": wall / max(steps, 1) * 1e3, "steps": steps,
           "warmup": warmup, "segments_timed": seg,
           "kernel_segments_per_s": st["segments"] / st["track_kernel_seconds"] if st["track_kernel_seconds"] else None,
           "phase_seconds": {"between_tracking_launches": st["init_seconds"], "tracking": st["tracking_seconds"],
                             "track_kernels": st["track_kernel_seconds"], "finalize": st["finalize_seconds"],
                             "wall": wall},
           "tuning": tune}
    return res, st


def copy_probe(gib: float = 2.0, reps: int = 3):
    """A plain device-to-device copy of known size, run under the same
    rocprofv3 counters as the tracking kernels: FETCH_SIZE / WRITE_SIZE read
    for it (per dispatch) calibrate what those counters report for a known
    number of bytes read and written."""
    import torch
    n = int(gib * (1 << 30)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "dispatches": reps}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed steps (cycles of every problem)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed steps before the timed ones")
    ap.add_argument("--only", choices=["batched", "plain", "large"],
                    help="run one workload alone, printing its JSON (for profiling runs)")
    ap.add_argument("--eager", action="store_true",
                    help="with --only batched: launch each step's kernels one by one, not from graphs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--cpu-budget-segments", type=float, default=1.2e8)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 2 GiB device copy (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the library drives one GPU per "
              "process (across GPUs a problem would be split by particles, a batch by whole problems, with no "
              "exchange)", file=sys.stderr)
        sys.exit(2)

    import quicksilver_amd as q
    q.lib()
    if args.only == "batched":
        from quicksilver_amd import autotune, decks, workloads
        ds = decks.mixed_batch(count=64)
        tune = autotune.lookup(autotune.signature("batch", [d.text() for d in ds])) or dict(autotune.DEFAULT)
        if args.eager:
            tune = dict(tune, graphs=0)
        r = workloads.run_batched(ds, tune, cycles=args.steps, warmup=args.warmup)
        print(json.dumps({"workload": "mixed_batch batched", "segments_per_s": r["segments_per_s"],
                          "tuning": tune, "stats": r["stats"]}), flush=True)
        return
    if args.only == "plain":
        r, _ = plain(NESTED_DECK, args.steps, args.warmup)
        print(json.dumps({"workload": "nested_boxes plain", **r}), flush=True)
        return

    from quicksilver_amd import decks
    if args.only == "large":
        res, _ = large(args)
        print(json.dumps({"workload": "checkerboard plain", **res}), flush=True)
        return

    ds, batched, alone, best, batch_tune = mixed_batch(args.steps, args.warmup)
    st = batched["stats"]
    n_iso = sum(int(m["nIsotopes"]) for d in ds for m in d.materials) / sum(len(d.materials) for d in ds)
    line = {
        "metric": METRIC,
        "value": batched["segments_per_s"],
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batched["solve_s"] / max(args.steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded decks (quicksilver_amd.decks), generated in-process",
        "config": {
            "workload": "mixed_batch: 64 independent problems; mesh edge 8-39 cells (log-uniform, capped so every "
                        "cell sources >= 1.5 particles per step), 24 facets per cell; 1.3e4-1.5e6 particles "
                        "(log-uniform 1e4-2e6); 1-3 nested-box materials x 5-20 isotopes x 3 reactions, seeded "
                        "polynomial cross sections over 230 groups; dt 1e-8 s",
            "mode": "batched: every cycle of all 64 problems in one step of three captured HIP graphs (set-up, "
                    "tracking, flux reduction)",
            "timed": f"steps {args.warmup}..{args.warmup + args.steps - 1} of every problem (after "
                     f"{args.warmup} untimed warm-up steps); problem construction and release excluded",
            "segments_timed": batched["segments"],
            "tuning": batch_tune,
        },
        "one_at_a_time": alone,
        "speedup_vs_one_at_a_time": batched["segments_per_s"] / best["segments_per_s"],
        "end_to_end_segments_per_s": batched["end_to_end_segments_per_s"],
        "end_to_end_speedup": batched["end_to_end_segments_per_s"] / best["end_to_end_segments_per_s"],
        "batched_kernel_segments_per_s": st["segments"] / st["track_kernel_seconds"]
        if st["track_kernel_seconds"] > 0 else None,
        "roofline": roofline("k_batch_track", "mixed_batch", st["segments"], st["track_launches"],
                             st["track_kernel_seconds"], n_iso, 3, st),
        "cpu_baseline": None,
    }
    # batched capacity in problems of 10 steps per second, end to end
    # (construction, 10 steps, release)
    capacity = 64 / (batched["build_s"] + batched["release_s"] + batched["solve_s"] * 10.0 / max(args.steps, 1))
    if not args.no_serve:
        line["serving"] = serving(args.serve_requests, capacity, batch_tune)
    if not args.no_plain:
        progress("nested_boxes: plain tracking")
        res, pst = plain(NESTED_DECK, args.steps, args.warmup)
        res["workload"] = ("nested_boxes: 40^3 cells, 3 materials x 10 isotopes x 3 reactions, 230 groups, "
                           "1e6 particles, dt 1e-8 s; one problem, cycles enqueued ahead")
        res["roofline"] = roofline("k_track", "nested_boxes", pst["segments"], pst["track_launches"],
                                   pst["track_kernel_seconds"], 10, 3, pst)
        line["plain_nested_boxes"] = res
    if not args.no_large:
        line["large_checkerboard"], _ = large(args)
    if not args.no_cpu_baseline and os.path.exists(QS_REF):
        cpus = host_cpus()
        threads = reference_threads(cpus)
        progress(f"reference CPU path: mixed_batch sample on {threads} threads")
        line["cpu_baseline"] = cpu_mixed_sample(ds, batched["balances"], args.warmup, threads,
                                                args.cpu_budget_segments)
        progress("reference CPU path: small_cube, all cores and one thread")
        line["cpu_reference_small_cube"] = {
            "deck": "tests/golden/decks/small_cube.inp (10^3 cells, 1 material, flat XS, 1e5 particles, 10 steps)",
            **cpu_small_cube(threads)}
        line["host_cpus"] = cpus
    print(json.dumps(line), flush=True)


def large(args):
    """checkerboard: 80^3 cells, 2 materials x 20 isotopes in a seeded 8^3
    checkerboard, 4e7 particles, 20 steps: 2 untimed, 18 timed."""
    from quicksilver_amd import decks
    progress("checkerboard 80^3, 4e7 particles: plain tracking")
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "checkerboard.inp")
        decks.checkerboard().write(path)
        res, st = plain(path, 18, 2)
    res["workload"] = ("checkerboard: 80^3 cells, 2 materials x 20 isotopes x 3 reactions in a seeded 8^3 "
                       "checkerboard, 230 groups, 4e7 particles, 20 steps (2 untimed, 18 timed)")
    return res, st


if __name__ == "__main__":
    main()

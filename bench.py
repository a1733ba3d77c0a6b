"""Benchmark of the per-cycle particle-tracking path on one MI355X.

Workload (one "step" = one time step / cycle of the transport problem):
  nested_boxes deck (tests/golden/decks/nested_boxes.inp): 40^3-cell mesh
  (24 facets per cell), three materials in nested boxes, 10 isotopes x 3
  reactions each with seeded polynomial cross sections over 230 groups,
  1e6 particles, dt = 1e-8 s, FP64.  The first W cycles are warm-up, the next
  K cycles are timed (the deck's 10 steps for the defaults W=2, K=8).

Printed: one JSON line with
  value      segments/s over the timed cycles, whole cycle (source +
             population control + tracking + tally reduction), wall clock,
             particle state resident in HBM throughout
  roofline   tracking-pass kernel: algorithmic bytes (per-segment model below)
             / summed kernel time from HIP events on the library's stream
  cpu_baseline  the reference CPU path (oracle/_ref/qs_ref, OpenMP) on the
             first steps of the same deck, timed on this host

Algorithmic bytes per segment (reference data model):
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
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DECK = os.path.join(ROOT, "tests", "golden", "decks", "nested_boxes.inp")
QS_REF = os.path.join(ROOT, "oracle", "_ref", "qs_ref")
METRIC = "segments/sec on 1 MI355X, batched vs one-at-a-time; p50/p99 solve latency"


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_track_nested_boxes_traffic.md")


def measured_traffic_per_segment():
    """HBM bytes per segment of k_track on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated against a copy of
    known size); None when the profile is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_segment:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def bytes_per_segment(stats: dict, n_iso: int, n_react: int) -> float:
    S = max(stats["segments"], 1)
    f_fc = stats["facet_crossings"] / S
    f_col = stats["collisions"] / S
    loads = stats["particle_loads"] / S
    return (8 + 768 + 336 + 144 + 16 + 40 * f_fc + f_col * (16 * n_iso + 8 * n_iso * n_react)
            + 272 * loads)


def progress(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(steps: int):
    if not os.path.exists(QS_REF):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    with open(DECK) as f:
        text = f.read()
    text = text.replace("nSteps: 10", f"nSteps: {steps}")
    with tempfile.TemporaryDirectory() as tmp:
        deck = os.path.join(tmp, "sample.inp")
        with open(deck, "w") as f:
            f.write(text)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        out = subprocess.run([QS_REF, "-i", deck], capture_output=True, text=True, env=env, cwd=tmp,
                             timeout=600, check=True).stdout
    for line in out.splitlines():
        tok = line.split()
        if tok and tok[0] == "QSREF_FOM":
            return {"value": float(tok[3]), "unit": "segments/s", "cores": int(tok[4]),
                    "kind": "reference",
                    "sample": f"first {steps} of the 10 steps of the same deck, reference OpenMP CPU "
                              f"path, {tok[1]} segments / {float(tok[2]):.3f} s tracking time "
                              "(the reference FOM; excludes source and population control)"}
    return None


def run_reference(deck_text: str):
    """Runs the reference CPU path (oracle/_ref/qs_ref, OpenMP on this host)
    on one deck; returns its per-cycle balances, segments and tracking
    seconds."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    fields = ["start", "source", "rr", "split", "absorb", "scatter", "fission", "produce", "collision",
              "escape", "census", "num_segments", "end"]
    with tempfile.TemporaryDirectory() as tmp:
        deck = os.path.join(tmp, "deck.inp")
        with open(deck, "w") as f:
            f.write(deck_text)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        out = subprocess.run([QS_REF, "-i", deck], capture_output=True, text=True, env=env, cwd=tmp,
                             timeout=600, check=True).stdout
    bals, seg, secs = [], 0, 0.0
    for line in out.splitlines():
        tok = line.split()
        if tok and tok[0] == "QSREF":
            bals.append(dict(zip(fields, (int(t) for t in tok[2:15]))))
        elif tok and tok[0] == "QSREF_FOM":
            seg, secs = int(tok[1]), float(tok[2])
    return bals, seg, secs, threads


def cpu_batch_sample(ds, gpu_balances, budget_segments: float = 1.2e8):
    """The reference CPU path on a bounded sample of the mix (decks in order,
    each taken if its segments still fit the budget): segments/s over the sum
    of their tracking times, and whether every balance tally equals the
    GPU's."""
    if not os.path.exists(QS_REF):
        return None
    seg = 0
    secs = 0.0
    used = 0
    same = True
    threads = 0
    names = []
    for i, (d, gb) in enumerate(zip(ds, gpu_balances)):
        if seg + sum(b["num_segments"] for b in gb) > budget_segments:
            continue
        bals, s, t, threads = run_reference(d.text())
        same = same and bals == gb
        seg += s
        secs += t
        used += 1
        names.append(i)
    if used == 0:
        return None
    return {"value": seg / secs if secs > 0 else None, "unit": "segments/s", "cores": threads,
            "kind": "reference",
            "sample": f"{used} of the {len(ds)} decks (indices {names}: taken in order while their GPU "
                      f"segment counts fit {budget_segments:.2g} segments), all their steps, run serially; "
                      "segments / sum of the reference's tracking times (its FOM; excludes source and "
                      "population control)",
            "segments": seg, "tallies_identical_to_gpu": same}


def batched_vs_plain(n_problems: int, with_cpu: bool):
    """mixed_batch workload: the same decks one at a time vs in one batch;
    per-problem tallies must be identical."""
    from quicksilver_amd import autotune, decks, workloads
    ds = decks.mixed_batch(count=n_problems)
    texts = [d.text() for d in ds]
    plain_tune = autotune.lookup(autotune.signature("plain", [open(DECK).read()]))
    batch_tune = autotune.lookup(autotune.signature("batch", texts))
    progress(f"mixed_batch: {n_problems} problems one at a time")
    plain = workloads.run_one_at_a_time(ds, plain_tune)
    progress(f"mixed_batch: one at a time {plain['segments_per_s']:.4g} segments/s; batched")
    batch = workloads.run_batched(ds, batch_tune)
    progress(f"mixed_batch: batched {batch['segments_per_s']:.4g} segments/s")
    same = plain["balances"] == batch["balances"]
    st = batch["stats"]
    out = {
        "workload": f"mixed_batch: {n_problems} problems, 1e4-2e6 particles (log-uniform), 8-48 cells per "
                    "edge, 1-3 materials x 5-20 isotopes x 3 reactions, polynomial XS, 10 steps each",
        "timing": "segments/s over the solve phase (all steps of every problem; problem construction "
                  "and release excluded from both sides); end_to_end includes them",
        "one_at_a_time_segments_per_s": plain["segments_per_s"],
        "batched_segments_per_s": batch["segments_per_s"],
        "speedup": batch["segments_per_s"] / plain["segments_per_s"],
        "one_at_a_time_solve_s": plain["solve_s"],
        "batched_solve_s": batch["solve_s"],
        "one_at_a_time_end_to_end_segments_per_s": plain["end_to_end_segments_per_s"],
        "batched_end_to_end_segments_per_s": batch["end_to_end_segments_per_s"],
        "end_to_end_speedup": batch["end_to_end_segments_per_s"] / plain["end_to_end_segments_per_s"],
        "segments": batch["segments"],
        "tallies_identical": same,
        "batched_kernel_segments_per_s": st["segments"] / st["track_kernel_seconds"]
        if st["track_kernel_seconds"] > 0 else None,
    }
    if not same:
        raise SystemExit("batched tallies differ from one-at-a-time tallies")
    if with_cpu:
        progress("mixed_batch: reference CPU path on a sample of the decks")
        out["cpu_reference"] = cpu_batch_sample(ds, batch["balances"])
    return out, n_problems / batch["wall_s"]


def serving(n_requests: int, capacity_problems_per_s: float):
    """serve_stream workload: Poisson arrivals at 0.75 x the batched capacity
    (problems/s on mixed_batch), served with continuous batching and, as the
    baseline, one at a time in arrival order."""
    from quicksilver_amd import decks, workloads
    rate = 0.75 * capacity_problems_per_s
    reqs = decks.serve_stream(count=n_requests, rate_per_s=rate)
    progress(f"serve_stream: {n_requests} requests at {rate:.3g}/s, continuous batching")
    served = workloads.serve_stream(reqs, batched=True)
    progress("serve_stream: one at a time")
    base = workloads.serve_stream(reqs, batched=False)
    keys = ("segments_per_s", "p50_s", "p99_s", "mean_s", "wall_s")
    return {
        "workload": f"serve_stream: {n_requests} requests drawn like mixed_batch with 5-20 steps, Poisson "
                    f"arrivals at {rate:.3g}/s (0.75 x batched capacity), HBM budget 200 GiB",
        "continuous_batching": {k: served[k] for k in keys},
        "one_at_a_time": {k: base[k] for k in keys},
        "throughput_ratio": served["segments_per_s"] / base["segments_per_s"],
    }


def copy_probe(gib: float = 2.0, reps: int = 3):
    """A plain device-to-device copy of known size, run under the same
    rocprofv3 counters as the tracking kernel: FETCH_SIZE / WRITE_SIZE read
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
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched (mixed_batch) comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-stream measurement")
    ap.add_argument("--batch-problems", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 2 GiB device copy (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the library drives one "
              "GPU per process (across GPUs a problem would be split by particles, a batch by whole "
              "problems)",
              file=sys.stderr)
        sys.exit(2)

    import quicksilver_amd as q
    from quicksilver_amd import autotune
    shape = autotune.lookup(autotune.signature("plain", [open(DECK).read()]))
    progress("nested_boxes: plain tracking")
    with q.Problem.from_file(DECK) as p:
        if shape:
            p.set_tuning(**shape)
        if args.warmup:
            p.run(args.warmup)
        p.reset_stats()
        t0 = time.perf_counter()
        out = p.run(args.steps)
        t1 = time.perf_counter()
        stats = p.stats()
        tune = p.tuning()
    wall = t1 - t0
    segments = sum(b["num_segments"] for b, _, _ in out)
    value = segments / wall
    bseg = bytes_per_segment(stats, n_iso=10, n_react=3)
    kernel_s = stats["track_kernel_seconds"]
    achieved = bseg * stats["segments"] / kernel_s / 1e9 if kernel_s > 0 else 0.0
    launches = max(stats["track_launches"], 1)
    traffic_seg = measured_traffic_per_segment()

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / max(args.steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "nested_boxes: 40^3 cells x 24 facets, 3 materials x 10 isotopes x 3 "
                        "reactions, 230 groups, polynomial XS, 1e6 particles, dt 1e-8 s",
            "deck": "tests/golden/decks/nested_boxes.inp",
            "segments_timed": segments,
            "mode": "plain (one problem, persistent tracking passes)",
            "tuning": tune.as_dict(),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_seg * stats["segments"] / launches if traffic_seg else None,
            "traffic_source": "profiles/k_track_nested_boxes_traffic.md: rocprofv3 FETCH_SIZE x 2.0 + "
                              "WRITE_SIZE x 1.0 (factors from a 2 GiB copy), bytes per segment x the "
                              "segments of this run's average launch",
            "kernel": "k_track",
            "bytes_per_segment": bseg,
            "traffic_bytes_per_segment": traffic_seg,
            "traffic_gbs": traffic_seg * stats["segments"] / kernel_s / 1e9 if traffic_seg and kernel_s > 0 else None,
            "traffic_frac": traffic_seg * stats["segments"] / kernel_s / 1e9 / HBM_PEAK_GBS
            if traffic_seg and kernel_s > 0 else None,
            "note": "achieved uses the reference data model's bytes per segment; the kernel moves far fewer "
                    "(traffic_bytes_per_segment, measured), so frac can exceed 1: HBM bandwidth is not what "
                    "bounds it (traffic_frac is the measured share of the HBM peak)",
            "segments_per_launch": stats["segments"] / launches,
            "avg_launch_ms": kernel_s / launches * 1e3,
        },
        "tracking_kernel_segments_per_s": stats["segments"] / kernel_s if kernel_s > 0 else None,
        "phase_seconds": {"init": stats["init_seconds"], "tracking": stats["tracking_seconds"],
                          "track_kernels": kernel_s, "finalize": stats["finalize_seconds"],
                          "wall": wall},
        "cpu_baseline": None,
    }
    if not args.no_batch:
        line["batched"], capacity = batched_vs_plain(args.batch_problems, not args.no_cpu_baseline)
        if not args.no_serve:
            line["serving"] = serving(args.serve_requests, capacity)
    if not args.no_cpu_baseline:
        progress("reference CPU path sample")
        line["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

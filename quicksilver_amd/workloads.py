"""Drivers for the batched and serving workloads (used by bench.py).

* ``run_one_at_a_time(decks)`` -- each deck alone on the GPU, in order
  (the plain method, one problem per launch sequence).
* ``run_batched(decks, tuning)`` -- all decks in one qsg_batch: one set of
  launches per cycle for every problem.
* ``serve_stream(requests, ...)`` -- Poisson arrivals served by the
  continuous-batching qsg_server (or, for the baseline, one at a time in
  arrival order); reports aggregate segments/s and p50/p99 latency.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from typing import Dict, List, Optional, Sequence

import numpy as np

import quicksilver_amd as q
from quicksilver_amd.decks import Deck, Request


BUILD_THREADS = 8


def build_problems(texts: Sequence[str], tuning: Optional[Dict[str, int]] = None,
                   threads: int = BUILD_THREADS) -> List[q.Problem]:
    """Constructs one problem per deck on `threads` host threads (the C ABI
    builds independent problems concurrently, and ctypes releases the GIL
    for each call).  On any failure every problem built is released."""
    q.lib()

    def one(text: str) -> q.Problem:
        p = q.Problem(text)
        if tuning:
            p.set_tuning(**tuning)
        return p

    with ThreadPoolExecutor(max(1, threads)) as ex:
        futures = [ex.submit(one, t) for t in texts]
    probs, err = [], None
    for f in futures:
        try:
            probs.append(f.result())
        except Exception as e:  # noqa: BLE001 -- re-raised below
            err = err or e
    if err is not None:
        for p in probs:
            p.close()
        raise err
    return probs


def run_one_at_a_time(decks: Sequence[Deck], tuning: Optional[Dict[str, int]] = None,
                      cycles: Optional[int] = None, warmup: int = 0) -> Dict:
    """Each deck alone: every step of one problem, then the next.  The
    problems are constructed up front, the same way as for run_batched, and
    each is released after its run.  Each problem runs `warmup` untimed
    cycles, then `cycles` timed ones (default: the deck's nSteps).  solve_s
    counts only the timed cycles (what the batched form is compared on);
    wall_s also counts construction, warm-up and release.  balances /
    fluxes hold the timed cycles; kernel_seconds is the tracking kernels'
    device time over them."""
    t0 = time.perf_counter()
    segments = 0
    solve = 0.0
    warm = 0.0
    kernel = 0.0
    per = []
    fluxes = []
    probs = build_problems([d.text() for d in decks], tuning)
    try:
        for p, d in zip(probs, decks):
            if warmup:
                w0 = time.perf_counter()
                p.run(warmup)
                warm += time.perf_counter() - w0
            p.reset_stats()
            n = int(d.simulation["nSteps"]) if cycles is None else cycles
            s0 = time.perf_counter()
            out = p.run(n)
            solve += time.perf_counter() - s0
            kernel += p.stats()["track_kernel_seconds"]
            p.close()
            per.append([b for b, _, _ in out])
            fluxes.append([f for _, f, _ in out])
            segments += sum(b["num_segments"] for b, _, _ in out)
    finally:
        for p in probs:
            p.close()
    wall = time.perf_counter() - t0
    return {"segments": segments, "wall_s": wall, "solve_s": solve, "segments_per_s": segments / solve,
            "end_to_end_segments_per_s": segments / (wall - warm), "kernel_seconds": kernel,
            "balances": per, "fluxes": fluxes}


def run_batched(decks: Sequence[Deck], tuning: Optional[Dict[str, int]] = None,
                cycles: Optional[int] = None, warmup: int = 0) -> Dict:
    """All decks in one batch: `warmup` untimed steps, then the timed ones
    (`cycles` steps of every problem, default each deck's own nSteps, each
    problem leaving the batch when it is done).  solve_s counts the timed
    steps only; stats are the batch's device timings over them."""
    t0 = time.perf_counter()
    probs = build_problems([d.text() for d in decks])
    try:
        with q.Batch(probs) as b:
            if tuning:
                b.set_tuning(**tuning)
            t1 = time.perf_counter()
            for _ in range(warmup):
                b.step()
            st0 = b.stats()
            steps = [int(d.simulation["nSteps"]) if cycles is None else cycles for d in decks]
            s0 = time.perf_counter()
            out = b.run(steps)
            solve = time.perf_counter() - s0
            st1 = b.stats()
            st = {k: st1[k] - st0[k] for k in st1}
    finally:
        r0 = time.perf_counter()
        for p in probs:
            p.close()
    wall = time.perf_counter() - t0
    segments = sum(bal["num_segments"] for per in out for bal, _ in per)
    return {"segments": segments, "wall_s": wall, "solve_s": solve, "segments_per_s": segments / solve,
            "end_to_end_segments_per_s": segments / (wall - (s0 - t1)), "steps": max(steps) if steps else 0,
            "build_s": t1 - t0, "warmup_s": s0 - t1, "release_s": time.perf_counter() - r0,
            "balances": [[bal for bal, _ in per] for per in out],
            "fluxes": [[f for _, f in per] for per in out], "stats": st}


def _percentiles(lat: List[float]) -> Dict[str, float]:
    a = np.array(lat)
    return {"p50_s": float(np.percentile(a, 50)), "p99_s": float(np.percentile(a, 99)),
            "mean_s": float(a.mean()), "max_s": float(a.max())}


def serve_stream(requests: Sequence[Request], batched: bool = True, hbm_budget_bytes: int = 200 << 30,
                 tuning: Optional[Dict[str, int]] = None) -> Dict:
    """Replays the arrival times in real time.  Latency = completion time -
    arrival time of each request."""
    arrivals = [r.arrival_s for r in requests]
    t0 = time.perf_counter()
    lat: List[float] = []
    segments = 0
    if not batched:
        clock = 0.0
        for r in requests:
            now = time.perf_counter() - t0
            if now < r.arrival_s:
                time.sleep(r.arrival_s - now)
            with q.Problem(r.deck.text()) as p:
                if tuning:
                    p.set_tuning(**tuning)
                out = p.run()
            segments += sum(b["num_segments"] for b, _, _ in out)
            lat.append(time.perf_counter() - t0 - r.arrival_s)
        wall = time.perf_counter() - t0
        return {"requests": len(requests), "segments": segments, "wall_s": wall,
                "segments_per_s": segments / wall, **_percentiles(lat), "mode": "one-at-a-time"}
    with q.Server(hbm_budget_bytes) as s:
        if tuning:
            s.set_tuning(**tuning)
        pending = list(range(len(requests)))
        running: Dict[int, int] = {}
        nxt = 0
        done = 0
        while done < len(requests):
            now = time.perf_counter() - t0
            while nxt < len(requests) and arrivals[nxt] <= now:
                running[s.submit(requests[nxt].deck.text())] = nxt
                nxt += 1
            if not running:
                time.sleep(max(0.0, arrivals[nxt] - now))
                continue
            s.step()
            now = time.perf_counter() - t0
            for rid, idx in list(running.items()):
                st, _ = s.status(rid)
                if st in (q.Server.DONE, q.Server.FAILED):
                    if st == q.Server.FAILED:
                        raise RuntimeError(f"request {idx} failed")
                    segments += sum(b["num_segments"] for b, _ in s.result(rid))
                    lat.append(now - arrivals[idx])
                    s.release(rid)
                    del running[rid]
                    done += 1
        stats = s.stats()
    wall = time.perf_counter() - t0
    del pending
    return {"requests": len(requests), "segments": segments, "wall_s": wall, "segments_per_s": segments / wall,
            **_percentiles(lat), "mode": "continuous batching", "server": stats}

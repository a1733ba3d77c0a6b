"""Autotuner for the tracking-pass launch shape.

Knobs (qsg_tuning):
  block_size     threads per workgroup (64..256)
  waves_per_cu   persistent waves launched per compute unit
  refill_min     idle lanes that make a wave refill from its queue
                 (how full a wave is kept with particles)
  chunk          queue entries a wave claims per atomic
  stack_entries  per-wave secondary stack (0: secondaries in extra passes)
  sort_interval  sort the population by cell every n cycles (0: never)
  waves_per_simd register budget of the tracking kernels (2, 3 or 4
                 resident waves per SIMD: occupancy against spills)

Method: coordinate descent from the current best.  Each knob in turn is swept
over its candidate values with the others fixed, scoring every candidate by
segments per second of device time (HIP events on the library's stream)
over a few cycles of the workload itself: tracking plus cycle set-up for
one problem (so a sort pays for itself), the whole step (set-up, tracking
passes and flux reduction) for a batch.  Each candidate runs on a fresh copy of the workload
over the same cycles, so every candidate does identical work (the tallies
do not depend on the shape).  The sweep repeats until a full round changes
nothing (at most ``rounds`` rounds).  Results are cached in a JSON table
keyed by a workload signature, so later runs start from the tuned shape.

Usage:
    python -m quicksilver_amd.autotune deck.inp                 (plain form)
    python -m quicksilver_amd.autotune --batch a.inp b.inp ...  (batched form)
    python -m quicksilver_amd.autotune --mixed-batch 64         (the 64-problem mix)
    python -m quicksilver_amd.autotune --mixed-batch 64 --sequence
                                        (the same mix one problem at a time)
"""
from __future__ import annotations

import argparse
import json
import os
import time
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import quicksilver_amd as q

SPACE = {
    "block_size": [64, 128, 256],
    "waves_per_cu": [4, 8, 12, 16, 24, 32],
    "refill_min": [1, 4, 8, 16, 32],
    "chunk": [16, 32, 64, 128, 256],
    "stack_entries": [0, 512, 1024, 2048],
    "sort_interval": [0, 1, 2, 4],
    "waves_per_simd": [2, 3, 4],
}
ORDER = ["waves_per_simd", "waves_per_cu", "refill_min", "chunk", "block_size", "stack_entries",
         "sort_interval"]
DEFAULT = dict(q.DEFAULT_TUNING)
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")

Shape = Dict[str, int]


def signature(kind: str, decks: Sequence[str]) -> str:
    """Workload class: plain/batched, mesh size and population of the decks."""
    import re
    sizes = []
    for text in decks:
        nx = re.search(r"\bnx:\s*(\d+)", text)
        npart = re.search(r"\bnParticles:\s*(\d+)", text)
        sizes.append((int(nx.group(1)) if nx else 10, int(npart.group(1)) if npart else 1000000))
    if kind == "plain":
        n, p = sizes[0]
        return f"plain:n{n}:p{p}"
    tot = sum(p for _, p in sizes)
    if kind == "sequence":
        return f"sequence:{len(sizes)}:p{tot}"
    return f"batch:{len(sizes)}:p{tot}"


def load_table(path: str = TABLE) -> Dict[str, Shape]:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def lookup(sig: str, path: str = TABLE) -> Optional[Shape]:
    t = load_table(path)
    if sig in t:
        return {**DEFAULT, **{k: int(v) for k, v in t[sig].items() if k in DEFAULT}}
    return None


def store(sig: str, shape: Shape, score: float, path: str = TABLE):
    t = load_table(path)
    t[sig] = {**shape, "segments_per_s": score, "tuned_at": time.strftime("%Y-%m-%d")}
    with open(path, "w") as f:
        json.dump(t, f, indent=1, sort_keys=True)


def coordinate_descent(score: Callable[[Shape], float], start: Optional[Shape] = None, rounds: int = 2,
                       log: Callable[[str], None] = print) -> Tuple[Shape, float]:
    best = dict(start or DEFAULT)
    best_score = score(best)
    log(f"start {best} -> {best_score:.4g}")
    seen = {tuple(sorted(best.items())): best_score}
    for _ in range(rounds):
        changed = False
        for knob in ORDER:
            for v in SPACE[knob]:
                cand = dict(best, **{knob: v})
                key = tuple(sorted(cand.items()))
                if key in seen:
                    continue
                s = score(cand)
                seen[key] = s
                log(f"  {knob}={v}: {s:.4g}")
                if s > best_score * 1.01:
                    best, best_score, changed = cand, s, True
        if not changed:
            break
    log(f"best {best} -> {best_score:.4g}")
    return best, best_score


def tune_plain(deck_text: str, warmup: int = 2, cycles: int = 2, rounds: int = 2,
               log: Callable[[str], None] = print) -> Tuple[Shape, float]:
    """Every candidate runs on a fresh copy of the problem over the same
    cycles (warm-up, then `cycles` timed), so all candidates do identical
    work.  Score: segments per second of device time in cycle set-up
    (source, population control, sort) plus tracking."""

    def score(s: Shape) -> float:
        with q.Problem(deck_text) as p:
            p.set_tuning(**s)
            p.run(warmup)
            p.reset_stats()
            p.run(cycles)
            st = p.stats()
        t = st["track_kernel_seconds"] + st["init_seconds"]
        return st["segments"] / t if t > 0 else 0.0

    start = lookup(signature("plain", [deck_text])) or dict(DEFAULT)
    return coordinate_descent(score, start, rounds, log)


def tune_batch(deck_texts: Sequence[str], warmup: int = 1, cycles: int = 2, rounds: int = 1,
               log: Callable[[str], None] = print) -> Tuple[Shape, float]:
    """As tune_plain, on a fresh batch of the decks per candidate.  Score:
    segments per second of the timed steps' device time, first event to
    last (set-up, source, population control, sort, tracking passes and
    flux reduction)."""

    def score(s: Shape) -> float:
        probs = [q.Problem(t) for t in deck_texts]
        try:
            with q.Batch(probs) as b:
                b.set_tuning(**s)
                for _ in range(warmup):
                    b.step()
                st0 = b.stats()
                for _ in range(cycles):
                    b.step()
                st = b.stats()
        finally:
            for p in probs:
                p.close()
        dt = st["tracking_seconds"] - st0["tracking_seconds"]
        return (st["segments"] - st0["segments"]) / dt if dt > 0 else 0.0

    start = lookup(signature("batch", deck_texts)) or dict(DEFAULT)
    return coordinate_descent(score, start, rounds, log)


def tune_sequence(deck_texts: Sequence[str], warmup: int = 1, cycles: int = 2, rounds: int = 1,
                  log: Callable[[str], None] = print) -> Tuple[Shape, float]:
    """The plain form run one problem at a time over a whole set of decks
    (the baseline the batched form is compared with): one shape for every
    problem of the set.  Score: the set's segments per second of device time
    in cycle set-up plus tracking, each problem on a fresh copy over the same
    cycles."""

    def score(s: Shape) -> float:
        seg, t = 0, 0.0
        for text in deck_texts:
            with q.Problem(text) as p:
                p.set_tuning(**s)
                p.run(warmup)
                p.reset_stats()
                p.run(cycles)
                st = p.stats()
            seg += st["segments"]
            t += st["track_kernel_seconds"] + st["init_seconds"]
        return seg / t if t > 0 else 0.0

    start = lookup(signature("sequence", deck_texts)) or dict(DEFAULT)
    return coordinate_descent(score, start, rounds, log)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("decks", nargs="*")
    ap.add_argument("--batch", action="store_true", help="tune the batched form over all decks")
    ap.add_argument("--mixed-batch", type=int, default=0, metavar="N",
                    help="tune the batched form on the first N mixed_batch decks (quicksilver_amd.decks)")
    ap.add_argument("--sequence", action="store_true",
                    help="tune the plain form run one problem at a time over all decks (with --mixed-batch: "
                         "the one-at-a-time baseline of that mix)")
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--no-store", action="store_true")
    a = ap.parse_args()
    if a.mixed_batch:
        from quicksilver_amd import decks as D
        texts = [d.text() for d in D.mixed_batch(count=a.mixed_batch)]
        a.batch = not a.sequence
    elif a.decks:
        texts = [open(d).read() for d in a.decks]
    else:
        ap.error("give deck files or --mixed-batch N")
    if a.sequence:
        shape, sc = tune_sequence(texts, rounds=a.rounds)
        sig = signature("sequence", texts)
    elif a.batch:
        shape, sc = tune_batch(texts, rounds=a.rounds)
        sig = signature("batch", texts)
    else:
        shape, sc = tune_plain(texts[0], rounds=a.rounds)
        sig = signature("plain", texts)
    if not a.no_store:
        store(sig, shape, sc)
    print(json.dumps({"signature": sig, "shape": shape, "segments_per_s": sc}))


if __name__ == "__main__":
    main()

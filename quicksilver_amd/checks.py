"""Size-independent checks on per-cycle balance tallies.

These hold for every run of the reference's algorithm, whatever its size,
so they check full-size device runs that no CPU run can reproduce in
seconds:

* particle conservation per cycle (the per-cycle form of
  MissingParticleTest, CoralBenchmark.cc:152-174):
  start + source + produce + split == absorb + census + escape + rr + fission
* collision == absorb + scatter + fission (CollisionEvent.cc:107-118 counts
  each collision once, as exactly one of the three reactions)
* end == census, and each cycle starts from the previous cycle's census
  (Tallies.cc:67-73)
* facet crossings = segments - collisions - census are non-negative
  (every segment ends in exactly one event, MC_Segment_Outcome.cc)
"""
from __future__ import annotations

from typing import Dict, List, Sequence


def balance_violations(cycles: Sequence[Dict[str, int]], source_per_cycle: int = None) -> List[str]:
    """Returns a description of every violated identity (empty: all hold)."""
    bad = []
    prev_end = None
    for c, b in enumerate(cycles):
        gains = b["start"] + b["source"] + b["produce"] + b["split"]
        losses = b["absorb"] + b["census"] + b["escape"] + b["rr"] + b["fission"]
        if gains != losses:
            bad.append(f"cycle {c}: gains {gains} != losses {losses}")
        if b["collision"] != b["absorb"] + b["scatter"] + b["fission"]:
            bad.append(f"cycle {c}: collision {b['collision']} != absorb + scatter + fission")
        if b["end"] != b["census"]:
            bad.append(f"cycle {c}: end {b['end']} != census {b['census']}")
        if b["num_segments"] < b["collision"] + b["census"]:
            bad.append(f"cycle {c}: fewer segments than collisions + census")
        if prev_end is not None and b["start"] != prev_end:
            bad.append(f"cycle {c}: start {b['start']} != previous end {prev_end}")
        if source_per_cycle is not None and b["source"] != source_per_cycle:
            bad.append(f"cycle {c}: source {b['source']} != {source_per_cycle}")
        prev_end = b["end"]
    return bad

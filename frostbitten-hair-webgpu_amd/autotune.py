"""Autotuner of the render configuration: tile size, depth-bin count and
slices per pixel.

The reference fixes these in CONFIG.hairRender (constants.ts:193-265; tile 16,
bins 32, slices 8 there). Here each instance gets the fastest of a candidate
set measured on the device. All candidates keep the same depth resolution per
tile (bins x slices = 32 ordered depth slices, the 8 x 4 of the 1080p
character workload), so the candidates trade the cost of the tile and fine
passes against each other without coarsening the blend order:
tile in {8, 16, 32} x (bins, slices) in {(4, 8), (8, 4), (16, 2)}.

The simulation step does not depend on these values, so a candidate is timed
on the render passes only (tile clear/draw, sort, fine pass, shading), after
one untimed run. Instances are grouped by resolution and strand-count octave;
the instance with the median strand count stands for its group.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from . import assets as A
from . import batch as B
from . import passes as PS

CANDIDATES = [(t, b, s) for t in (8, 16, 32) for (b, s) in ((4, 8), (8, 4), (16, 2))]


def apply(wl, cand) -> None:
    """Sets the candidate (tile, bins, slices) on a workload's render config,
    with one fine-pass processor per tile up to the configured count."""
    t, b, s = cand
    rc = dataclasses.replace(wl.render, tileSize=t, tileDepthBins=b, slicesPerPixel=s,
                             avgSegmentsPerTile=A.segment_list_budget(wl.strands, wl.points, wl.width, wl.height, t))
    rc.processorCount = B.default_processor_count(dataclasses.replace(rc, processorCount=512), wl.width,
                                                  wl.height)
    wl.render = rc


def group_key(wl) -> tuple:
    return (wl.width, wl.height, int(math.log2(max(1.0, wl.strands / 1000.0))))


def measure(wl, inputs: dict, cand, reps: int = 3, device="cuda") -> float:
    """Milliseconds per render (tile, sort, fine, shading) of `wl` with the
    candidate configuration."""
    w = dataclasses.replace(wl)
    apply(w, cand)
    r = PS.HairRenderer(w, device, inputs=inputs)
    r.beforeFirstFrame()
    ctx = r.ctx()
    r.drawHair(ctx)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        r.drawHair(ctx)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    del r
    return ms


def tune(workloads: list, inputs: list, reps: int = 3, log=None) -> list:
    """The chosen candidate of every workload (also applied to it)."""
    groups = {}
    for k, wl in enumerate(workloads):
        groups.setdefault(group_key(wl), []).append(k)
    choice = [None] * len(workloads)
    table = {}
    for key, members in groups.items():
        members = sorted(members, key=lambda k: workloads[k].strands)
        rep = members[len(members) // 2]
        times = {c: measure(workloads[rep], inputs[rep], c, reps) for c in CANDIDATES}
        best = min(times, key=times.get)
        table[key] = (best, times)
        if log:
            log(f"# tune {key}: {best} {times[best]:.3f} ms (tile 16/8x4: {times[(16, 8, 4)]:.3f} ms)")
        for k in members:
            choice[k] = best
    for wl, c in zip(workloads, choice):
        apply(wl, c)
    return choice


def summary(choice: list) -> dict:
    vals, counts = np.unique(np.array([list(c) for c in choice]), axis=0, return_counts=True)
    return {f"tile{t}_bins{b}_slices{s}": int(n) for (t, b, s), n in zip(vals.tolist(), counts.tolist())}

"""Served form of the hair chain: continuous batching of simulate-and-render
requests.

A request names an asset of the pool and a number of frames; its result is
the image after its last frame. The server keeps each asset's simulation
state in HBM between requests (positions, grid, shading), so a later request
for the same asset continues where the previous one stopped; requests for one
asset run one after another, in arrival order.

Every tick runs one frame of every admitted request, as one merged launch
sequence (batch.py): the frame of each resident instance is recorded once per
parity of its position buffers, and the recorded frames of the tick's
instances are merged (merges are cached per set of (asset, parity)). Requests
that arrive while others run are admitted at the next tick, as long as a slot
is free and the instance fits the HBM budget; idle instances are evicted
least-recently-used first (their state moves to host memory and comes back
bit for bit when the asset is requested again).

With max_active = 1 the same loop serves one request at a time: the plain,
sequential way to serve the same stream.

Pipelined (`pipelined=True`), each instance's frame is recorded as a
simulation part and a render part (batch.record_parts), and a tick runs the
merged simulation parts of the requests with steps left on a second stream
beside the merged render parts of the requests whose last simulated frame is
not yet rendered: a request's first tick only simulates, its last only
renders, and every image is the one its plain frames give.
"""
from __future__ import annotations

import dataclasses
import math
import time
from collections import OrderedDict

import numpy as np
import torch

from . import batch as B
from . import layouts as L
from . import native as N
from . import passes as PS


@dataclasses.dataclass
class Request:
    rid: int
    asset: int
    steps: int
    arrival: float               # seconds after the start of the run
    admitted: float | None = None
    done: float | None = None
    frames_done: int = 0         # frames rendered
    sims_done: int = 0           # simulation steps run (ahead by one while pipelined)
    image: torch.Tensor | None = None   # rasterizer result after the last frame
    screen: torch.Tensor | None = None  # presented rgba8 sRGB image after the last frame


def poisson_requests(n: int, rate: float, assets: list, max_steps: int, seed: int) -> list:
    """`n` requests with exponential inter-arrival times of mean 1/rate, each
    for a uniformly drawn asset and U{1..max_steps} frames."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    return [Request(i, int(rng.choice(assets)), int(rng.integers(1, max_steps + 1)), float(t[i]))
            for i in range(n)]


def estimate_bytes(wl) -> int:
    """Device bytes of one instance: its buffers (as HairRenderer allocates
    them) plus a bound on the library scratch."""
    rc, N_ = wl.render, wl.strands * wl.points
    W, H = wl.width, wl.height
    T = -(-W // rc.tileSize) * -(-H // rc.tileSize)
    entries = L.segments_buffer_entries(rc, W, H)
    per_proc = rc.avgFragmentsPerSlice * rc.slicesPerPixel * rc.tileSize ** 2
    sm = 1024
    buffers = (4 * 16 * N_ + 4 * N_                      # positions x3, tangents, segment lengths
               + 2 * 16 * wl.sim.gridDims ** 3 + 4 * wl.sdf_dims ** 3
               + 16 * T * rc.tileDepthBins + 8 * entries + 8 * entries + 8 * T
               + 16 * rc.processorCount * per_proc + 16 * W * H
               + 8 * W * H + W * H + 4 * sm * sm + 16 * wl.strands * rc.shadingPoints
               + 24 * W * H)                                # HDR scene + HDR texture + screen + mesh normals
    scratch = 36 * N_ + 16 * entries + 8 * T * rc.tileDepthBins + (32 << 20)
    return int(buffers + scratch)


class _Instance:
    """A resident asset: its renderer, uniform slot and recorded frames."""

    def __init__(self, asset: int, r: PS.HairRenderer, slot: int, progs: list, base: int, nbytes: int):
        # progs: per parity, the whole frame, or (simulation part, render part)
        self.asset, self.r, self.slot, self.progs, self.base, self.nbytes = asset, r, slot, progs, base, nbytes
        self.last_used = 0.0

    def parity(self) -> int:
        return self.r.hair.current ^ self.base

    def program(self, part: str = "frame") -> N.Program:
        p = self.progs[self.parity()]
        return p if part == "frame" else p[0 if part == "sim" else 1]

    def close(self) -> None:
        for p in self.progs:
            for q in (p if isinstance(p, tuple) else (p,)):
                q.close()


class HairServer:
    """Continuous batching over a pool of assets (asset id -> (workload,
    host inputs)), within `budget_bytes` of device memory for the resident
    instances and at most `max_active` requests per tick."""

    def __init__(self, pool: dict, budget_bytes: int, max_active: int = 64, device="cuda", keep_images=False,
                 max_resident: int = 256, pipelined: bool = False):
        self.pool = pool
        self.pipelined = pipelined
        self.budget = int(budget_bytes)
        self.max_active = max_active
        self.device = torch.device(device)
        self.dev = self.device.index or 0
        self.keep_images = keep_images
        self.native = N.Context(self.dev)
        self.native.set_stream(torch.cuda.current_stream(self.device))
        self.sim_bytes = L.SIM_UNIFORMS_DTYPE.itemsize
        self.slots = torch.zeros((max_resident, self.sim_bytes), dtype=torch.uint8, device=self.device)
        self.host_slots = torch.zeros((max_resident, self.sim_bytes), dtype=torch.uint8, pin_memory=True)
        self.free_slots = list(range(max_resident - 1, -1, -1))
        self.resident: OrderedDict = OrderedDict()   # asset -> _Instance, least recently used first
        self.saved = {}                                # asset -> host copy of an evicted instance's state
        self.sizes = {}                                # asset -> measured device bytes
        self.merged: OrderedDict = OrderedDict()       # (part, ((asset, parity), ...)) -> merged Program
        self.used = 0
        self.stats = dict(ticks=0, frames=0, evictions=0, materializations=0, merges=0)
        # host seconds spent per activity of the loop
        self.times = dict(materialize=0.0, evict=0.0, merge=0.0, uniforms=0.0, launch=0.0, wait=0.0)

    # -- residency -------------------------------------------------------
    def _bytes_of(self, asset: int) -> int:
        return self.sizes.get(asset) or estimate_bytes(self.pool[asset][0])

    def _evict(self, asset: int) -> None:
        t = time.perf_counter()
        inst = self.resident.pop(asset)
        r = inst.r
        torch.cuda.synchronize(self.device)
        self.saved[asset] = dict(positions=[p.cpu() for p in r.hair.positions], current=r.hair.current,
                                 frameIdx=r.frameIdx, dv=r.grid.densityVelocity.cpu(), shading=r.hair.shading.cpu())
        inst.close()
        for key in [k for k in self.merged if any(a == asset for a, _ in k[1])]:
            self.merged.pop(key).close()
        self.free_slots.append(inst.slot)
        self.used -= inst.nbytes
        r.native.close()
        del inst, r
        self.stats["evictions"] += 1
        self.times["evict"] += time.perf_counter() - t

    def _make_room(self, need: int, busy: set) -> bool:
        for asset in list(self.resident):
            if self.used + need <= self.budget:
                break
            if asset not in busy:
                self._evict(asset)
        return self.used + need <= self.budget

    def _materialize(self, asset: int) -> _Instance:
        t = time.perf_counter()
        wl, inputs = self.pool[asset]
        ctx = N.Context(self.dev)
        ctx.set_stream(torch.cuda.current_stream(self.device))
        r = PS.HairRenderer(wl, self.device, native=ctx, inputs=inputs)
        slot = self.free_slots.pop()
        self.slots[slot].copy_(r.simulationUniforms)
        r.simulationUniforms = self.slots[slot]
        st = self.saved.pop(asset, None)
        if st is None:
            r.beforeFirstFrame()
        else:
            for dst, src in zip(r.hair.positions, st["positions"]):
                dst.copy_(src)
            r.hair.current = st["current"]
            r.frameIdx = st["frameIdx"]
            r.grid.densityVelocity.copy_(st["dv"])
            r.hair.shading.copy_(st["shading"])
        base = r.hair.current
        # parity of `base`, then the other one
        rec = B.record_parts if self.pipelined else B.record_frame
        progs = [rec(r), rec(r)]
        nbytes = B.instance_bytes(r) + ctx.scratch_bytes()
        self.sizes[asset] = nbytes
        inst = _Instance(asset, r, slot, progs, base, nbytes)
        self.resident[asset] = inst
        self.used += nbytes
        self.stats["materializations"] += 1
        self.times["materialize"] += time.perf_counter() - t
        return inst

    # -- one tick ----------------------------------------------------------
    def _merged_for(self, insts: list, part: str = "frame") -> N.Program:
        key = (part, tuple(sorted((i.asset, i.parity()) for i in insts)))
        prog = self.merged.get(key)
        if prog is None:
            t = time.perf_counter()
            ordered = sorted(insts, key=lambda i: i.asset)
            prog = self.native.program_merge([i.program(part) for i in ordered])
            self.merged[key] = prog
            self.stats["merges"] += 1
            while len(self.merged) > 128:
                self.merged.popitem(last=False)[1].close()
            self.times["merge"] += time.perf_counter() - t
        else:
            self.merged.move_to_end(key)
        return prog

    def _tick(self, active: dict) -> None:
        """One frame of every active request; pipelined, the simulation step
        of the requests with steps left beside the render of those with a
        simulated frame not yet rendered."""
        t0 = time.perf_counter()
        if self.pipelined:
            sim = [self.resident[a] for a, q in active.items() if q.sims_done < q.steps]
            ren = [self.resident[a] for a, q in active.items() if q.sims_done > q.frames_done]
        else:
            sim = ren = [self.resident[a] for a in active]
        for inst in sim:
            self.host_slots[inst.slot].numpy()[:] = inst.r.sim_uniforms(inst.r.frameIdx).view(np.uint8).reshape(-1)
        self.slots.copy_(self.host_slots, non_blocking=True)
        t1 = time.perf_counter()
        if self.pipelined:
            sp = self._merged_for(sim, "sim") if sim else None
            rp = self._merged_for(ren, "render") if ren else None
        else:
            prog = self._merged_for(sim)
        t2 = time.perf_counter()
        if self.pipelined:
            if sp is not None:
                self.native.branch_begin()
                sp.launch(self.native)
                self.native.branch_end()
            if rp is not None:
                rp.launch(self.native)
            self.native.branch_join()
        else:
            prog.launch(self.native)
        for inst in sim:
            inst.r.hair.swapPositionBuffersAfterSimIntegration()
            inst.r.frameIdx += 1
        t3 = time.perf_counter()
        torch.cuda.synchronize(self.device)
        t4 = time.perf_counter()
        self.times["uniforms"] += t1 - t0
        self.times["launch"] += t3 - t2
        self.times["wait"] += t4 - t3
        self.stats["ticks"] += 1
        self.stats["frames"] += len(ren)

    # -- the request loop --------------------------------------------------
    def run(self, requests: list, realtime: bool = True) -> list:
        """Serves `requests` (sorted by arrival). With `realtime`, a request
        becomes visible at its arrival time on the wall clock; otherwise all
        are visible at once (arrival order kept)."""
        pending = sorted(requests, key=lambda q: (q.arrival, q.rid))
        waiting = []
        active = {}   # asset -> request running this tick
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: math.inf)
        k = 0
        while k < len(pending) or waiting or active:
            now = clock()
            while k < len(pending) and pending[k].arrival <= now:
                waiting.append(pending[k])
                k += 1
            # admission, in arrival order; one request per asset at a time
            still = []
            for q in waiting:
                if len(active) >= self.max_active or q.asset in active:
                    still.append(q)
                    continue
                if q.asset not in self.resident:
                    if not self._make_room(self._bytes_of(q.asset), set(active)):
                        still.append(q)
                        continue
                    self._materialize(q.asset)
                    if self.used > self.budget:  # the estimate was low: undo
                        self._evict(q.asset)
                        still.append(q)
                        continue
                self.resident.move_to_end(q.asset)
                q.admitted = clock() if realtime else None
                active[q.asset] = q
            waiting = still
            if not active:
                if k < len(pending):
                    time.sleep(max(0.0, min(pending[k].arrival - clock(), 0.002)))
                    continue
                raise RuntimeError("requests that cannot be admitted within the memory budget")
            rendering = {a for a, q in active.items() if not self.pipelined or q.sims_done > q.frames_done}
            self._tick(active)
            t = time.perf_counter() - t0
            for a in list(active):
                q = active[a]
                if q.sims_done < q.steps:
                    q.sims_done += 1
                if a in rendering:
                    q.frames_done += 1
                if q.frames_done == q.steps:
                    q.done = t
                    if self.keep_images:
                        r = self.resident[a].r
                        q.image = r.finePass.image(r.wl.width, r.wl.height).clone()
                        q.screen = r.screenTexture.clone()
                    del active[a]
        self.elapsed = time.perf_counter() - t0
        return requests

    def close(self) -> None:
        for p in self.merged.values():
            p.close()
        self.merged.clear()
        for a in list(self.resident):
            inst = self.resident.pop(a)
            inst.close()
            inst.r.native.close()


def summarize(requests: list, elapsed: float) -> dict:
    lat = np.array([q.done - q.arrival for q in requests])
    frames = sum(q.steps for q in requests)
    return {"requests": len(requests), "frames": frames, "elapsed_s": round(elapsed, 3),
            "frames_per_s": round(frames / elapsed, 2),
            "latency_s": {"p50": round(float(np.percentile(lat, 50)), 4),
                          "p99": round(float(np.percentile(lat, 99)), 4),
                          "mean": round(float(lat.mean()), 4)}}

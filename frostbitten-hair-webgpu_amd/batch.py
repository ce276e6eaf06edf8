"""Batched form of the hair chain: many instances, one launch per kernel.

Each instance is an ordinary HairRenderer (its own buffers, configuration and
library context, so its own scratch). Its frame is recorded through the same
pass calls as the plain path (Context.record_begin/end: nothing runs, the
kernel launches are kept), once per parity of the tick-tock position buffers
(hairObject.ts:102-106). The recorded frames of all instances are merged into
one sequence of batched launches (fh_program_merge): every kernel of the chain
runs once for all instances that launch it, each instance on its own block
range with its own arguments, so each instance computes exactly what its plain
frame computes. The merged frame is then captured as one HIP graph per parity.
Pipelined, each instance's frame is recorded as a simulation part and a render
part; the merged simulation step of frame f then runs on a second stream
beside the merged render passes of frame f-1.

The simulation uniforms of all instances live in one packed device array, so a
frame's per-instance inputs (frameIdx, wind timer) arrive in one copy.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import layouts as L
from . import native as N
from . import passes as PS


def record_frame(r: PS.HairRenderer) -> N.Program:
    """The launches of one frame (sim step + render) of `r`, recorded without
    running; the position buffers swap as in a real frame."""
    for _ in range(2):
        ctx = r.ctx()
        r.native.record_begin()
        try:
            r.simulateHair(ctx)
            r.drawHair(ctx)
        finally:
            prog = r.native.record_end()
        if prog is not None:
            return prog
        # the scratch grew while recording: undo the swap and record again
        r.hair.swapPositionBuffersAfterSimIntegration()
    raise N.FhError("recording failed twice")


def record_parts(r: PS.HairRenderer) -> tuple:
    """The simulation step and the render passes of one frame of `r`,
    recorded separately at the same parity of the position buffers: the
    render part reads the positions that the simulation part reads as "now"
    (frame f-1's) and the simulation part writes the other buffer, so the two
    can run side by side (the pipelined frame of HairRenderer). The grid
    accumulation records on the instance's branch scratch, so the two parts
    share no temporaries."""
    assert not r.hw_passes, "the hardware-raster passes share the branch scratch with the grid accumulation"
    c = r.hair.current
    for _ in range(2):
        ctx = r.ctx()
        r.hair.current = c
        r.native.record_begin()
        try:
            r.simulateHair(ctx)
        finally:
            sim = r.native.record_end()
        r.hair.current = c
        r.native.record_begin()
        try:
            r.drawHair(ctx)
        finally:
            ren = r.native.record_end()
        if sim is not None and ren is not None:
            r.hair.current = 1 - c
            return sim, ren
        for prog in (sim, ren):  # a scratch grew while recording: record again
            if prog is not None:
                prog.close()
    r.hair.current = c
    raise N.FhError("recording failed twice")


class BatchRenderer:
    """Instances rendered together; frame() advances every instance by one
    frame (sim step + render) with one graph launch."""

    def __init__(self, workloads: list, device="cuda", shadow_size: int = 1024, inputs: list | None = None,
                 use_graph: bool = True):
        self.device = torch.device(device)
        dev = self.device.index or 0
        self.renderers = []
        for k, wl in enumerate(workloads):
            ctx = N.Context(dev)
            r = PS.HairRenderer(wl, self.device, native=ctx, shadow_size=shadow_size,
                                inputs=None if inputs is None else inputs[k])
            self.renderers.append(r)
        n = len(self.renderers)
        self.sim_bytes = self.renderers[0].simulationUniforms.numel()
        # all instances' simulation uniforms, one row each
        self.simPacked = torch.zeros((n, self.sim_bytes), dtype=torch.uint8, device=self.device)
        for i, r in enumerate(self.renderers):
            self.simPacked[i].copy_(r.simulationUniforms)
            r.simulationUniforms = self.simPacked[i]
        self.native = N.Context(dev)
        self.native.set_stream(torch.cuda.current_stream(self.device))
        self.frameIdx = 0
        self.use_graph = use_graph
        self.programs = None
        self.graphs = None
        self.build_current = 0

    def __len__(self) -> int:
        return len(self.renderers)

    def sim_uniforms(self, frame_idx: int) -> np.ndarray:
        """Packed simulation uniforms of every instance for frame `frame_idx`."""
        return np.stack([r.sim_uniforms(frame_idx).view(np.uint8).reshape(-1) for r in self.renderers])

    def upload_sim_uniforms(self) -> None:
        self.simPacked.copy_(torch.from_numpy(self.sim_uniforms(self.frameIdx)))

    def beforeFirstFrame(self) -> None:
        for r in self.renderers:
            r.beforeFirstFrame()

    def build(self, device_blocks: int | None = None) -> None:
        """Records both parities of every instance, merges them and (with
        graphs) captures the merged frames. The grid-stride kernels of each
        instance get its share (by point count) of `device_blocks`
        workgroups (default 8 per CU), at least 64."""
        self._share_blocks(device_blocks)
        recorded = [[], []]
        for r in self.renderers:
            for parity in range(2):
                recorded[parity].append(record_frame(r))
        self.programs = [self.native.program_merge(recorded[p]) for p in range(2)]
        for lst in recorded:
            for p in lst:
                p.close()
        self.build_current = self.renderers[0].hair.current  # program 0 reads positions[build_current]
        if self.use_graph:
            self.graphs = self._capture(lambda p: self.programs[p].launch(self.native))

    def _share_blocks(self, device_blocks: int | None) -> None:
        if device_blocks is None:
            device_blocks = 8 * torch.cuda.get_device_properties(self.device).multi_processor_count
        pts = np.array([r.wl.strands * r.wl.points for r in self.renderers], dtype=np.float64)
        for r, w in zip(self.renderers, pts / pts.sum()):
            r.native.set_persistent_blocks(max(64, int(math.ceil(device_blocks * w))))

    def _capture(self, launch) -> list:
        main = torch.cuda.current_stream(self.device)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(main)
        torch.cuda.synchronize(self.device)
        self.native.set_stream(side)
        graphs = []
        try:
            for p in range(2):
                self.native.capture_begin()
                try:
                    launch(p)
                finally:
                    g = self.native.capture_end()
                graphs.append(g)
        finally:
            self.native.set_stream(main)
        return graphs

    # Pipelined batched frames: the merged simulation step of frame f on the
    # batch context's branch beside the merged render passes of frame f-1
    # (HairRenderer's pipelined frame, for every instance at once).

    def build_pipelined(self, device_blocks: int | None = None) -> None:
        """Records both parities of every instance as a simulation part and a
        render part, merges each part over the instances and (with graphs)
        captures the pipelined frame per parity."""
        self._share_blocks(device_blocks)
        sims, rens = [[], []], [[], []]
        for r in self.renderers:
            for parity in range(2):
                a, b = record_parts(r)
                sims[parity].append(a)
                rens[parity].append(b)
        self.psim = [self.native.program_merge(sims[p]) for p in range(2)]
        self.prender = [self.native.program_merge(rens[p]) for p in range(2)]
        for lst in sims + rens:
            for prog in lst:
                prog.close()
        self.pbuild_current = self.renderers[0].hair.current
        self.pgraphs = None
        if self.use_graph:
            self.pgraphs = self._capture(self._pipelined_launch)
        self.pipelined = False

    def _pipelined_launch(self, p: int) -> None:
        self.native.branch_begin()
        self.psim[p].launch(self.native)
        self.native.branch_end()
        self.prender[p].launch(self.native)
        self.native.branch_join()

    def _pparity(self) -> int:
        parity = self.renderers[0].hair.current ^ self.pbuild_current
        assert all(r.hair.current ^ self.pbuild_current == parity for r in self.renderers), \
            "instances must advance together"
        return parity

    def _advance(self) -> None:
        for r in self.renderers:
            r.hair.swapPositionBuffersAfterSimIntegration()
            r.frameIdx += 1
        self.frameIdx += 1

    def pipeline_begin(self, upload_uniforms: bool = True) -> None:
        """The prologue: the simulation step of every instance alone."""
        if upload_uniforms:
            self.upload_sim_uniforms()
        self.psim[self._pparity()].launch(self.native)
        self._advance()
        self.pipelined = True

    def frame_pipelined(self, upload_uniforms: bool = True) -> None:
        """The simulation step of frame frameIdx beside the render passes of
        frame frameIdx - 1, for every instance."""
        assert self.pipelined, "pipeline_begin() first"
        if upload_uniforms:
            self.upload_sim_uniforms()
        p = self._pparity()
        if self.pgraphs is not None:
            self.pgraphs[p].launch()
        else:
            self._pipelined_launch(p)
        self._advance()

    def pipeline_end(self) -> None:
        """The epilogue: the render passes of the last simulated frame."""
        assert self.pipelined, "pipeline_begin() first"
        self.prender[self._pparity()].launch(self.native)
        self.pipelined = False

    def launches_per_frame(self) -> int:
        return self.programs[0].launches()

    def frame(self, upload_uniforms: bool = True) -> None:
        """One frame of every instance."""
        if upload_uniforms:
            self.upload_sim_uniforms()
        parity = self.renderers[0].hair.current ^ self.build_current
        assert all(r.hair.current ^ self.build_current == parity for r in self.renderers), \
            "instances must advance together"
        if self.graphs is not None:
            self.graphs[parity].launch()
        else:
            self.programs[parity].launch(self.native)
        for r in self.renderers:
            r.hair.swapPositionBuffersAfterSimIntegration()
            r.frameIdx += 1
        self.frameIdx += 1

    def images(self) -> list:
        return [r.finePass.image(r.wl.width, r.wl.height) for r in self.renderers]

    def memory_bytes(self) -> int:
        return sum(instance_bytes(r) for r in self.renderers)


def instance_bytes(r: PS.HairRenderer) -> int:
    """Device memory held by one instance's buffers (the library scratch
    excluded)."""
    ts = list(r.inputs.values()) + [v for v in vars(r).values() if isinstance(v, torch.Tensor)]
    for obj in (r.hair, r.grid, r.sdf, r.tilesPass, r.sortPass, r.finePass):
        for v in vars(obj).values():
            if isinstance(v, torch.Tensor):
                ts.append(v)
            elif isinstance(v, (list, tuple)):
                ts += [x for x in v if isinstance(x, torch.Tensor)]
    seen, total = set(), 0
    for t in ts:
        if t.is_cuda and t.data_ptr() not in seen:
            seen.add(t.data_ptr())
            total += t.numel() * t.element_size()
    return total


def default_processor_count(rc: L.RenderConfig, width: int, height: int) -> int:
    """Fine-pass processors: one per tile up to the configured count (a
    processor takes whole tiles, so more than the tiles would idle)."""
    tiles = -(-width // rc.tileSize) * -(-height // rc.tileSize)
    return max(1, min(rc.processorCount, tiles))

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
        if device_blocks is None:
            device_blocks = 8 * torch.cuda.get_device_properties(self.device).multi_processor_count
        pts = np.array([r.wl.strands * r.wl.points for r in self.renderers], dtype=np.float64)
        for r, w in zip(self.renderers, pts / pts.sum()):
            r.native.set_persistent_blocks(max(64, int(math.ceil(device_blocks * w))))
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
            main = torch.cuda.current_stream(self.device)
            side = torch.cuda.Stream(self.device)
            side.wait_stream(main)
            torch.cuda.synchronize(self.device)
            self.native.set_stream(side)
            try:
                self.graphs = []
                for p in range(2):
                    self.native.capture_begin()
                    try:
                        self.programs[p].launch(self.native)
                    finally:
                        g = self.native.capture_end()
                    self.graphs.append(g)
            finally:
                self.native.set_stream(main)

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

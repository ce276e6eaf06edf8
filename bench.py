"""Benchmark of the per-frame hair compute chain on one MI355X.

A step is one frame of the chain on the 1080p character-scale workload
(10,000 strands x 32 points, 64^3 grid, 64^3 SDF, 1920x1080, 16x16 tiles,
8 depth bins x 4 slices, 8 shading points): the simulation step (grid
pre-pass, integration, particle-to-grid), then the tile pass, the tile sort,
the fine pass and the shading pass, with inputs resident in HBM.

By default frames are pipelined and replayed as HIP graphs (one graph per
parity of the tick-tock position buffers), in two forms: each frame runs the
simulation step of frame f on a second stream beside the render passes of
frame f-1 (`pipelined`), or the simulation step of frame f, the tile and
sort passes of frame f-1 and the fine, combine/present and shading passes of
frame f-2 on three streams at once (`deep`); either way every frame does one
simulation step and one render of the chain, and the images are the plain
frames'. `value` is the faster form's (named by `frame_form`; `forms` reports
both). Prints one JSON line. `value` comes from a timed region
of --steps frames with no events between them, the frame latencies from a
second region of as many frames with events at the frame boundaries; the
latency from an image's simulation step to its present, and the rate of one
graph per frame without pipelining (`unpipelined`), are reported beside
them. `roofline` is for the kernel that takes the most time in a
second timed region of as many frames, from its algorithmic bytes per launch
and its average launch duration, measured with HIP events that the library
records around each of its kernels on its own stream (per-kernel events cost
frame time, so they are kept out of the first region). `cpu_baseline`
is the CPU oracle (a scalar C restatement of the reference's kernels) timed
on full frames of the same workload, rank 0 only.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sim+raster frames/s at 1080p, batched vs plain; p50/p99 frame latency"
HBM_PEAK_GBS = 8000.0
# wave64 f32 VALU instructions per second at full rate: 1,024 SIMDs (256 CUs)
# at the 2.4 GHz peak engine clock, one non-packed instruction per 4 cycles
VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4


def kernel_bytes(wl, n_pairs: int, tiles: int) -> dict:
    """Compulsory HBM bytes of one launch of each main kernel: every distinct
    input byte read once, every output byte written once (re-reads, atomics
    and scratch excluded)."""
    N = wl.strands * wl.points
    d3 = wl.sim.gridDims ** 3
    V = wl.sdf_dims ** 3
    WH = wl.width * wl.height
    B = wl.render.tileDepthBins
    S, Ps = wl.strands, wl.render.shadingPoints
    return {
        # prev + now + initial (16 B each), rest length, write of the new position
        "sim_integrate": 68 * N + 32 * d3 + 4 * V,
        "grid_pre_sim": 32 * d3 + 4 * V,
        # P2G: binning reads each position once per pass and writes a pair index;
        # the accumulation reads positions and pair indices, and adds the grid once
        "p2g_clear": 16 * d3,
        "p2g_count": 16 * N,
        "p2g_scatter": 20 * N,
        "grid_post_sim": 36 * N + 16 * d3,
        # positions + tangents, mesh depth, list element + slot tag per pair, tile headers
        "tiles_draw": 32 * N + 4 * WH + 12 * n_pairs + 16 * tiles * B,
        # slot tag + list element read, key written per pair; per range the
        # scan's end and the list length read and the start written
        "bin_scatter": 12 * n_pairs + 12 * tiles * B,
        # every range header and start read; each key read and written once
        "bin_sort": 8 * n_pairs + 20 * tiles * B,
        # tile headers, sorted keys, positions + tangents, mesh depth, shading
        # points, framebuffer read + write, tile list
        "fine_raster": 16 * tiles * B + 4 * n_pairs + 32 * N + 4 * WH + 16 * S * Ps + 32 * WH + 4 * tiles,
        "shading": 32 * N + 4 * WH + WH // 2 + 2048 * 2048 * 4 + 16 * S * Ps,
        # HDR scene restored into the HDR texture (8 B read + 8 B written per texel)
        "copy_u32": 16 * WH,
        # rasterizer pixel read, HDR texel read and written
        "hair_combine": 32 * WH,
        # HDR texel read, rgba8 texel written
        "present": 12 * WH,
        # rasterizer pixel and scene texel read, HDR and rgba8 texels written
        "combine_present": 36 * WH,
    }


def stage_bytes(wl, n_pairs: int, tiles: int) -> dict:
    """Compulsory HBM bytes per frame per stage."""
    N = wl.strands * wl.points
    d = wl.sim.gridDims
    V = wl.sdf_dims ** 3
    WH = wl.width * wl.height
    B = wl.render.tileDepthBins
    S, Ps = wl.strands, wl.render.shadingPoints
    return {
        "sim": 100 * N + 96 * d ** 3 + 8 * V,
        "tiles": 32 * N + 4 * WH + 24 * n_pairs + 16 * tiles * B + 4 * tiles,
        "sort": 12 * tiles + 1024,
        "fine": 16 * tiles * B + 8 * n_pairs + 32 * N + 4 * WH + 16 * S * Ps + 32 * WH + 4 * tiles,
        "shading": 32 * N + 4 * WH + WH // 2 + 2048 * 2048 * 4 + 16 * S * Ps,
        "post": 36 * WH,
    }


def full_frame(wl, r0, steps: int, dev, branches: bool = True, mesh_passes: bool = True) -> dict:
    """Frames of the same workload with the passes that feed the shading pass
    computed on the device every frame (ShadowMapPass's hair draw, HwHairPass,
    AoPass) and, with mesh_passes, the background, the meshes' shadow map and
    the meshes' depth, normals and colour (DrawBackgroundGradientPass,
    ShadowMapPass's mesh draw, DrawMeshesPass): every pass of the reference's
    renderer; replayed as HIP graphs."""
    import torch
    from frostbitten_hair_webgpu_amd import passes as PS
    inputs = {k: v.cpu().numpy() for k, v in r0.inputs.items()}
    r = PS.HairRenderer(wl, dev, inputs=inputs, hw_passes=True, branches=branches, mesh_passes=mesh_passes)
    ring = torch.from_numpy(np.concatenate([r.sim_uniforms(f).view(np.uint8).reshape(-1)
                                            for f in range(2 * steps + 4)])).to(dev).view(2 * steps + 4, -1)
    r.use_uniform_ring(ring, r.frameIdx)
    r.beforeFirstFrame()
    r.frame(upload_uniforms=False)
    torch.cuda.synchronize()
    pipelined = branches
    if branches:  # pipelined frames, as the main measurement
        r.pipeline_begin(upload_uniforms=False)
        r.frame_pipelined(upload_uniforms=False)
        r.capture_pipelined()
    else:
        r.capture_frames()

    def step():
        if branches:
            r.frame_pipelined_graph(upload_uniforms=False)
        else:
            r.frame_graph(upload_uniforms=False)

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    if branches:
        r.pipeline_end()
        branches = False  # the kernel timing below replays one graph per frame
    r.native.profiler_enable(64 * steps)
    r.capture_frames()
    torch.cuda.synchronize()
    r.native.profiler_read()
    for _ in range(steps - 1):
        step()
    torch.cuda.synchronize()
    kernels = r.native.profiler_read()
    r.native.profiler_enable(0)
    extra = ("hair_raster", "hair_raster_big", "hair_resolve", "gtao", "background", "mesh_raster",
             "mesh_raster_big", "mesh_resolve")
    passes = ("chain + ShadowMapPass hair draw (2048^2) + HwHairPass (depth + normals) + AoPass (GTAO, "
              "half resolution)")
    if mesh_passes:
        passes = ("chain + DrawBackgroundGradientPass + ShadowMapPass (meshes and hair, 2048^2) + DrawMeshesPass "
                  f"(depth, normals, PBR colour; {r.meshes.triangles()} triangles) + HwHairPass + AoPass")
    return {"value": round(1e3 / ms, 3), "unit": "frames/s", "ms_per_step": round(ms, 4), "steps": steps,
            "passes": passes, "pipelined": pipelined,
            "kernel_us": {k: round(t / n * 1e3, 2) for k, (n, t) in sorted(kernels.items(), key=lambda kv: -kv[1][1])
                          if k in extra or k in ("copy_u32", "fill_u32")}}


def read_valu_instructions(kernel: str, workload: str):
    """VALU instructions per launch of `kernel` (waves x SQ_INSTS_VALU per
    wave) from the rocprofv3 counter table committed under profiles/ (None
    when absent)."""
    path = os.path.join(ROOT, "profiles", f"counters_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) > 5 and cells[0] == kernel:
            try:
                return float(cells[3]) * float(cells[5])
            except ValueError:
                return None
    return None


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary committed under profiles/ (None when absent)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return int(float(line.split()[1]))
    return None


VULKAN_ICD_DIRS = ("/usr/share/vulkan/icd.d", "/etc/vulkan/icd.d")


def reference_cpu_status() -> str:
    """Why the reference's own CPU path (its WGSL under Deno on lavapipe, a
    software Vulkan adapter) is not timed, from what this machine has."""
    import shutil
    deno = shutil.which("deno")
    icds = sorted(f for d in VULKAN_ICD_DIRS if os.path.isdir(d) for f in os.listdir(d) if f.endswith(".json"))
    cache = os.environ.get("DENO_DIR") or os.path.expanduser("~/.cache/deno")
    return ("unavailable: the reference's CPU path is its WGSL under Deno on a software Vulkan adapter "
            f"(lavapipe); this machine has {'deno at ' + deno if deno else 'no deno binary'}, "
            f"{'Vulkan ICDs ' + ', '.join(icds) if icds else 'no Vulkan ICD'} (in {', '.join(VULKAN_ICD_DIRS)}) "
            f"and {'a Deno module cache at ' + cache if os.path.isdir(cache) else 'no Deno module cache'}; "
            "the reference's modules are remote and nothing can be fetched offline")


def host_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(wl, r, frames: int = 8, budget_s: float = 15.0) -> dict:
    """Times the oracle on full frames of the workload on every host core:
    up to `frames` frames, stopping once `budget_s` seconds are spent. The
    oracle's independent loops (strands, grid nodes, tiles) run on one
    thread per core; its results do not depend on the thread count."""
    import __graft_entry__ as G
    O = G.load_oracle()
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import layouts as L
    threads = host_threads()
    O.set_threads(threads)
    S, P = wl.strands, wl.points
    i = {k: np.ascontiguousarray(v.cpu().numpy()) for k, v in r.inputs.items()}
    d = A.derived(wl)
    sl, tan = d["seglen"], d["tangents"]
    hd = L.pack_hair_data(r.hair.sphere[0], r.hair.sphere[1], S, P)
    cap = r.tilesPass.segmentsCapacity
    prev = wl.positions.copy()
    now = wl.positions.copy()
    shading = A.initial_shading(S, wl.render.shadingPoints)
    dv = O.grid_post_sim(wl.sim.gridDims, r.sim_uniforms(0), prev, now, S, P)
    times = []
    try:
        for f in range(frames):
            t0 = time.perf_counter()
            su = r.sim_uniforms(f)
            dgw = O.grid_pre_sim(wl.sim.gridDims, su, dv, wl.sdf, wl.sdf_dims)
            nxt = O.sim_integrate(wl.sim.gridDims, su, prev, now, sl, wl.sdf, wl.sdf_dims, dv, dgw, wl.positions,
                                  S, P)
            prev, now = now, nxt
            dv = O.grid_post_sim(wl.sim.gridDims, su, prev, now, S, P)
            tiles, seg, counts, _ = O.tiles_draw(wl.render, wl.width, wl.height, r.render_uniforms, hd, now, tan,
                                                 i["meshDepth"], cap, S, P)
            tl, _ = O.tiles_sort(wl.render, wl.width, wl.height, counts)
            O.fine_raster(wl.render, wl.width, wl.height, r.render_uniforms, now, tiles, seg, cap, i["meshDepth"],
                          shading, tan, tl, S, P, canonical=False)
            shading = O.shading(wl.render, r.render_uniforms, now, tan, i["shadowMap"], i["ao"], i["hairDepth"],
                                S, P).reshape(-1, 4)
            times.append(time.perf_counter() - t0)
            if sum(times) > budget_s:
                break
    finally:
        O.set_threads(1)
    frames = len(times)
    t = float(np.median(times))
    return {"value": round(1.0 / t, 4), "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"{frames} full frames (sim step + tile + sort + fine + shading) of the same workload, "
                      f"median {t:.3f} s/frame, the C oracle (a restatement of the reference's kernels) with its "
                      f"strand, grid-node and tile loops on {threads} host threads",
            "reference": reference_cpu_status()}


def batched_vs_plain(n: int, steps: int, warmup: int, tune: bool, dev, log, gap: bool = False,
                     instances: tuple | None = None) -> dict:
    """The batched form against plain runs on the same instances: `n`
    instances of uneven sizes (strands log-uniform in [1k, 50k], points from
    {8, 16, 24, 32}, resolutions from {512^2, 1280x720, 1920x1080}, each with
    its own camera, wind and 64^3 SDF), render configuration from the
    autotuner. Plain: each instance replays its own frame graph, one instance
    after another. Batched: one graph per frame runs every kernel of the chain
    once for all instances. Both are timed over `steps` rounds (one frame of
    every instance) with the per-frame simulation uniforms already on the
    device.

    gap: also time each instance's plain frame alone and, for the instances
    that take the larger half of the plain round, the batched frame of that
    half alone: with the other instances free, batching could not beat
    (whole plain round) / (batched frame of that half). `instances` =
    (workloads, inputs) skips generation and tuning (the workloads keep the
    configuration applied to them)."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import autotune as AT
    from frostbitten_hair_webgpu_amd import batch as B
    from frostbitten_hair_webgpu_amd import passes as PS
    t0 = time.perf_counter()
    if instances is not None:
        wls, inputs = instances
        n = len(wls)
        tune = False
    else:
        wls = [A.make_instance(i) for i in range(n)]
        inputs = [PS.make_inputs(wl, PS.workload_render_uniforms(wl), 1024) for wl in wls]
    t1 = time.perf_counter()
    log(f"# batch: {n} instances generated in {t1 - t0:.1f} s")
    if instances is not None:
        choice = None
    elif tune:
        choice = AT.tune(wls, inputs, log=log)
    else:
        choice = [(16, 8, 4)] * n
        for wl in wls:
            AT.apply(wl, choice[0])
    t2 = time.perf_counter()
    log(f"# batch: tuned in {t2 - t1:.1f} s")
    br = B.BatchRenderer(wls, dev, inputs=inputs)
    br.beforeFirstFrame()
    for r in br.renderers:  # every call once (scratch sizes), then the plain graphs
        r.frame()
    for r in br.renderers:
        r.capture_frames()
    br.frameIdx = br.renderers[0].frameIdx
    br.build()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    log(f"# batch: built in {t3 - t2:.1f} s, {br.launches_per_frame()} launches per batched frame")
    total = 6 * (warmup + steps) + 16
    ring = torch.from_numpy(np.stack([br.sim_uniforms(br.frameIdx + f) for f in range(total)])).to(dev)

    def plain_round(f):
        for i, r in enumerate(br.renderers):
            r.simulationUniforms.copy_(ring[f, i])
            r.frame_graph(upload_uniforms=False)

    def batch_frame(f):
        br.simPacked.copy_(ring[f])
        br.frame(upload_uniforms=False)

    def timed(fn, first):
        for k in range(warmup):
            fn(first + k)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        torch.cuda.synchronize()
        c0 = time.perf_counter()
        ev[0].record()
        for k in range(steps):
            fn(first + warmup + k)
            ev[k + 1].record()
        torch.cuda.synchronize()
        el = time.perf_counter() - c0
        per = np.array([ev[k].elapsed_time(ev[k + 1]) for k in range(steps)])
        return el, per

    el_p, per_p = timed(plain_round, 0)
    el_b, per_b = timed(batch_frame, warmup + steps)
    k_end = 2 * (warmup + steps)  # ring rows used so far, before the kernel timing region
    # per-kernel times of the batched frame: its launches issued directly
    # with HIP events around each (a third region, kept out of the timing above)
    graphs, br.graphs = br.graphs, None
    br.native.profiler_enable(64 * 4)
    br.native.profiler_read()
    for k in range(4):
        batch_frame(k_end + k)
    torch.cuda.synchronize()
    kernels = br.native.profiler_read()
    br.native.profiler_enable(0)
    br.graphs = graphs
    f0 = k_end + 4

    # Pipelined: each instance's simulation step of frame f beside its render
    # passes of frame f-1. Plain: every instance replays its own pipelined
    # graphs, one after another; batched: one pipelined graph per frame.
    def plain_pipelined_round(f):
        for i, r in enumerate(br.renderers):
            r.simulationUniforms.copy_(ring[f, i])
            r.frame_pipelined_graph(upload_uniforms=False)

    def batch_pipelined_frame(f):
        br.simPacked.copy_(ring[f])
        br.frame_pipelined(upload_uniforms=False)

    for i, r in enumerate(br.renderers):
        r.simulationUniforms.copy_(ring[f0, i])
        r.pipeline_begin(upload_uniforms=False)
        r.simulationUniforms.copy_(ring[f0 + 1, i])
        r.frame_pipelined(upload_uniforms=False)  # every call once before the capture
        r.capture_pipelined()
    el_pp, per_pp = timed(plain_pipelined_round, f0 + 2)
    for r in br.renderers:
        r.pipeline_end()
    f1 = f0 + 2 + warmup + steps
    br.build_pipelined()
    br.simPacked.copy_(ring[f1])
    br.pipeline_begin(upload_uniforms=False)
    el_bp, per_bp = timed(batch_pipelined_frame, f1 + 1)
    br.pipeline_end()
    torch.cuda.synchronize()
    S = np.array([w.strands for w in wls])
    P = np.array([w.points for w in wls])
    plain_fps = n * steps / el_p
    batch_fps = n * steps / el_b
    plain_pp_fps = n * steps / el_pp
    batch_pp_fps = n * steps / el_bp

    def ms(el, per):
        return {"mean": round(el / steps * 1e3, 3), "p50": round(float(np.percentile(per, 50)), 3),
                "p99": round(float(np.percentile(per, 99)), 3)}

    best_plain = max(plain_fps, plain_pp_fps)
    gap_out = None
    if gap:
        # each instance's plain frame alone (events between the instances of
        # a round; one stream, so the durations add up to the round)
        f2 = f1 + 2 + warmup + steps
        rounds = 5
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(n + 1)] for _ in range(rounds)]
        for q in range(rounds):
            ev[q][0].record()
            for i, r in enumerate(br.renderers):
                r.simulationUniforms.copy_(ring[f2 + q, i])
                r.frame_graph(upload_uniforms=False)
                ev[q][i + 1].record()
        torch.cuda.synchronize()
        t_inst = np.median([[ev[q][i].elapsed_time(ev[q][i + 1]) for i in range(n)] for q in range(rounds)], axis=0)
        order = np.argsort(-t_inst)
        cum = np.cumsum(t_inst[order])
        k = int(np.searchsorted(cum, 0.5 * cum[-1])) + 1  # the largest instances, half of the plain round
        big = sorted(order[:k].tolist())
        sub = batched_vs_plain(k, steps, warmup, False, dev, log,
                               instances=([wls[i] for i in big], [inputs[i] for i in big]))
        round_ms = float(t_inst.sum())
        big_batch_ms = min(sub["batch_frame_ms"]["mean"], sub["batch_unpipelined_frame_ms"]["mean"])
        gap_out = {
            "plain_instance_ms": {"sum": round(round_ms, 3), "median": round(float(np.median(t_inst)), 3),
                                  "max": round(float(t_inst.max()), 3), "min": round(float(t_inst.min()), 3)},
            "largest_half": {"instances": k, "plain_ms": round(float(t_inst[big].sum()), 3),
                             "batched_frame_ms": big_batch_ms,
                             "speedup_vs_plain": sub["speedup_vs_plain"],
                             "strands": int(sum(wls[i].strands for i in big)),
                             "pixels": int(sum(wls[i].width * wls[i].height for i in big))},
            "rest": {"instances": n - k, "plain_ms": round(float(round_ms - t_inst[big].sum()), 3)},
            "speedup_bound": round(round_ms / big_batch_ms, 3),
            "bound": "plain round (sum of the instances' plain frames) / batched frame of the largest half "
                     "alone: the speedup if batching made every other instance free",
        }
    return {
        "instances": n, "steps": steps,
        "value": round(batch_pp_fps, 2), "unit": "instance-frames/s",
        "form": "batched, pipelined (merged simulation step of frame f beside the merged render of frame f-1)",
        "plain_value": round(best_plain, 2),
        "speedup_vs_plain": round(batch_pp_fps / best_plain, 3),
        "speedup_vs": "the faster plain form (each instance replaying its own graphs, one after another)",
        "forms_instance_frames_per_s": {"plain": round(plain_fps, 2), "plain_pipelined": round(plain_pp_fps, 2),
                                        "batched": round(batch_fps, 2), "batched_pipelined": round(batch_pp_fps, 2)},
        "batch_frame_ms": ms(el_bp, per_bp),
        "batch_unpipelined_frame_ms": ms(el_b, per_b),
        "plain_round_ms": ms(el_p, per_p),
        "plain_pipelined_round_ms": ms(el_pp, per_pp),
        "launches_per_batched_frame": br.launches_per_frame(),
        "kernel_us": {k: round(t / c * 1e3, 1) for k, (c, t) in sorted(kernels.items(), key=lambda kv: -kv[1][1])},
        "strands_total": int(S.sum()), "points_total": int((S * P).sum()),
        "pixels_total": int(sum(w.width * w.height for w in wls)),
        "particle_steps_per_s": round(float((S * P).sum()) * steps / el_bp, 1),
        "segments_rasterized_per_s": round(float((S * (P - 1)).sum()) * steps / el_bp, 1),
        "autotune": (AT.summary(choice) if tune else "off (tile 16, 8 bins x 4 slices)") if instances is None
        else "as configured by the caller",
        "device_bytes": br.memory_bytes(),
        "setup_s": round(time.perf_counter() - t0, 1),
        **({"gap": gap_out} if gap_out is not None else {}),
    }


def served(n_req: int, rate: float, budget_gib: float, dev, log) -> dict:
    """The served form against sequential serving of the same request stream:
    `n_req` Poisson arrivals at `rate` per second, each for an asset of the
    256-asset pool (strands log-uniform in [1k, 100k], 16-32 points, 720p /
    1080p / 4K) and U{1..120} frames; instance state stays in HBM between
    requests, within a `budget_gib` GiB budget. Continuous batching runs one
    merged frame of every admitted request per tick (pipelined: the merged
    simulation steps beside the merged renders of the previous steps);
    sequential serving runs
    one request at a time through the same loop. Asset data (strands and the
    precomputed depth / AO / shadow inputs) is generated on the host first;
    instance creation on the device is part of serving."""
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    from frostbitten_hair_webgpu_amd import serve as SV
    t0 = time.perf_counter()
    reqs = SV.poisson_requests(n_req, rate, list(range(256)), 120, seed=A.SERVE_SEED)
    pool = {}
    for a in sorted({q.asset for q in reqs}):
        wl = A.make_pool_asset(a)
        pool[a] = (wl, PS.make_inputs(wl, PS.workload_render_uniforms(wl), 1024))
    log(f"# serve: {len(pool)} assets generated in {time.perf_counter() - t0:.1f} s")
    out = {"requests": n_req, "rate_per_s": rate, "assets": len(pool), "budget_GiB": budget_gib}
    for mode, max_active, pipelined in (("continuous_pipelined", 64, True), ("continuous", 64, False),
                                        ("sequential", 1, False)):
        srv = SV.HairServer(pool, int(budget_gib * (1 << 30)), max_active=max_active, device=dev,
                            pipelined=pipelined)
        rs = srv.run([dataclasses.replace(q) for q in reqs])
        out[mode] = dict(SV.summarize(rs, srv.elapsed), **{k: v for k, v in srv.stats.items()},
                         host_s={k: round(v, 3) for k, v in srv.times.items()})
        srv.close()
        del srv
        log(f"# serve {mode}: {out[mode]['frames_per_s']} frames/s, p50 {out[mode]['latency_s']['p50']} s")
    best = max(out["continuous"]["frames_per_s"], out["continuous_pipelined"]["frames_per_s"])
    out["speedup_vs_sequential"] = round(best / out["sequential"]["frames_per_s"], 3)
    # latency at loads the device sustains: the same request mix (assets and
    # frame counts) arriving at 25%, 50% and 90% of the measured capacity in
    # requests per second (the continuous rate / mean frames per request)
    mean_frames = out["continuous"]["frames"] / n_req
    capacity = best / mean_frames
    out["capacity_requests_per_s"] = round(capacity, 3)
    sweep = []
    for load in (0.25, 0.5, 0.9):
        lam = load * capacity
        qs = SV.poisson_requests(n_req, lam, list(range(256)), 120, seed=A.SERVE_SEED)
        srv = SV.HairServer(pool, int(budget_gib * (1 << 30)), max_active=64, device=dev, pipelined=True)
        rs = srv.run(qs)
        row = dict(SV.summarize(rs, srv.elapsed), rate_per_s=round(lam, 3), offered_load=load,
                   ticks=srv.stats.get("ticks"))
        srv.close()
        del srv
        sweep.append(row)
        log(f"# serve at {load:.0%} of capacity ({lam:.2f}/s): p50 {row['latency_s']['p50']} s, "
            f"p99 {row['latency_s']['p99']} s")
    out["latency_vs_load"] = sweep
    out["speedup_vs"] = ("the faster continuous form (pipelined ticks: the merged simulation step beside the "
                         "merged render of the previous step) over one request at a time")
    return out


FORM_TEXT = {
    "pipelined": "pipelined: simulation step of frame f beside the render passes of frame f-1; shading beside "
                 "the combine/present pass",
    "deep": "deep-pipelined: simulation step of frame f, tile and sort passes of frame f-1, fine, "
            "combine/present and shading passes of frame f-2, on three streams at once",
}


def plain_run(kind: str, steps: int, warmup: int, args, dev) -> tuple:
    """Frames of the plain drop-in on workload `kind`: the timed regions, the
    per-kernel durations and the roofline of the dominant kernel. Returns the
    result dict, the renderer and the workload."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    t_setup = time.perf_counter()
    wl = A.make_workload(kind)
    r = PS.HairRenderer(wl, dev, branches=not args.no_branches)
    total = warmup + 7 * steps + 8  # seven timed regions, pipeline prologues and sizing frames
    # per-frame simulation uniforms (frameIdx changes), uploaded once up front
    ring = torch.from_numpy(np.concatenate([r.sim_uniforms(f).view(np.uint8).reshape(-1) for f in range(total)]))
    ring = ring.to(dev).view(total, -1)
    # each simulation step copies its row on the device (fh_copy_indexed), so
    # no host call sits between two frames' graphs
    r.use_uniform_ring(ring, r.frameIdx)
    print(f"# setup {time.perf_counter() - t_setup:.1f} s, workload {wl.name}: {wl.strands}x{wl.points} "
          f"{wl.width}x{wl.height}", file=sys.stderr, flush=True)

    stages = ["sim", "tiles", "sort", "fine", "post", "shading"]

    def frame(events=None):
        ctx = r.ctx()
        rec = (lambda k: events[k].record()) if events else (lambda k: None)
        rec(0)
        r.simulateHair(ctx)
        rec(1)
        r.tilesPass.cmdClearBeforeRender(ctx)
        r.finePass.cmdClearBeforeRender(ctx)
        r.sortPass.cmdClearBeforeRender(ctx)
        r.tilesPass.cmdDrawHairToTiles(ctx, r.hair)
        rec(2)
        r.sortPass.cmdSortHairTiles(ctx)
        rec(3)
        r.finePass.cmdRasterizeSlicesHair(ctx, r.hair, r.tilesPass)
        rec(4)
        r.hairCombinePass.cmdCombineAndPresent(ctx, r.presentPass, r.screenTexture, r.inputs["hdrScene"])
        rec(5)
        r.shadingPass.cmdComputeShadingPoints(ctx, r.hair)
        r.join_branch()  # the grid accumulation ran beside the render passes
        rec(6)
        r.frameIdx += 1

    def graph_frame(events=None):
        if events:
            events[0].record()
        r.frame_graph(upload_uniforms=False)
        if events:
            events[6].record()

    def pipelined_frame(events=None):
        if events:
            events[0].record()
        r.frame_pipelined_graph(upload_uniforms=False)
        if events:
            events[6].record()

    def deep_frame(events=None):
        if events:
            events[0].record()
        r.frame_deep_graph(upload_uniforms=False)
        if events:
            events[6].record()

    def region(step, events=True):
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(7)] for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            step(evs[k] if events else None)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, evs

    def timed_pair(step):
        """Wall time of `steps` frames with no events between them (value),
        then as many frames with an event at every frame boundary (latency)."""
        el, _ = region(step, events=False)
        _, evs = region(step, events=True)
        return el, evs

    use_graph = not args.no_graph
    pipelined = use_graph and not args.no_branches and not args.no_pipeline
    r.beforeFirstFrame()
    for _ in range(warmup):
        frame()
    torch.cuda.synchronize()
    step = graph_frame if use_graph else frame

    # Timed region 1: the frames alone (events only at frame boundaries) give
    # value and the frame latencies. Pipelined, each frame runs the simulation
    # step of frame f beside the render passes of frame f-1.
    unpipelined = None
    forms = {}
    if pipelined:
        r.pipeline_begin(upload_uniforms=False)
        r.frame_pipelined(upload_uniforms=False)  # every call once before capture
        r.capture_pipelined()
        el_p, evs_p = timed_pair(pipelined_frame)
        r.pipeline_end()
        # sim-to-present: from the start of the frame that ran an image's
        # simulation step to the end of the frame that presented it
        forms["pipelined"] = (el_p, evs_p, np.array([evs_p[k - 1][0].elapsed_time(evs_p[k][6])
                                                     for k in range(1, steps)]))
        # deep-pipelined: simulation step f, tile stage f-1 and fine stage f-2
        # at once (three streams)
        r.deep_pipeline_begin(upload_uniforms=False)
        r.frame_deep(upload_uniforms=False)  # every call once before capture
        r.capture_deep()
        el_d, evs_d = timed_pair(deep_frame)
        r.deep_pipeline_end()
        forms["deep"] = (el_d, evs_d, np.array([evs_d[k - 2][0].elapsed_time(evs_d[k][6])
                                                for k in range(2, steps)]))
        form = min(forms, key=lambda k: forms[k][0])
        elapsed, evs, s2p = forms[form]
        r.capture_frames()
        el_u, evs_u = timed_pair(graph_frame)
        per_u = np.array([evs_u[k][0].elapsed_time(evs_u[k][6]) for k in range(steps)])
        unpipelined = {"value": round(steps / el_u, 3), "ms_per_step": round(el_u / steps * 1e3, 4),
                       "latency_ms": {"p50": round(float(np.percentile(per_u, 50)), 4),
                                      "p99": round(float(np.percentile(per_u, 99)), 4)},
                       "frame": "one graph per frame: simulation step, then the render passes"}
    else:
        if use_graph:
            r.capture_frames()
            elapsed, evs = timed_pair(step)
        else:  # eager frames: the events also time the stages
            elapsed, evs = region(step)

    # Timed region 2: as many further frames with HIP events around every
    # kernel on the library's stream (in graph replays, event-record nodes
    # around every kernel node): per-kernel durations and the roofline. The
    # events cost frame time, so they stay out of region 1.
    kernels = {}
    evs2 = [[torch.cuda.Event(enable_timing=True) for _ in range(7)] for _ in range(steps)]
    if not args.no_kernel_timing:
        r.native.profiler_enable(64 * steps)
        if use_graph:
            r.capture_frames()
        torch.cuda.synchronize()
        r.native.profiler_read()
        for k in range(steps):
            step(evs2[k])
        torch.cuda.synchronize()
        kernels = r.native.profiler_read()
        r.native.profiler_enable(0)
    per_frame = np.array([evs[k][0].elapsed_time(evs[k][6]) for k in range(steps)])  # ms
    if use_graph:
        # stage sums from the per-kernel events
        groups = {"sim": ("grid_pre_sim", "sim_integrate", "p2g_clear", "p2g_count", "p2g_scan", "p2g_scatter",
                         "grid_post_sim"),
                  "tiles": ("fill_many", "tile_depth_bounds", "tiles_draw", "tiles_pieces", "tile_counts"),
                  "sort": ("sort_tiles",),
                  "fine": ("fill_u32", "bin_scan_local", "bin_scan_add", "bin_scatter", "bin_sort", "fine_raster"),
                  "post": ("copy_u32", "hair_combine", "present", "combine_present"),
                  "shading": ("shading",)}
        stage_ms = {g: sum(kernels[k][1] for k in ks if k in kernels) / steps for g, ks in groups.items()}
        if not kernels:
            stage_ms = {}
    elif kernels:
        stage_ms = {s: float(np.mean([evs2[k][i].elapsed_time(evs2[k][i + 1]) for k in range(steps)]))
                    for i, s in enumerate(stages)}
    else:
        stage_ms = {}
    # drawnSegmentsCount counts every pair; the list keeps the first
    # `capacity` of them (hairTilesPass.wgsl.ts:221-237) and the bytes below
    # count only the kept ones
    drawn_pairs = int(r.tilesPass.hairTileSegmentsBuffer[0].item())
    capacity = r.tilesPass.segmentsCapacity
    n_pairs = min(drawn_pairs, capacity)
    tiles = (wl.width + wl.render.tileSize - 1) // wl.render.tileSize * \
            ((wl.height + wl.render.tileSize - 1) // wl.render.tileSize)
    drawn_tiles = int(r.sortPass.tileListBuffer[0].item())
    kb = kernel_bytes(wl, n_pairs, tiles)
    kernel_ms = {k: t / n for k, (n, t) in kernels.items()}
    timed = [k for k in kernel_ms if k in kb]
    dom = max(timed, key=lambda k: kernels[k][1]) if timed else None
    achieved = kb[dom] / (kernel_ms[dom] * 1e-3) / 1e9 if dom else None
    traffic = read_traffic(dom, wl.name) if dom else None
    rooflines = {}
    for k in ("fine_raster", "tiles_draw"):
        if k in kernel_ms:
            a = kb[k] / (kernel_ms[k] * 1e-3) / 1e9
            rooflines[k] = {"achieved": round(a, 2), "frac": round(a / HBM_PEAK_GBS, 5),
                            "avg_launch_us": round(kernel_ms[k] * 1e3, 2), "algorithmic_bytes": int(kb[k]),
                            "traffic": read_traffic(k, wl.name)}
            # the issue bound these kernels sit against: their instructions at
            # full rate (transcendentals take four times as long, so this is
            # a lower bound of the VALU time and frac a lower bound of the
            # VALU's use)
            n_valu = read_valu_instructions(k, wl.name)
            if n_valu:
                bound_us = n_valu / VALU_ISSUE_PER_S * 1e6
                rooflines[k]["valu"] = {"instructions": int(n_valu), "issue_bound_us": round(bound_us, 1),
                                        "frac": round(bound_us / (kernel_ms[k] * 1e3), 3)}
    img = r.finePass.image(wl.width, wl.height)
    assert torch.isfinite(img).all()
    exhausted_bins = r.native.fine_exhausted_bins()  # of the last frame's fine pass

    ms = elapsed / steps * 1e3
    out = {
        "metric": METRIC,
        "value": round(1e3 / ms, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded strands on a scalp cap, analytic colliders; depth/shadow/AO inputs precomputed)",
        "config": {"workload": f"plain drop-in{' replayed as a HIP graph' if use_graph else ''}, "
                               f"{wl.strands} strands x {wl.points} points, {wl.width}x{wl.height}, "
                               f"tile {wl.render.tileSize}, {wl.render.tileDepthBins} bins x "
                               f"{wl.render.slicesPerPixel} slices, grid {wl.sim.gridDims}^3, SDF {wl.sdf_dims}^3, "
                               f"{wl.sim.constraintIterations} constraint iterations",
                   "global_batch": 1, "parallelism": "none (1 GPU)",
                   "overlap": "none (one stream)" if args.no_branches else
                              (FORM_TEXT[form] if pipelined else
                               "grid accumulation beside tiles/sort/fine, shading beside the combine/present pass")},
        "latency_ms": {"p50": round(float(np.percentile(per_frame, 50)), 4),
                       "p99": round(float(np.percentile(per_frame, 99)), 4)},
        "pipelined": pipelined,
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "stage_hbm_GBps": {k: round(v / (stage_ms[k] * 1e-3) / 1e9, 2)
                           for k, v in stage_bytes(wl, n_pairs, tiles).items() if stage_ms.get(k)},
        "throughput": {"particle_steps_per_s": round(wl.strands * wl.points * 1e3 / ms, 1),
                       "segments_rasterized_per_s": round(wl.strands * (wl.points - 1) * 1e3 / ms, 1),
                       "segment_tile_pairs": drawn_pairs, "segment_list_capacity": capacity,
                       "dropped_pairs": max(0, drawn_pairs - capacity), "drawn_tiles": drawn_tiles,
                       "tile_coverage": round(drawn_tiles / tiles, 4),
                       "exhausted_bins": exhausted_bins},
        "kernel_us": {k: round(v * 1e3, 2) for k, v in sorted(kernel_ms.items(), key=lambda kv: -kernels[kv[0]][1])},
        "kernel_timing": f"a further timed region of {steps} frames with HIP events around every kernel "
                         f"(one graph per frame, kernels one after another)",
        "roofline": None if dom is None else {
            "bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "algorithmic_bytes": int(kb[dom]), "avg_launch_us": round(kernel_ms[dom] * 1e3, 2)},
        "rooflines": rooflines,
    }
    if pipelined:
        out["frame_form"] = form
        out["latency_ms"]["frame"] = "per pipelined frame (frame period)"
        out["latency_ms"]["sim_to_present_p50"] = round(float(np.percentile(s2p, 50)), 4)
        out["latency_ms"]["sim_to_present_p99"] = round(float(np.percentile(s2p, 99)), 4)
        out["forms"] = {}
        for k, (el, ev, sp) in forms.items():
            per = np.array([ev[i][0].elapsed_time(ev[i][6]) for i in range(steps)])
            out["forms"][k] = {"value": round(steps / el, 3), "ms_per_step": round(el / steps * 1e3, 4),
                               "latency_ms": {"p50": round(float(np.percentile(per, 50)), 4),
                                              "p99": round(float(np.percentile(per, 99)), 4),
                                              "sim_to_present_p50": round(float(np.percentile(sp, 50)), 4)},
                               "frame": FORM_TEXT[k]}
        out["unpipelined"] = unpipelined
    return out, r, wl


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2", choices=["small", "c2", "c3"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--c3-steps", type=int, default=20,
                    help="frames of the 4K, 100,000-strand workload timed beside the main line (0: skip it)")
    ap.add_argument("--full-frame", type=int, default=1,
                    help="also time frames that draw the background, the meshes, the shadow map, the hair depth / "
                         "normals and AO every frame, and frames with the meshes precomputed (0 to skip)")
    ap.add_argument("--batch-gap", action="store_true",
                    help="also time each instance's plain frame and the batched frame of the largest half alone")
    ap.add_argument("--batch-instances", type=int, default=64,
                    help="instances of the batched-versus-plain comparison (0: skip it)")
    ap.add_argument("--no-tune", action="store_true", help="batched comparison without the autotuner")
    ap.add_argument("--serve-requests", type=int, default=32,
                    help="requests of the served-versus-sequential comparison (0: skip it)")
    ap.add_argument("--serve-rate", type=float, default=40.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--serve-budget-gib", type=float, default=8.0, help="HBM budget of the resident instances")
    ap.add_argument("--no-graph", action="store_true", help="launch every call eagerly instead of replaying "
                    "the frame as a HIP graph")
    ap.add_argument("--no-branches", action="store_true",
                    help="run every pass of the frame on one stream (no overlap of the grid accumulation, "
                         "hardware-raster and shading passes with the render chain)")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="one graph per frame (simulation step then render passes) instead of pipelined frames "
                         "(the simulation step of frame f beside the render passes of frame f-1)")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="no per-kernel HIP events (roofline omitted); frame time only")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the hair chain has no multi-GPU path here (asked for --gpus "
              f"{args.gpus}). Sharded, the simulation would split strands (one int32 all-reduce of the grid "
              f"per step) and the render would split screen tiles.", file=sys.stderr)
        sys.exit(2)

    import torch

    import __graft_entry__ as G
    G.load_package()
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        print(json.dumps({"copy_probe_mib": args.copy_probe, "ms": ms,
                          "GB_per_s": 2 * n * 4 / (ms * 1e-3) / 1e9}), flush=True)
        return
    out, r, wl = plain_run(args.workload, args.steps, args.warmup, args, dev)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, r)
        out["speedup_vs_cpu_baseline"] = round(out["value"] / out["cpu_baseline"]["value"], 1)
    if args.full_frame:
        out["full_frame"] = full_frame(wl, r, min(args.steps, 30), dev, branches=not args.no_branches)
        out["full_frame_precomputed_meshes"] = full_frame(wl, r, min(args.steps, 30), dev,
                                                          branches=not args.no_branches, mesh_passes=False)
    del r
    if args.c3_steps and args.workload != "c3":
        # the 4K, 100,000-strand workload (the dense fine-pass case) beside the 1080p line
        torch.cuda.empty_cache()
        full, r3, wl3 = plain_run("c3", args.c3_steps, 3, args, dev)
        c3 = {k: full[k] for k in ("value", "unit", "ms_per_step", "latency_ms", "frame_form", "forms",
                                   "unpipelined", "throughput",
                                   "roofline", "rooflines", "stage_ms", "kernel_us") if k in full}
        c3["config"] = full["config"]["workload"]
        c3["steps"] = args.c3_steps
        if not args.no_cpu_baseline:
            c3["cpu_baseline"] = cpu_baseline(wl3, r3, frames=3, budget_s=12.0)
            c3["speedup_vs_cpu_baseline"] = round(c3["value"] / c3["cpu_baseline"]["value"], 1)
        out["c3"] = c3
        del r3
        torch.cuda.empty_cache()
    if args.batch_instances:
        out["batched"] = batched_vs_plain(args.batch_instances, min(args.steps, 20), 2, not args.no_tune, dev,
                                          lambda m: print(m, file=sys.stderr, flush=True), gap=args.batch_gap)
    if args.serve_requests:
        out["served"] = served(args.serve_requests, args.serve_rate, args.serve_budget_gib, dev,
                               lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Batched form: every instance computes in the merged launches exactly what
its plain frame computes (positions, tile lists, images), over several frames
and both parities of the position buffers, with and without the graph."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _workloads():
    """Uneven instances: strand and point counts, resolutions and all three
    tile sizes (so the merged frame has kernels that only some instances
    launch, e.g. the fine pass per tile size)."""
    from frostbitten_hair_webgpu_amd import assets as A
    specs = [dict(kind="small", strands=300, width=320, height=240, points=None, tile=16, bins=4, slices=4),
             dict(kind="c2", strands=900, width=400, height=300, points=8, tile=8, bins=8, slices=4),
             dict(kind="c2", strands=1500, width=512, height=512, points=24, tile=32, bins=4, slices=8),
             dict(kind="small", strands=64, width=160, height=120, points=None, tile=16, bins=16, slices=2)]
    out = []
    for i, sp in enumerate(specs):
        wl = A.make_workload(sp["kind"], seed=1000 + i, strands=sp["strands"], width=sp["width"],
                             height=sp["height"], points=sp["points"])
        wl.render.tileSize, wl.render.tileDepthBins, wl.render.slicesPerPixel = sp["tile"], sp["bins"], sp["slices"]
        out.append(wl)
    return out


def _state(r):
    """Everything a frame leaves behind. The list heads and element order of
    the tile lists follow the order of the atomic appends, so the lists are
    compared as per-(tile, bin) sets (with the depth range and count)."""
    from gpu_util import walk_lists
    tiles = _np(r.tilesPass.hairTilesBuffer).reshape(-1, 4)
    seg = _np(r.tilesPass.hairTileSegmentsBuffer).view(np.uint32)
    lists = walk_lists(tiles.view(np.uint32), seg, r.tilesPass.segmentsCapacity)
    return dict(now=_np(r.hair.positionsNow), prev=_np(r.hair.positionsPrev), dv=_np(r.grid.densityVelocity),
                depth=tiles[:, [0, 1, 3]], lists=sorted(lists.items()), drawn=int(seg[0]),
                counts=_np(r.tilesPass.segmentCountPerTileBuffer),
                img=_np(r.finePass.image(r.wl.width, r.wl.height)), shading=_np(r.hair.shading),
                hdr=_np(r.hdrRenderTexture), screen=_np(r.screenTexture))


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_plain(gpu, use_graph):
    import torch
    from frostbitten_hair_webgpu_amd import batch as B
    from frostbitten_hair_webgpu_amd import passes as PS
    wls = _workloads()
    plain = [PS.HairRenderer(wl, "cuda:0", shadow_size=256) for wl in wls]
    inputs = [{k: v.cpu().numpy() for k, v in r.inputs.items()} for r in plain]
    br = B.BatchRenderer(wls, "cuda:0", inputs=inputs, use_graph=use_graph)
    for r in plain:
        r.beforeFirstFrame()
    br.beforeFirstFrame()
    br.build()
    for frame in range(3):
        for r in plain:
            r.frame()
        br.frame()
        torch.cuda.synchronize()
        for k, (rp, rb) in enumerate(zip(plain, br.renderers)):
            a, b = _state(rp), _state(rb)
            for key in a:
                same = a[key] == b[key] if key in ("lists", "drawn") else np.array_equal(a[key], b[key])
                assert same, f"frame {frame} instance {k}: {key} differs"
            assert a["img"][..., 3].max() > 0.0


SIM_KEYS = ("now", "prev", "dv")


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_pipelined_matches_plain(gpu, use_graph):
    """Pipelined batched frames (the merged simulation step of frame f beside
    the merged render passes of frame f-1) leave, for every instance, the
    simulation state of its plain frame f and the render state of its plain
    frame f-1; the epilogue leaves the render state of the last frame."""
    import torch
    from frostbitten_hair_webgpu_amd import batch as B
    from frostbitten_hair_webgpu_amd import passes as PS
    wls = _workloads()
    print(f"batch pipelined: graph={use_graph}, {len(wls)} instances", flush=True)
    plain = [PS.HairRenderer(wl, "cuda:0", shadow_size=256) for wl in wls]
    inputs = [{k: v.cpu().numpy() for k, v in r.inputs.items()} for r in plain]
    br = B.BatchRenderer(wls, "cuda:0", inputs=inputs, use_graph=use_graph)
    for r in plain + br.renderers:
        r.beforeFirstFrame()
    for r in br.renderers:  # every call once (scratch sizes), as a plain frame
        r.frame()
    frames = 4
    ref = [[] for _ in plain]
    for _ in range(frames + 1):
        for k, r in enumerate(plain):
            r.frame()
            torch.cuda.synchronize()
            ref[k].append(_state(r))
    br.frameIdx = br.renderers[0].frameIdx
    br.build_pipelined()
    br.pipeline_begin()  # simulation step of plain frame 1 (plain frame 0 ran eagerly above)

    def check(k, st, want, what):
        for key in want:
            same = want[key] == st[key] if key in ("lists", "drawn") else np.array_equal(want[key], st[key])
            assert same, f"{what}, instance {k}: {key} differs"

    for f in range(2, frames + 1):
        br.frame_pipelined()
        torch.cuda.synchronize()
        for k, rb in enumerate(br.renderers):
            st = _state(rb)
            check(k, st, {x: ref[k][f][x] for x in SIM_KEYS}, f"simulation state of frame {f}")
            check(k, st, {x: v for x, v in ref[k][f - 1].items() if x not in SIM_KEYS},
                  f"render state of frame {f - 1}")
    br.pipeline_end()
    torch.cuda.synchronize()
    for k, rb in enumerate(br.renderers):
        st = _state(rb)
        check(k, st, {x: v for x, v in ref[k][frames].items() if x not in SIM_KEYS}, f"render state of frame {frames}")


def test_batch_merge_is_shorter(gpu):
    """The merged frame has one launch per kernel position, not one per
    instance: far fewer launches than the instances' frames together."""
    from frostbitten_hair_webgpu_amd import batch as B
    from frostbitten_hair_webgpu_amd import passes as PS
    wls = _workloads()
    rs = [PS.HairRenderer(wl, "cuda:0", shadow_size=256) for wl in wls]
    for r in rs:
        r.beforeFirstFrame()
    progs = [B.record_frame(r) for r in rs]
    per = [p.launches() for p in progs]
    merged = rs[0].native.program_merge(progs)
    assert max(per) <= merged.launches() < sum(per)
    assert merged.launches() <= max(per) + 4  # tile-size variants of the fine pass and optional scans


def test_batch_generator_instances(gpu):
    """Instances of the batching workload's own generator (make_instance:
    seeded strands, points, resolution, camera, wind and SDF per instance;
    the five smallest of its 64), with render configurations from the
    autotuner's candidates (tile 8 / 16 / 32): the batched frames equal each
    instance's plain frames, replayed as graphs."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import autotune as AT
    from frostbitten_hair_webgpu_amd import batch as B
    from frostbitten_hair_webgpu_amd import passes as PS
    wls = [A.make_instance(i) for i in (50, 63, 33, 37, 30)]
    for wl, cand in zip(wls, AT.CANDIDATES[::2]):
        AT.apply(wl, cand)
    print("instances:", [(wl.strands, wl.points, wl.width, wl.height, wl.render.tileSize) for wl in wls], flush=True)
    plain = [PS.HairRenderer(wl, "cuda:0", shadow_size=256) for wl in wls]
    inputs = [{k: v.cpu().numpy() for k, v in r.inputs.items()} for r in plain]
    br = B.BatchRenderer(wls, "cuda:0", inputs=inputs, use_graph=True)
    for r in plain:
        r.beforeFirstFrame()
    br.beforeFirstFrame()
    br.build()
    for frame in range(2):
        for r in plain:
            r.frame()
        br.frame()
        torch.cuda.synchronize()
        for k, (rp, rb) in enumerate(zip(plain, br.renderers)):
            a, b = _state(rp), _state(rb)
            for key in a:
                same = a[key] == b[key] if key in ("lists", "drawn") else np.array_equal(a[key], b[key])
                assert same, f"frame {frame} instance {k}: {key} differs"
            assert a["img"][..., 3].max() > 0.0


@pytest.mark.parametrize("seed", range(3))
def test_batch_random_instances(gpu, seed):
    """Six instances drawn from the seed (strand and point counts,
    resolution, tile size, bins, slices, fiber width), merged and launched
    as one batched frame from a captured graph: after each of three frames
    every instance's state (positions, grid, tile lists, image, shading,
    screen) equals its plain frame's."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import batch as B
    from frostbitten_hair_webgpu_amd import passes as PS
    rng = np.random.default_rng(0xBA7C0 + seed)
    wls = []
    for i in range(6):
        wl = A.make_workload(str(rng.choice(["small", "c2"])), seed=100 * seed + i,
                             strands=int(rng.integers(32, 1500)), width=int(rng.integers(64, 520)),
                             height=int(rng.integers(48, 400)), points=int(rng.choice([5, 8, 16, 24, 32])))
        rc = wl.render
        rc.tileSize = int(rng.choice([8, 16, 32]))
        rc.tileDepthBins = int(rng.choice([1, 2, 4, 8, 16]))
        rc.slicesPerPixel = int(rng.choice([1, 2, 4, 8]))
        rc.fiberRadius *= float(rng.uniform(0.5, 3.0))
        rc.avgSegmentsPerTile = 1024
        wls.append(wl)
    print(f"batch sweep {seed}: " + "; ".join(f"{wl.strands}x{wl.points} {wl.width}x{wl.height} "
                                               f"t{wl.render.tileSize} b{wl.render.tileDepthBins} "
                                               f"s{wl.render.slicesPerPixel}" for wl in wls), flush=True)
    plain = [PS.HairRenderer(wl, "cuda:0", shadow_size=256) for wl in wls]
    inputs = [{k: v.cpu().numpy() for k, v in r.inputs.items()} for r in plain]
    br = B.BatchRenderer(wls, "cuda:0", inputs=inputs, use_graph=True)
    for r in plain:
        r.beforeFirstFrame()
    br.beforeFirstFrame()
    br.build()
    for frame in range(3):
        for r in plain:
            r.frame()
        br.frame()
        torch.cuda.synchronize()
        for k, (rp, rb) in enumerate(zip(plain, br.renderers)):
            a, b = _state(rp), _state(rb)
            for key in a:
                same = a[key] == b[key] if key in ("lists", "drawn") else np.array_equal(a[key], b[key])
                assert same, f"frame {frame} instance {k}: {key} differs"

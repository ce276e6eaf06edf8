"""Branches (fh_branch_*): frames whose grid accumulation, hardware-raster and
shading passes run on the context's second stream beside the tile, sort and
fine passes leave exactly what frames on one stream leave: positions, grids,
tile lists, images, shading, depth / normals / AO and the presented screen,
eagerly and replayed as HIP graphs, over both parities of the position
buffers, with the precomputed scene inputs, with the hardware-raster passes
and with the background and mesh passes drawn every frame. Pipelined frames
(the simulation step of frame f beside the render passes of frame f-1) leave, frame for frame, the simulation state of frame f
and the render state of frame f-1 that plain frames leave, and deep-pipelined
frames (simulation step f, tile stage f-1, fine stage f-2 on three streams)
the simulation state of frame f, the tile lists of frame f-1 and the image,
screen and shading points of frame f-2."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


SIM_KEYS = ("now", "prev", "dv", "dgw")


def _state(r):
    from gpu_util import walk_lists
    tiles = _np(r.tilesPass.hairTilesBuffer).reshape(-1, 4)
    seg = _np(r.tilesPass.hairTileSegmentsBuffer).view(np.uint32)
    lists = walk_lists(tiles.view(np.uint32), seg, r.tilesPass.segmentsCapacity)
    out = dict(now=_np(r.hair.positionsNow), prev=_np(r.hair.positionsPrev), dv=_np(r.grid.densityVelocity),
               dgw=_np(r.grid.densityGradAndWind), depth=tiles[:, [0, 1, 3]], lists=sorted(lists.items()),
               counts=_np(r.tilesPass.segmentCountPerTileBuffer),
               img=_np(r.finePass.image(r.wl.width, r.wl.height)), shading=_np(r.hair.shading),
               hdr=_np(r.hdrRenderTexture), screen=_np(r.screenTexture))
    if r.hw_passes:
        out.update(hair_depth=_np(r.depthTexture), normals=_np(r.normalsTexture), ao=_np(r.aoTexture),
                   shadow=_np(r.shadowTexture))
    if r.mesh_passes:
        out.update(mesh_depth=_np(r.meshDepthTexture), mesh_normals=_np(r.meshNormalsTexture))
    return out


PASSES = {"none": dict(), "hw": dict(hw_passes=True), "mesh": dict(mesh_passes=True)}


def _same(a, b, where):
    assert a.keys() == b.keys()
    for k in a:
        if isinstance(a[k], np.ndarray):
            x, y = np.ascontiguousarray(a[k]), np.ascontiguousarray(b[k])
            assert x.dtype == y.dtype and np.array_equal(x.view(np.uint8), y.view(np.uint8)), f"{where}: {k} differs"
        else:
            assert a[k] == b[k], f"{where}: {k} differs"


@pytest.mark.parametrize("passes", list(PASSES))
@pytest.mark.parametrize("use_graph", [False, True])
def test_branched_frames_match_one_stream(gpu, passes, use_graph):
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload("c2", seed=77, strands=1200, width=480, height=320, points=16)
    print(f"branches: passes={passes} graph={use_graph} {wl.strands}x{wl.points} {wl.width}x{wl.height}",
          flush=True)
    serial = PS.HairRenderer(wl, "cuda:0", shadow_size=512, branches=False, **PASSES[passes])
    inputs = {k: v.cpu().numpy() for k, v in serial.inputs.items()}
    branched = PS.HairRenderer(wl, "cuda:0", shadow_size=512, inputs=inputs, branches=True, **PASSES[passes])
    for r in (serial, branched):
        r.beforeFirstFrame()
        r.frame()  # every call once before capture
    torch.cuda.synchronize()
    _same(_state(serial), _state(branched), "first frame")
    if use_graph:
        for r in (serial, branched):
            r.capture_frames()
    for f in range(3):
        for r in (serial, branched):
            if use_graph:
                r.frame_graph()
            else:
                r.frame()
        torch.cuda.synchronize()
        _same(_state(serial), _state(branched), f"frame {f + 2}")
    assert branched.native.scratch_bytes() > 0


def _split(st):
    sim = {k: v for k, v in st.items() if k in SIM_KEYS}
    render = {k: v for k, v in st.items() if k not in SIM_KEYS}
    return sim, render


@pytest.mark.parametrize("passes", list(PASSES))
@pytest.mark.parametrize("use_graph", [False, True])
def test_pipelined_frames_match_plain(gpu, passes, use_graph):
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload("c2", seed=78, strands=1000, width=400, height=288, points=24)
    print(f"pipelined: passes={passes} graph={use_graph} {wl.strands}x{wl.points} {wl.width}x{wl.height}",
          flush=True)
    plain = PS.HairRenderer(wl, "cuda:0", shadow_size=512, branches=False, **PASSES[passes])
    inputs = {k: v.cpu().numpy() for k, v in plain.inputs.items()}
    pipe = PS.HairRenderer(wl, "cuda:0", shadow_size=512, inputs=inputs, branches=True, **PASSES[passes])
    frames = 6
    plain.beforeFirstFrame()
    ref = []
    for _ in range(frames):
        plain.frame()
        torch.cuda.synchronize()
        ref.append(_split(_state(plain)))
    pipe.beforeFirstFrame()
    pipe.pipeline_begin()  # simulation step of frame 0
    for f in range(1, frames):
        if use_graph and f >= 2:
            if f == 2:
                pipe.capture_pipelined()
            pipe.frame_pipelined_graph()
        else:
            pipe.frame_pipelined()  # frame 1 runs every call once before the capture
        torch.cuda.synchronize()
        sim, render = _split(_state(pipe))
        _same(ref[f][0], sim, f"simulation state of frame {f}")
        _same(ref[f - 1][1], render, f"render state of frame {f - 1}")
    pipe.pipeline_end()
    torch.cuda.synchronize()
    _same(ref[frames - 1][1], _split(_state(pipe))[1], f"render state of frame {frames - 1}")


def _tile_state(tp, sp):
    from gpu_util import walk_lists
    tiles = _np(tp.hairTilesBuffer).reshape(-1, 4)
    seg = _np(tp.hairTileSegmentsBuffer).view(np.uint32)
    lists = walk_lists(tiles.view(np.uint32), seg, tp.segmentsCapacity)
    return dict(depth=tiles[:, [0, 1, 3]], lists=sorted(lists.items()), counts=_np(tp.segmentCountPerTileBuffer),
                tile_list=sorted(_np(sp.tileListBuffer).view(np.uint32)[1:1 + int(_np(sp.tileListBuffer)[0])]))


@pytest.mark.parametrize("use_graph", [False, True])
def test_deep_pipelined_frames_match_plain(gpu, use_graph):
    """Deep-pipelined frames (simulation step f, tile stage f-1 and fine
    stage f-2 at once, on three streams): after frame f the simulation state
    is plain frame f's, the tile set of frame f-1 holds plain frame f-1's
    tile lists, and the image, screen and shading points are plain frame
    f-2's, frame for frame, over both parities of the position pair and of
    the tile sets."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload("c2", seed=79, strands=1100, width=416, height=272, points=24)
    print(f"deep pipeline: graph={use_graph} {wl.strands}x{wl.points} {wl.width}x{wl.height}", flush=True)
    plain = PS.HairRenderer(wl, "cuda:0", shadow_size=512, branches=False)
    inputs = {k: v.cpu().numpy() for k, v in plain.inputs.items()}
    deep = PS.HairRenderer(wl, "cuda:0", shadow_size=512, inputs=inputs, branches=True)
    frames = 7
    plain.beforeFirstFrame()
    ref = []
    for _ in range(frames):
        plain.frame()
        torch.cuda.synchronize()
        st = _state(plain)
        ref.append(dict(sim={k: st[k] for k in SIM_KEYS}, tiles=_tile_state(plain.tilesPass, plain.sortPass),
                        fine=dict(img=st["img"], shading=st["shading"], screen=st["screen"])))

    def fine_state(r):
        return dict(img=_np(r.finePass.image(wl.width, wl.height)), shading=_np(r.hair.shading),
                    screen=_np(r.screenTexture))

    deep.beforeFirstFrame()
    deep.deep_pipeline_begin()  # frame 0's simulation step, frame 1's beside frame 0's tile stage
    torch.cuda.synchronize()
    _same(ref[1]["sim"], {k: _state(deep)[k] for k in SIM_KEYS}, "simulation state of frame 1")
    _same(ref[0]["tiles"], _tile_state(*deep.deep_tile_set(0)), "tile lists of frame 0")
    for f in range(2, frames):
        if use_graph and f >= 3:
            if f == 3:
                deep.capture_deep()
            deep.frame_deep_graph()
        else:
            deep.frame_deep()  # frame 2 runs every call once before the capture
        torch.cuda.synchronize()
        _same(ref[f]["sim"], {k: _state(deep)[k] for k in SIM_KEYS}, f"simulation state of frame {f}")
        _same(ref[f - 1]["tiles"], _tile_state(*deep.deep_tile_set(f - 1)), f"tile lists of frame {f - 1}")
        _same(ref[f - 2]["fine"], fine_state(deep), f"image, screen and shading of frame {f - 2}")
    deep.deep_pipeline_end()
    torch.cuda.synchronize()
    _same(ref[frames - 1]["tiles"], _tile_state(*deep.deep_tile_set(frames - 1)), f"tile lists of frame {frames - 1}")
    _same(ref[frames - 1]["fine"], fine_state(deep), f"image, screen and shading of frame {frames - 1}")


def test_branch_errors(gpu):
    """Misuse of the branch calls fails with an error and leaves the context
    usable: end without begin, nested begin, join or set_stream or profiling
    while a branch is open or not joined, a capture ended with a branch open."""
    import torch
    from frostbitten_hair_webgpu_amd import native as N
    ctx = N.Context(0)
    ctx.set_stream(torch.cuda.current_stream())
    with pytest.raises(N.FhError, match="no branch is open"):
        ctx.branch_end()
    ctx.branch_begin()
    with pytest.raises(N.FhError, match="already open"):
        ctx.branch_begin()
    with pytest.raises(N.FhError, match="still open"):
        ctx.branch_join()
    with pytest.raises(N.FhError, match="branch is open"):
        ctx.set_stream(torch.cuda.current_stream())
    buf = torch.full((1024,), 9, dtype=torch.int32, device="cuda")
    ctx.call("fh_clear", N.ptr(buf), 1024 * 4)  # runs on the branch
    ctx.branch_end()
    with pytest.raises(N.FhError, match="not joined"):
        ctx.profiler_enable(8)
    ctx.branch_join()
    ctx.sync()
    assert (buf.cpu().numpy() == 0).all()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    ctx.set_stream(side)
    ctx.capture_begin()
    ctx.branch_begin()
    ctx.call("fh_clear", N.ptr(buf), 1024 * 4)
    with pytest.raises(N.FhError, match="branch was still open"):
        ctx.capture_end()
    ctx.set_stream(torch.cuda.current_stream())
    buf.fill_(5)
    ctx.call("fh_clear", N.ptr(buf), 512 * 4)  # the context still works
    ctx.sync()
    b = buf.cpu().numpy()
    assert (b[:512] == 0).all() and (b[512:] == 5).all()


@pytest.mark.parametrize("pipelined", [False, True])
def test_uniform_ring_matches_uploads(gpu, pipelined):
    """Frames whose simulation uniforms come from a device ring
    (fh_copy_indexed at the start of each simulation step, inside the
    replayed graphs) leave what frames with a host upload per frame leave."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload("c2", seed=79, strands=900, width=400, height=288, points=16)
    print(f"uniform ring: pipelined={pipelined}", flush=True)
    plain = PS.HairRenderer(wl, "cuda:0", shadow_size=512, branches=False)
    inputs = {k: v.cpu().numpy() for k, v in plain.inputs.items()}
    ringed = PS.HairRenderer(wl, "cuda:0", shadow_size=512, inputs=inputs, branches=True)
    frames = 6
    rows = np.stack([ringed.sim_uniforms(f).view(np.uint8).reshape(-1) for f in range(frames + 2)])
    ringed.use_uniform_ring(torch.from_numpy(rows).to("cuda:0"), 0)
    plain.beforeFirstFrame()
    ref = []
    for _ in range(frames):
        plain.frame()
        torch.cuda.synchronize()
        ref.append(_state(plain))
    ringed.beforeFirstFrame()
    if pipelined:
        ringed.pipeline_begin(upload_uniforms=False)
        ringed.frame_pipelined(upload_uniforms=False)
        ringed.capture_pipelined()
        for f in range(2, frames):
            ringed.frame_pipelined_graph(upload_uniforms=False)
            torch.cuda.synchronize()
            sim, _ = _split(_state(ringed))
            _same(_split(ref[f])[0], sim, f"simulation state of frame {f}")
    else:
        ringed.frame(upload_uniforms=False)
        ringed.capture_frames()
        for f in range(1, frames):
            ringed.frame_graph(upload_uniforms=False)
            torch.cuda.synchronize()
            _same(ref[f], _state(ringed), f"frame {f}")
    assert int(ringed.uniformCounter.item()) == frames

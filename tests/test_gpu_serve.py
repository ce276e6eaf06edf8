"""Served form: each request's image is the same alone (one request at a
time) and in the mix (continuous batching), also when a tight memory budget
evicts instances between requests and restores their state."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pool():
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    specs = [dict(kind="small", strands=300, width=320, height=240, points=None),
             dict(kind="c2", strands=800, width=400, height=300, points=8),
             dict(kind="c2", strands=1200, width=512, height=384, points=24),
             dict(kind="small", strands=100, width=160, height=120, points=None)]
    pool = {}
    for i, sp in enumerate(specs):
        wl = A.make_workload(sp["kind"], seed=2000 + i, strands=sp["strands"], width=sp["width"],
                             height=sp["height"], points=sp["points"])
        wl.render.tileSize = (16, 8, 32, 16)[i]
        pool[10 + i] = (wl, PS.make_inputs(wl, PS.workload_render_uniforms(wl), 256))
    return pool


def _requests():
    from frostbitten_hair_webgpu_amd import serve as SV
    order = [(10, 2), (11, 3), (10, 1), (12, 2), (13, 4), (11, 1), (12, 3), (10, 2)]
    return [SV.Request(i, a, s, 0.0) for i, (a, s) in enumerate(order)]


def _serve(pool, max_active, budget, pipelined=False, requests=None):
    from frostbitten_hair_webgpu_amd import serve as SV
    print(f"serve: max_active={max_active} budget={budget} pipelined={pipelined}", flush=True)
    srv = SV.HairServer(pool, budget, max_active=max_active, device="cuda:0", keep_images=True,
                        pipelined=pipelined)
    reqs = srv.run(requests() if requests else _requests(), realtime=False)
    stats = dict(srv.stats)
    srv.close()
    return reqs, stats


@pytest.mark.parametrize("pipelined", [False, True])
def test_serve_alone_equals_mix(gpu, pipelined):
    """Each request's image is the same served alone (one request at a time,
    plain ticks) and in the mix (continuous batching, plain or pipelined
    ticks)."""
    pool = _pool()
    alone, st_alone = _serve(pool, 1, 1 << 40)
    mix, st_mix = _serve(pool, 8, 1 << 40, pipelined)
    assert st_mix["ticks"] < st_alone["ticks"]  # requests overlapped
    assert st_alone["frames"] == st_mix["frames"] == sum(q.steps for q in _requests())
    for a, b in zip(alone, mix):
        assert a.image is not None and b.image is not None
        assert np.array_equal(a.image.cpu().numpy(), b.image.cpu().numpy()), f"request {a.rid} differs"
        assert np.array_equal(a.screen.cpu().numpy(), b.screen.cpu().numpy()), f"request {a.rid} screen differs"
        assert a.image[..., 3].max() > 0


@pytest.mark.parametrize("pipelined", [False, True])
def test_serve_eviction_restores_state(gpu, pipelined):
    from frostbitten_hair_webgpu_amd import serve as SV
    pool = _pool()
    alone, _ = _serve(pool, 1, 1 << 40)
    # room for about two instances at a time: later requests evict idle ones
    sizes = sorted(SV.estimate_bytes(wl) for wl, _ in pool.values())
    tight, st = _serve(pool, 8, sizes[-1] + sizes[-2], pipelined)
    assert st["evictions"] > 0
    for a, b in zip(alone, tight):
        assert np.array_equal(a.image.cpu().numpy(), b.image.cpu().numpy()), f"request {a.rid} differs"


def test_serve_generator_assets(gpu):
    """Assets of the served workload's own pool generator (make_pool_asset:
    seeded strands, 16-32 points, 720p here; the four smallest of its 256)
    and a Poisson request stream of its own generator (poisson_requests,
    frame counts U{1..12}): each request's image is the same alone and in
    the mix, with pipelined ticks."""
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    from frostbitten_hair_webgpu_amd import serve as SV
    pool = {}
    for a in (51, 111, 74, 116):
        wl = A.make_pool_asset(a)
        pool[a] = (wl, PS.make_inputs(wl, PS.workload_render_uniforms(wl), 256))

    def requests():
        return SV.poisson_requests(10, 50.0, sorted(pool), 12, seed=A.SERVE_SEED)

    alone, st_alone = _serve(pool, 1, 1 << 40, requests=requests)
    mix, st_mix = _serve(pool, 8, 1 << 40, pipelined=True, requests=requests)
    assert st_mix["ticks"] < st_alone["ticks"]
    for a, b in zip(alone, mix):
        assert np.array_equal(a.image.cpu().numpy(), b.image.cpu().numpy()), f"request {a.rid} differs"
        assert a.image[..., 3].max() > 0

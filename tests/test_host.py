"""Host-side logic on the CPU: layouts, packers, formats, generators, and the
C-ABI library (loads and exports every symbol of include/fh_hair.h)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def _header_symbols():
    text = (ROOT / "include" / "fh_hair.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|uint32_t|void\s*\*|const char\s*\*)\s*(fh_\w+)\s*\(", text, re.M)))


def test_library_exports_every_header_symbol():
    from frostbitten_hair_webgpu_amd import native as N
    lib = N.load()
    syms = _header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert lib.fh_abi_version() == 1
    assert set(N.SIGNATURES) >= set(syms)


def test_header_struct_sizes_match_wgsl_layouts(tmp_path):
    """The C structs mirror the WGSL uniform blocks byte for byte."""
    src = tmp_path / "s.c"
    src.write_text(
        '#include "%s"\n#include <stdio.h>\n#include <stddef.h>\n'
        'int main(){printf("%%zu %%zu %%zu %%zu %%zu %%zu %%zu %%zu\\n", sizeof(fh_sim_uniforms), '
        'sizeof(fh_render_uniforms), offsetof(fh_render_uniforms, shadows), offsetof(fh_render_uniforms, ao), '
        'offsetof(fh_render_uniforms, hairMaterial), offsetof(fh_render_uniforms, maxDrawnHairSegments), '
        'offsetof(fh_sim_uniforms, deltaTime), sizeof(fh_tile_result)); return 0;}\n' % (ROOT / "include/fh_hair.h"))
    exe = tmp_path / "s"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert list(map(int, out)) == [304, 1184, 736, 960, 992, 1128, 224, 16]


def test_numpy_layouts_match_header():
    from frostbitten_hair_webgpu_amd import layouts as L
    assert L.SIM_UNIFORMS_DTYPE.itemsize == 304
    assert L.RENDER_UNIFORMS_DTYPE.itemsize == 1184
    f = L.RENDER_UNIFORMS_DTYPE.fields
    assert f["viewport"][1] == 576 and f["cameraPosition"][1] == 592 and f["lights"][1] == 640
    assert f["shadowSourceMVP"][1] == 864 and f["fiberRadius"][1] == 1120 and f["maxDrawnHairSegments"][1] == 1128
    s = L.SIM_UNIFORMS_DTYPE.fields
    assert s["frameIdx"][1] == 236 and s["windStrengthJitter"][1] == 292


def test_render_cfg_struct_matches_header():
    from frostbitten_hair_webgpu_amd import native as N
    assert C.sizeof(N.RenderCfg) == 48
    assert C.sizeof(N.Tex2D) == 16


def test_segment_buffer_sizes():
    from frostbitten_hair_webgpu_amd import layouts as L
    from frostbitten_hair_webgpu_amd import native as N
    rc = L.RenderConfig(tileSize=16)
    assert L.segments_buffer_entries(rc, 1920, 1080) == 1 + 120 * 68 * 128
    lib = N.load()
    cfg = N.RenderCfg.from_config(rc)
    assert lib.fh_tile_count(C.byref(cfg), 1920, 1080) == 120 * 68
    assert lib.fh_segments_buffer_entries(C.byref(cfg), 1920, 1080, 128.0) == 1 + 120 * 68 * 128
    assert lib.fh_slice_data_per_processor(C.byref(cfg)) == 16 * 8 * 16 * 16


def test_sim_uniform_packing():
    from frostbitten_hair_webgpu_amd import layouts as L
    sc = L.SimConfig(windDirPhi=0.0, windDirTheta=90.0, windStrength=2.0, constraintIterations=3)
    u = L.pack_sim_uniforms(sc, np.eye(4), (np.zeros(3), np.ones(3)), (np.full(3, -1.0), np.full(3, 1.0)), 7)
    assert u["frameIdx"][0] == 7 and u["constraintIterations"][0] == 3
    assert np.allclose(u["wind"][0], [1.0, 0.0, 0.0, 2.0], atol=1e-7)
    assert u["sdfBoundsMax"][0][3] == np.float32(sc.sdfDistanceOffset)
    assert np.array_equal(u["modelMatrix"][0], np.eye(4, dtype=np.float32).reshape(16))


def test_render_uniform_packing_projects_like_reference():
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import layouts as L
    wl = A.make_workload("small", strands=64)
    box = A.bounding_box(wl.positions)
    u = L.pack_render_uniforms(wl.render, wl.sim, wl.camera, 640, 480, np.eye(4), A.bounding_sphere(box),
                               [], wl.sdf_bounds, A.scale_box(box, 2.0))
    vp = u["vpMatrix"][0].reshape(4, 4).T.astype(np.float64)
    c = vp @ np.array([*wl.camera.target, 1.0])
    ndc = c[:3] / c[3]
    assert abs(ndc[0]) < 1e-5 and abs(ndc[1]) < 1e-5 and 0.0 < ndc[2] < 1.0  # target at the centre
    inv = u["vpMatrixInv"][0].reshape(4, 4).T.astype(np.float64)
    assert np.allclose(inv @ vp, np.eye(4), atol=1e-4)
    assert u["maxDrawnHairSegments"][0] == L.segments_buffer_entries(wl.render, 640, 480)


def test_tfx_roundtrip_and_reference_asset_header():
    from frostbitten_hair_webgpu_amd import formats as F
    pos = np.random.default_rng(1).normal(size=(6 * 4, 4)).astype(np.float32)
    t = F.parse_tfx(F.write_tfx(pos, 6, 4))
    assert t.num_strands == 6 and t.points_per_strand == 4 and np.array_equal(t.positions, pos)
    ref = Path("/root/reference/static/models/SintelHairOriginal-sintel_hair.16points.tfx")
    if ref.exists():  # format smoke test on the reference's shipped asset
        t = F.parse_tfx(ref.read_bytes())
        assert t.num_strands == 11400 and t.points_per_strand == 16
        assert np.all(t.positions.reshape(-1, 16, 4)[:, 0, 3] == 0.0)


def test_generator_is_seeded_and_well_formed():
    from frostbitten_hair_webgpu_amd import assets as A
    a = A.make_workload("c2", strands=300)
    b = A.make_workload("c2", strands=300)
    assert np.array_equal(a.positions, b.positions)
    p = a.positions.reshape(300, 32, 4)
    assert np.all(p[:, 0, 3] == 0) and np.all(p[:, 1:, 3] == 1)
    assert np.allclose(np.linalg.norm(p[:, 0, :3], axis=1), A.HEAD_RADIUS, atol=1e-6)
    polar = np.degrees(np.arccos(p[:, 0, 1] / A.HEAD_RADIUS))
    assert polar.max() <= 70.0 + 1e-3
    L = np.linalg.norm(np.diff(p[:, :, :3], axis=1), axis=2).sum(axis=1)
    assert 0.15 < np.median(L) < 0.45
    t = A.tangents(a.positions, 300, 32).reshape(300, 32, 4)
    assert np.allclose(np.linalg.norm(t[..., :3], axis=2), 1.0, atol=1e-5)
    assert np.array_equal(t[:, -1], t[:, -2])
    sl = A.segment_lengths(a.positions, 300, 32).reshape(300, 32)
    assert np.array_equal(sl[:, -1], sl[:, -2])


def test_sdf_bake_matches_analytic_distance(oracle):
    from frostbitten_hair_webgpu_amd import assets as A
    col = A.Collider(spheres=[((0.0, 0.0, 0.0), 0.1)])
    lo, hi, d = A.bake_sdf(col, 32, margin=0.05)
    for p in ([0.0, 0.0, 0.0], [0.12, 0.0, 0.0], [0.05, -0.07, 0.02]):
        v = oracle.sample_sdf(d, 32, lo, hi, np.array(p, np.float32))
        assert v == pytest.approx(np.linalg.norm(p) - 0.1, abs=0.01)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from frostbitten_hair_webgpu_amd import native as N
    monkeypatch.setattr(N, "_lib", None)
    with pytest.raises(N.FhError):
        N.load(tmp_path / "nope.so")

"""The library's uniform packers (fh_pack_render_uniforms,
fh_pack_sim_uniforms) against the Python packers (layouts.py), and their
matrices against what the projection geometry fixes.

Both packers build the matrices in double from the same inputs and store
f32; they differ in the order of a few double operations (the inverse:
LAPACK's LU here, Gauss-Jordan in the library), so float words may differ by
a unit or two in the last place. Integer words must be equal.
"""
import math

import numpy as np
import pytest

from frostbitten_hair_webgpu_amd import layouts as L

INT_FIELDS_RENDER = ("shadowUsePCSS", "shadowPCFRadius", "aoNumDirections", "aoNumSteps", "gridDebugDisplayValue",
                     "maxDrawnHairSegments", "displayMode", "gizmoActiveState")
INT_FIELDS_SIM = ("gridDebugDisplayValue", "constraintIterations", "frameIdx")


def _inputs(kind, rotate=False):
    from frostbitten_hair_webgpu_amd import assets as A
    wl = A.make_workload(kind, strands=64, width=333, height=171)
    model = np.eye(4)
    if rotate:
        c, s = math.cos(0.4), math.sin(0.4)
        model = np.array([[c, 0, s, 0.05], [0, 1, 0, -0.02], [-s, 0, c, 0.1], [0, 0, 0, 1.0]])
    box = A.derived(wl)["box"]
    return wl, model, A.bounding_sphere(box), A.scale_box(box, wl.sim.gridScale)


def _compare(a_bytes, b_arr, dtype, int_fields):
    a = np.frombuffer(a_bytes, dtype=np.uint8).view(dtype)[0]
    b = b_arr[0]
    for name in dtype.names:
        x, y = np.asarray(a[name]), np.asarray(b[name])
        if name in int_fields:
            assert np.array_equal(x, y), name
            continue
        xf, yf = x.astype(np.float32).reshape(-1), y.astype(np.float32).reshape(-1)
        ulp = np.abs(xf.view(np.int32).astype(np.int64) - yf.view(np.int32).astype(np.int64))
        close = (xf == yf) | (ulp <= 2) | (np.abs(xf - yf) <= 1e-7)
        assert close.all(), (name, xf[~close], yf[~close])


@pytest.mark.parametrize("kind,rotate", [("small", False), ("c2", False), ("c2", True)])
def test_render_uniforms_match_python_packer(kind, rotate):
    from frostbitten_hair_webgpu_amd import native as N
    wl, model, hs, grid = _inputs(kind, rotate)
    args = (wl.render, wl.sim, wl.camera, wl.width, wl.height, model, hs, [(np.zeros(3), 0.3)], wl.sdf_bounds, grid)
    ref = L.pack_render_uniforms(*args)
    got = N.pack_render_uniforms(L.render_params(*args))
    assert len(got) == L.RENDER_UNIFORMS_DTYPE.itemsize == 1184
    _compare(got, ref, L.RENDER_UNIFORMS_DTYPE, INT_FIELDS_RENDER)


@pytest.mark.parametrize("rotate", [False, True])
def test_sim_uniforms_match_python_packer(rotate):
    from frostbitten_hair_webgpu_amd import native as N
    wl, model, hs, grid = _inputs("c2", rotate)
    ref = L.pack_sim_uniforms(wl.sim, model, wl.sdf_bounds, grid, 37)
    got = N.pack_sim_uniforms(L.sim_params(wl.sim, model, wl.sdf_bounds, grid, 37))
    assert len(got) == L.SIM_UNIFORMS_DTYPE.itemsize == 304
    _compare(got, ref, L.SIM_UNIFORMS_DTYPE, INT_FIELDS_SIM)


def _mat(u, name):
    return np.asarray(u[name], dtype=np.float64).reshape(4, 4).T  # column-major -> [row, col]


def test_packed_matrices_geometry():
    """What projection geometry fixes: vp * vp^-1 = I; the camera target maps
    to the NDC centre; the near / far planes map to z = 0 / 1 (WebGPU); the
    view matrix takes the eye to the origin; the shadow source's box holds the
    hair's bounding sphere."""
    from frostbitten_hair_webgpu_amd import native as N
    wl, model, hs, grid = _inputs("c2", True)
    args = (wl.render, wl.sim, wl.camera, wl.width, wl.height, model, hs, [(np.zeros(3), 0.3)], wl.sdf_bounds, grid)
    u = np.frombuffer(N.pack_render_uniforms(L.render_params(*args)), np.uint8).view(L.RENDER_UNIFORMS_DTYPE)[0]
    vp, vpi = _mat(u, "vpMatrix"), _mat(u, "vpMatrixInv")
    assert np.allclose(vp @ vpi, np.eye(4), atol=2e-5)
    eye, target = np.asarray(wl.camera.eye), np.asarray(wl.camera.target)
    t = vp @ np.r_[target, 1.0]
    assert np.allclose(t[:2] / t[3], 0.0, atol=1e-5)
    view, proj = _mat(u, "viewMatrix"), _mat(u, "projMatrix")
    assert np.allclose(view @ np.r_[eye, 1.0], [0, 0, 0, 1], atol=1e-5)
    for dist, z in ((wl.render.camera_near, 0.0), (wl.render.camera_far, 1.0)):
        c = proj @ np.array([0.0, 0.0, -dist, 1.0])
        assert abs(c[2] / c[3] - z) < 1e-5
    smvp = _mat(u, "shadowSourceMVP")
    c = smvp @ np.r_[hs[0], 1.0]
    assert (np.abs(c[:2] / c[3]) < 1.0).all() and 0.0 < c[2] / c[3] < 1.0

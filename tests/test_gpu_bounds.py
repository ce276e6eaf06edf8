"""The pixel bounds the tile and fine passes enumerate never lose a pixel.

Both passes visit, for each projected segment quad, only the pixels of its
pixel box (seg_pixel_bbox) and, row by row, of a conservative x span
(span_frame / span_x), both widened by a margin (raster_margin) that covers
the f32 rounding of the edge tests (swRasterizeHair.wgsl.ts:127-129). The
reference tests every pixel of every tile of the segment's tile rectangle
(hairTilesPass.wgsl.ts:171-203) and every pixel of the tile in the fine
pass (processHairSegment.wgsl.ts:14-22), so a pixel that passes the edge
tests but falls outside these bounds would be a pixel the reference draws
and this build does not. fh_debug_pixel_bounds runs the same device
functions over every pixel within 3 px of each quad's box.

Quads: hair-like ones (thin, long, bent caps, at 4K and at coordinates up
to 40,000 px), arbitrary four-point quads (concave and self-intersecting
ones among them), and near-degenerate ones (slivers a thousandth of a pixel
wide, sharp corners, repeated vertices). Bar: no inside pixel outside the
bounds, on any quad.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hair_quads(rng, n, scale):
    c0 = rng.uniform(0.0, scale, (n, 2))
    ang = rng.uniform(0.0, 2 * np.pi, n)
    length = np.exp(rng.uniform(np.log(0.3), np.log(40.0), n))
    a = np.stack([np.cos(ang), np.sin(ang)], 1)
    c1 = c0 + a * length[:, None]
    w0 = np.exp(rng.uniform(np.log(0.02), np.log(6.0), n))
    w1 = w0 * rng.uniform(0.8, 1.2, n)
    n0 = np.stack([-a[:, 1], a[:, 0]], 1)
    tilt = np.clip(rng.normal(0.0, np.radians(25.0), n), -np.radians(80.0), np.radians(80.0))
    ct, st = np.cos(tilt), np.sin(tilt)
    n1 = np.stack([n0[:, 0] * ct - n0[:, 1] * st, n0[:, 0] * st + n0[:, 1] * ct], 1)
    # either winding: the edge tests accept one (swapping the sides flips it)
    side = np.where(rng.uniform(size=(n, 1)) < 0.5, 1.0, -1.0)
    v00, v01 = c0 - side * n0 * w0[:, None] / 2, c0 + side * n0 * w0[:, None] / 2
    v10, v11 = c1 - side * n1 * w1[:, None] / 2, c1 + side * n1 * w1[:, None] / 2
    return np.concatenate([v00, v01, v10, v11], 1)


def _free_quads(rng, n):
    off = rng.uniform(0.0, 4096.0, (n, 1, 2))
    size = np.exp(rng.uniform(np.log(0.1), np.log(30.0), (n, 1, 1)))
    return (off + rng.uniform(0.0, 1.0, (n, 4, 2)) * size).reshape(n, 8)


def _degenerate_quads(rng, n):
    q = _hair_quads(rng, n, 4096.0).reshape(n, 4, 2)
    k = n // 4
    # slivers: widths of 1e-4 .. 1e-2 px
    mid0, mid1 = (q[:k, 0] + q[:k, 1]) / 2, (q[:k, 2] + q[:k, 3]) / 2
    d0, d1 = q[:k, 1] - q[:k, 0], q[:k, 3] - q[:k, 2]
    s = np.exp(rng.uniform(np.log(1e-4), np.log(1e-2), (k, 1)))
    q[:k, 0], q[:k, 1] = mid0 - d0 * s, mid0 + d0 * s
    q[:k, 2], q[:k, 3] = mid1 - d1 * s, mid1 + d1 * s
    # sharp corners: the end cap nearly along the axis
    axis = q[k:2 * k, 2] - q[k:2 * k, 0]
    half = np.linalg.norm(q[k:2 * k, 3] - q[k:2 * k, 2], axis=1, keepdims=True) / 2
    u = axis / np.maximum(np.linalg.norm(axis, axis=1, keepdims=True), 1e-9)
    rot = rng.uniform(-0.05, 0.05, (k, 1))
    u2 = np.concatenate([u[:, :1] - rot * u[:, 1:], u[:, 1:] + rot * u[:, :1]], 1)
    centre = (q[k:2 * k, 2] + q[k:2 * k, 3]) / 2
    q[k:2 * k, 2], q[k:2 * k, 3] = centre - u2 * half, centre + u2 * half
    # a repeated vertex (a triangle)
    q[2 * k:3 * k, 3] = q[2 * k:3 * k, 2]
    # nearly collinear four points
    t = rng.uniform(0.0, 1.0, (n - 3 * k, 4, 1))
    p0 = q[3 * k:, 0][:, None, :]
    d = (q[3 * k:, 2] - q[3 * k:, 0])[:, None, :]
    jitter = rng.normal(0.0, 1e-3, (n - 3 * k, 4, 2))
    q[3 * k:] = p0 + t * d + jitter
    return q.reshape(n, 8)


def test_pixel_bounds_hold(gpu):
    import torch
    from frostbitten_hair_webgpu_amd import native as N
    rng = np.random.default_rng(0xB0DE5)
    sets = {
        "hair 4K": _hair_quads(rng, 300_000, 4096.0),
        "hair far": _hair_quads(rng, 100_000, 40_000.0),
        "four points": _free_quads(rng, 200_000),
        "near-degenerate": _degenerate_quads(rng, 200_000),
    }
    ctx = N.Context(0)
    for name, q in sets.items():
        quads = torch.from_numpy(np.ascontiguousarray(q.astype(np.float32))).to(gpu)
        out = torch.zeros((q.shape[0] * 4,), dtype=torch.int32, device=gpu)
        ctx.call("fh_debug_pixel_bounds", C.c_void_p(quads.data_ptr()), q.shape[0], C.c_void_p(out.data_ptr()))
        torch.cuda.synchronize()
        r = out.cpu().numpy().view(np.uint32).reshape(-1, 4).astype(np.uint64)
        checked = r[:, 3] != 0xFFFFFFFF
        box_miss, span_miss = int(r[checked, 0].sum()), int(r[checked, 1].sum())
        cand, inside = int(r[checked, 2].sum()), int(r[checked, 3].sum())
        print(f"{name}: {checked.sum()} of {q.shape[0]} quads checked, {inside} inside pixels, "
              f"{cand} candidates ({cand / max(inside, 1):.2f} per inside pixel), "
              f"outside the box {box_miss}, outside the span {span_miss}", flush=True)
        assert inside > 0
        assert box_miss == 0 and span_miss == 0

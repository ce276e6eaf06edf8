"""ctypes binding of libfhhip.so (include/fh_hair.h).

Loading fails loudly when the library is missing: there is no CPU fallback
for any pass.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libfhhip.so"

vp = C.c_void_p
u32 = C.c_uint32
i32 = C.c_int


class RenderCfg(C.Structure):
    """struct fh_render_cfg"""
    _fields_ = [
        ("tileSize", u32), ("tileDepthBins", u32), ("slicesPerPixel", u32), ("avgFragmentsPerSlice", u32),
        ("processorCount", u32), ("invalidTilesPerSegmentThreshold", u32), ("sortBuckets", u32),
        ("sortBucketSize", u32), ("shadingPoints", u32), ("alphaQuadratic", u32), ("alphaMultiplier", C.c_float),
        ("reserved", u32),
    ]

    @classmethod
    def from_config(cls, rc) -> "RenderCfg":
        return cls(rc.tileSize, rc.tileDepthBins, rc.slicesPerPixel, rc.avgFragmentsPerSlice, rc.processorCount,
                   rc.invalidTilesPerSegmentThreshold, rc.sortBuckets, rc.sortBucketSize, rc.shadingPoints,
                   1 if rc.alphaQuadratic else 0, rc.alphaMultiplier, 0)


class Tex2D(C.Structure):
    """struct fh_tex2d"""
    _fields_ = [("data", vp), ("width", u32), ("height", u32)]

    def ptr(self) -> vp:
        return vp(self.data)


d = C.c_double


class LightParams(C.Structure):
    """struct fh_light_params"""
    _fields_ = [("posPhi", d), ("posTheta", d), ("color", d * 3), ("energy", d)]


class RenderParams(C.Structure):
    """struct fh_render_params"""
    _fields_ = [
        ("width", u32), ("height", u32), ("cameraEye", d * 3), ("cameraTarget", d * 3), ("cameraUp", d * 3),
        ("fovDeg", d), ("near", d), ("far", d), ("modelMatrix", d * 16), ("hairSphere", d * 4),
        ("meshSpheres", C.POINTER(d)), ("meshSphereCount", u32), ("tileSize", u32), ("avgSegmentsPerTile", d),
        ("sdfBoundsMin", d * 3), ("sdfBoundsMax", d * 3), ("sdfDistanceOffset", d), ("gridBoundsMin", d * 3),
        ("gridBoundsMax", d * 3), ("collisionSphere", d * 4), ("fiberRadius", d), ("displayMode", u32),
        ("dbgTileModeMaxSegments", u32), ("dbgSlicesModeMaxSlices", u32), ("dbgShowTiles", u32), ("colors", d * 3),
        ("lightAmbient", d * 4), ("lights", LightParams * 3), ("shadowUsePCSS", u32), ("shadowPCFRadius", u32),
        ("shadowBias", d), ("shadowStrength", d), ("shadowPosPhi", d), ("shadowPosTheta", d),
        ("shadowDistance", d), ("shadowTarget", d * 3), ("aoRadius", d), ("aoDirectionOffset", d),
        ("aoFalloffStart2", d), ("aoFalloffEnd2", d), ("aoNumDirections", u32), ("aoNumSteps", u32),
        ("aoStrength", d), ("hairColor0", d * 3), ("hairColor1", d * 3), ("specular", d), ("shadows", d),
        ("weightTT", d), ("weightTRT", d), ("shift", d), ("roughness", d), ("colorRng", d), ("lumaRng", d),
        ("attenuation", d),
    ]


class SimParams(C.Structure):
    """struct fh_sim_params"""
    _fields_ = [
        ("modelMatrix", d * 16), ("collisionSphere", d * 4), ("windDirPhi", d), ("windDirTheta", d),
        ("windStrength", d), ("sdfBoundsMin", d * 3), ("sdfBoundsMax", d * 3), ("sdfDistanceOffset", d),
        ("gridBoundsMin", d * 3), ("gridBoundsMax", d * 3), ("frameIdx", u32), ("constraintIterations", u32),
        ("deltaTime", d), ("gravity", d), ("windStrengthLull", d), ("windColisionTraceOffset", d),
        ("stiffnessLengthConstr", d), ("stiffnessCollisions", d), ("stiffnessSDF", d), ("volumePreservation", d),
        ("friction", d), ("stiffnessGlobalConstr", d), ("globalConstrExtent", d), ("globalConstrFade", d),
        ("stiffnessLocalConstr", d), ("windPhaseOffset", d), ("windStrengthFrequency", d),
        ("windStrengthJitter", d),
    ]


class Mesh(C.Structure):
    """struct fh_mesh (device pointers)"""
    _fields_ = [("positions", vp), ("normals", vp), ("indices", vp), ("vertex_count", u32),
                ("triangle_count", u32), ("is_collider_preview", u32), ("reserved", u32)]


class KernelTime(C.Structure):
    """struct fh_kernel_time"""
    _fields_ = [("name", C.c_char * 48), ("launches", u32), ("reserved", u32), ("total_ms", C.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    "fh_create": (i32, [i32, C.POINTER(vp)]),
    "fh_destroy": (i32, [vp]),
    "fh_set_stream": (i32, [vp, vp]),
    "fh_get_stream": (vp, [vp]),
    "fh_sync": (i32, [vp]),
    "fh_last_error": (C.c_char_p, []),
    "fh_abi_version": (i32, []),
    "fh_alloc": (i32, [vp, C.c_uint64, C.POINTER(vp)]),
    "fh_free": (i32, [vp, vp]),
    "fh_upload": (i32, [vp, vp, vp, C.c_uint64]),
    "fh_download": (i32, [vp, vp, vp, C.c_uint64]),
    "fh_clear": (i32, [vp, vp, C.c_uint64]),
    "fh_tile_count": (u32, [C.POINTER(RenderCfg), u32, u32]),
    "fh_segments_buffer_entries": (u32, [C.POINTER(RenderCfg), u32, u32, C.c_float]),
    "fh_slice_data_per_processor": (u32, [C.POINTER(RenderCfg)]),
    "fh_grid_pre_sim": (i32, [vp, u32, vp, vp, vp, vp, u32]),
    "fh_sim_integrate": (i32, [vp, u32, vp, vp, vp, vp, vp, vp, u32, vp, vp, vp, u32, u32]),
    "fh_grid_post_sim": (i32, [vp, u32, vp, vp, vp, vp, vp, u32, u32]),
    "fh_grid_post_sim_noalloc": (i32, [vp, u32, vp, vp, vp, vp, vp, u32, u32]),
    "fh_tiles_clear": (i32, [vp, C.POINTER(RenderCfg), u32, u32, vp, vp, vp]),
    "fh_tiles_draw": (i32, [vp, C.POINTER(RenderCfg), u32, u32, vp, vp, vp, vp, vp, Tex2D, vp, u32, vp, vp,
                            u32, u32]),
    "fh_tiles_rebuild_slot_keys": (i32, [vp, C.POINTER(RenderCfg), u32, u32, vp, vp, u32, vp]),
    "fh_tiles_sort_clear": (i32, [vp, C.POINTER(RenderCfg), vp, vp]),
    "fh_tiles_sort": (i32, [vp, C.POINTER(RenderCfg), u32, u32, vp, vp, vp]),
    "fh_fine_clear": (i32, [vp, u32, u32, vp]),
    "fh_fine_raster": (i32, [vp, C.POINTER(RenderCfg), u32, u32, vp, vp, vp, vp, vp, u32, vp, vp, Tex2D, vp, vp,
                             vp, vp, u32, u32]),
    "fh_shading": (i32, [vp, C.POINTER(RenderCfg), vp, vp, vp, vp, vp, Tex2D, Tex2D, Tex2D, u32, u32]),
    "fh_debug_raster_barycentric": (i32, [vp, vp, vp, vp, u32, u32, vp]),
    "fh_hair_combine": (i32, [vp, C.POINTER(RenderCfg), u32, u32, vp, vp, vp, vp, vp, vp]),
    "fh_present": (i32, [vp, u32, u32, vp, vp, Tex2D, Tex2D, vp, u32, vp, u32]),
    "fh_combine_present": (i32, [vp, C.POINTER(RenderCfg), u32, u32, vp, vp, vp, vp, vp, vp, vp, Tex2D, Tex2D, vp,
                                 u32, vp, u32]),
    "fh_copy": (i32, [vp, vp, vp, C.c_uint64]),
    "fh_copy_indexed": (i32, [vp, vp, vp, C.c_uint64, u32, vp]),
    "fh_hw_hair": (i32, [vp, u32, u32, vp, vp, vp, vp, vp, u32, u32]),
    "fh_shadow_map_hair": (i32, [vp, vp, vp, vp, vp, u32, u32, u32]),
    "fh_calc_ao": (i32, [vp, vp, Tex2D, vp, vp, u32, u32]),
    "fh_meshes_create": (i32, [vp, C.POINTER(Mesh), u32, C.POINTER(vp)]),
    "fh_meshes_destroy": (i32, [vp]),
    "fh_meshes_triangles": (i32, [vp, C.POINTER(u32)]),
    "fh_draw_background": (i32, [vp, u32, u32, vp, vp]),
    "fh_draw_meshes": (i32, [vp, u32, u32, vp, vp, Tex2D, Tex2D, vp, vp, vp]),
    "fh_draw_meshes_visibility": (i32, [vp, u32, u32, vp, vp, vp, vp, vp]),
    "fh_shade_meshes": (i32, [vp, u32, u32, vp, vp, vp, Tex2D, Tex2D, vp]),
    "fh_shadow_map_meshes": (i32, [vp, vp, vp, vp, u32]),
    "fh_pack_render_uniforms": (i32, [C.POINTER(RenderParams), vp]),
    "fh_pack_sim_uniforms": (i32, [C.POINTER(SimParams), vp]),
    "fh_profiler_enable": (i32, [vp, u32]),
    "fh_capture_begin": (i32, [vp]),
    "fh_branch_begin": (i32, [vp]),
    "fh_branch_begin_lane": (i32, [vp, u32]),
    "fh_branch_end": (i32, [vp]),
    "fh_branch_join": (i32, [vp]),
    "fh_clear_batch_begin": (i32, [vp]),
    "fh_clear_batch_end": (i32, [vp]),
    "fh_capture_end": (i32, [vp, C.POINTER(vp)]),
    "fh_graph_launch": (i32, [vp, vp]),
    "fh_graph_destroy": (i32, [vp]),
    "fh_profiler_read": (i32, [vp, C.POINTER(KernelTime), u32, C.POINTER(u32)]),
    "fh_set_persistent_blocks": (i32, [vp, u32]),
    "fh_debug_pixel_bounds": (i32, [vp, vp, u32, vp]),
    "fh_set_raster_band_cap": (i32, [vp, u32]),
    "fh_scratch_bytes": (i32, [vp, C.POINTER(C.c_uint64)]),
    "fh_scratch_reserve": (i32, [vp, C.c_uint64]),
    "fh_fine_exhausted_bins": (i32, [vp, C.POINTER(u32)]),
    "fh_record_begin": (i32, [vp]),
    "fh_record_end": (i32, [vp, C.POINTER(vp)]),
    "fh_program_merge": (i32, [vp, C.POINTER(vp), u32, C.POINTER(vp)]),
    "fh_program_launches": (i32, [vp, C.POINTER(u32)]),
    "fh_program_launch": (i32, [vp, vp]),
    "fh_program_destroy": (i32, [vp]),
}

FH_ERR_CAPACITY = 4

_lib = None


class FhError(RuntimeError):
    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


def load(path: Path | None = None) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FhError(f"HIP library not built: {p} is missing (run __graft_entry__.build())")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = _lib.fh_last_error().decode(errors="replace") if _lib else "?"
        raise FhError(f"{what} failed with code {rc}: {msg}", rc)


def pack_render_uniforms(params: RenderParams) -> bytes:
    """fh_pack_render_uniforms: the 1184 bytes of struct Uniforms (host only)."""
    lib = load()
    out = C.create_string_buffer(1184)
    check(lib.fh_pack_render_uniforms(C.byref(params), C.cast(out, vp)), "fh_pack_render_uniforms")
    return out.raw


def pack_sim_uniforms(params: SimParams) -> bytes:
    """fh_pack_sim_uniforms: the 304 bytes of struct SimulationUniforms (host only)."""
    lib = load()
    out = C.create_string_buffer(304)
    check(lib.fh_pack_sim_uniforms(C.byref(params), C.cast(out, vp)), "fh_pack_sim_uniforms")
    return out.raw


def ptr(t) -> vp:
    """Device pointer of a torch tensor (or None)."""
    return vp(t.data_ptr()) if t is not None else vp(0)


def tex(t, width: int, height: int) -> Tex2D:
    return Tex2D(t.data_ptr(), width, height)


class Context:
    """fh_ctx bound to the current torch stream of `device`."""

    def __init__(self, device: int = 0, stream=None):
        self.lib = load()
        h = vp()
        check(self.lib.fh_create(device, C.byref(h)), "fh_create")
        self.handle = h
        if stream is not None:
            self.set_stream(stream)

    def set_stream(self, stream) -> None:
        s = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        check(self.lib.fh_set_stream(self.handle, vp(s)), "fh_set_stream")

    def call(self, name: str, *args) -> None:
        check(getattr(self.lib, name)(self.handle, *args), name)

    def sync(self) -> None:
        check(self.lib.fh_sync(self.handle), "fh_sync")

    def branch_begin(self, lane: int = 0) -> None:
        check(self.lib.fh_branch_begin_lane(self.handle, lane), "fh_branch_begin_lane")

    def branch_end(self) -> None:
        check(self.lib.fh_branch_end(self.handle), "fh_branch_end")

    def branch_join(self) -> None:
        check(self.lib.fh_branch_join(self.handle), "fh_branch_join")

    def clear_batch_begin(self) -> None:
        check(self.lib.fh_clear_batch_begin(self.handle), "fh_clear_batch_begin")

    def clear_batch_end(self) -> None:
        check(self.lib.fh_clear_batch_end(self.handle), "fh_clear_batch_end")

    def capture_begin(self) -> None:
        check(self.lib.fh_capture_begin(self.handle), "fh_capture_begin")

    def capture_end(self) -> "Graph":
        g = vp()
        check(self.lib.fh_capture_end(self.handle, C.byref(g)), "fh_capture_end")
        return Graph(self, g)

    def set_persistent_blocks(self, blocks: int) -> None:
        check(self.lib.fh_set_persistent_blocks(self.handle, blocks), "fh_set_persistent_blocks")

    def scratch_bytes(self) -> int:
        n = C.c_uint64(0)
        check(self.lib.fh_scratch_bytes(self.handle, C.byref(n)), "fh_scratch_bytes")
        return n.value

    def fine_exhausted_bins(self) -> int:
        """Bins whose slice memory the last fine pass used up (waits for the stream)."""
        n = u32(0)
        check(self.lib.fh_fine_exhausted_bins(self.handle, C.byref(n)), "fh_fine_exhausted_bins")
        return n.value

    def scratch_reserve(self, nbytes: int) -> None:
        check(self.lib.fh_scratch_reserve(self.handle, nbytes), "fh_scratch_reserve")

    def record_begin(self) -> None:
        check(self.lib.fh_record_begin(self.handle), "fh_record_begin")

    def record_end(self) -> "Program | None":
        """The recorded launches, or None if the recording had to be discarded
        because the context's scratch grew (record the same calls again)."""
        p = vp()
        rc = self.lib.fh_record_end(self.handle, C.byref(p))
        if rc == FH_ERR_CAPACITY:
            return None
        check(rc, "fh_record_end")
        return Program(self, p)

    def program_merge(self, programs: list) -> "Program":
        arr = (vp * len(programs))(*[p.handle for p in programs])
        out = vp()
        check(self.lib.fh_program_merge(self.handle, arr, len(programs), C.byref(out)), "fh_program_merge")
        return Program(self, out)

    def profiler_enable(self, max_launches: int) -> None:
        """Times the next `max_launches` kernel launches with HIP events (0 stops)."""
        check(self.lib.fh_profiler_enable(self.handle, max_launches), "fh_profiler_enable")

    def profiler_read(self) -> dict:
        """{kernel name: (launches, total ms)} of the launches timed since the last read."""
        out = (KernelTime * 64)()
        n = u32(0)
        check(self.lib.fh_profiler_read(self.handle, out, 64, C.byref(n)), "fh_profiler_read")
        return {out[i].name.decode(): (out[i].launches, out[i].total_ms) for i in range(n.value)}

    def close(self) -> None:
        if self.handle:
            self.lib.fh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Program:
    """fh_program: recorded launches of one context, or several merged into
    batched launches (launch() enqueues a merged one on the context's stream)."""

    def __init__(self, ctx: Context, handle: vp):
        self.ctx = ctx
        self.handle = handle

    def launches(self) -> int:
        n = u32(0)
        check(self.ctx.lib.fh_program_launches(self.handle, C.byref(n)), "fh_program_launches")
        return n.value

    def launch(self, ctx: Context | None = None) -> None:
        c = ctx or self.ctx
        check(c.lib.fh_program_launch(c.handle, self.handle), "fh_program_launch")

    def close(self) -> None:
        if self.handle:
            self.ctx.lib.fh_program_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """fh_graph: a recorded sequence of library calls, replayed on the context's stream."""

    def __init__(self, ctx: Context, handle: vp):
        self.ctx = ctx
        self.handle = handle

    def launch(self) -> None:
        check(self.ctx.lib.fh_graph_launch(self.ctx.handle, self.handle), "fh_graph_launch")

    def close(self) -> None:
        if self.handle:
            self.ctx.lib.fh_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Meshes:
    """An fh_meshes object list over device buffers: `meshes` is a list of
    dicts with positions and normals (V, 3) float32 and indices (T, 3) int32
    or uint32 tensors on the context's device, and `preview` (bool). The
    tensors are kept alive with the object."""

    def __init__(self, ctx: Context, meshes: list):
        self.ctx = ctx
        self._keep = []
        rows = []
        for m in meshes:
            pos, nrm, idx = m["positions"].contiguous(), m["normals"].contiguous(), m["indices"].contiguous()
            assert pos.shape == nrm.shape and pos.shape[-1] == 3 and idx.shape[-1] == 3
            self._keep += [pos, nrm, idx]
            rows.append(Mesh(pos.data_ptr(), nrm.data_ptr(), idx.data_ptr(), pos.shape[0], idx.shape[0],
                             1 if m.get("preview") else 0, 0))
        arr = (Mesh * max(1, len(rows)))(*rows)
        h = vp()
        check(load().fh_meshes_create(ctx.handle, arr, len(rows), C.byref(h)), "fh_meshes_create")
        self.handle = h

    def triangles(self) -> int:
        n = u32()
        check(load().fh_meshes_triangles(self.handle, C.byref(n)), "fh_meshes_triangles")
        return n.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(load().fh_meshes_destroy(self.handle), "fh_meshes_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

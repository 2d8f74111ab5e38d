"""Python handle on the MI355X-native sofren2 renderer.

The renderer is the C library ``lib/libsofren2_gpu.so`` (public API in
``include/sofren2.h``, additive capture/stats API in ``include/sofren2_ext.h``);
this module only loads it through ctypes for tests and benchmarks.  Loading
fails loudly when the library has not been built: there is no fallback path.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG, "lib")
GPU_LIB = os.path.join(LIB_DIR, "libsofren2_gpu.so")
SCENES_LIB = os.path.join(LIB_DIR, "libsofren2_scenes.so")

SS_MESHLET_BYTES = 2304
CLIP_RECORD_BYTES = 164
CAPTURE_VISIBILITY = 1
CAPTURE_SS_CACHE = 2
CAPTURE_CLIPS = 4
CAPTURE_SHADOWS = 8

SCENE_PLANE_SPHERES = 0
SCENE_BOX_ROOM = 1
SCENE_CROWD_SHADOWS = 2
SCENE_SERVED = 3
SCENE_MAP = 4
SCENE_FEATURES = 5
SCENE_INSTANCED = 6

# default resolution of each workload
SCENE_SIZE = {SCENE_PLANE_SPHERES: (640, 360), SCENE_BOX_ROOM: (1920, 1080), SCENE_CROWD_SHADOWS: (1920, 1080),
              SCENE_SERVED: (1920, 1080), SCENE_MAP: (1280, 720), SCENE_FEATURES: (640, 360),
              SCENE_INSTANCED: (960, 540)}


class FrameStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "meshlets_tested", "meshlets_visible", "meshlets_skinned", "batch_payloads", "clip_triangles",
        "tile_payloads", "tiles", "samples", "pixels", "shaded_pixels", "shading_groups", "tex_bilinear",
        "tex_nearest", "lightmapped_pixels", "occlusion_rays", "occlusion_nodes", "occlusion_tri_tests",
        "occlusion_invalid", "occlusion_overflow", "shadow_proxies")] + [(n, ctypes.c_float) for n in (
        "ms_upload", "ms_geometry", "ms_binning", "ms_raster_shade", "ms_download", "ms_raster", "ms_shade",
        "ms_total", "ms_shadows")] + [(n, ctypes.c_uint32) for n in (
        "tex_npot", "unlit_shadings", "dirlit_shadings", "light_cut_tiles", "bg_blend_pixels", "instances_dropped",
        "capsule_tests")]


_fp = ctypes.POINTER(ctypes.c_float)


class Arena(ctypes.Structure):
    """sft_arena (40 bytes): reserve, commit and offset of an arena."""
    _fields_ = [("base", ctypes.c_void_p), ("max_size", ctypes.c_uint64), ("committed_size", ctypes.c_uint64),
                ("offset", ctypes.c_uint64), ("commit_lock", ctypes.c_int32)]


class Mesh(ctypes.Structure):
    """sft_mesh with the reference layout (meshlet arrays are raw pointers)."""
    _fields_ = [("meshlets", ctypes.c_void_p), ("meshlet_attribs", ctypes.c_void_p), ("meshlet_skins", ctypes.c_void_p),
                ("meshlet_count", ctypes.c_uint32), ("model_meshlet_offset", ctypes.c_uint32),
                ("inds", ctypes.c_void_p)] + [(n, _fp) for n in (
                    "bounds_center_x", "bounds_center_y", "bounds_center_z", "bounds_radius", "bounds_min_x",
                    "bounds_max_x", "bounds_min_y", "bounds_max_y", "bounds_min_z", "bounds_max_z", "cone_axis_x",
                    "cone_axis_y", "cone_axis_z", "cone_cutoff")] + [
                ("culling_bone_inds", ctypes.POINTER(ctypes.c_int32)), ("material", ctypes.c_uint32),
                ("lightmap_page", ctypes.c_uint32)]


class Model(ctypes.Structure):
    """sft_model with the reference layout."""
    _fields_ = [("meshes", ctypes.POINTER(Mesh)), ("mesh_count", ctypes.c_uint32), ("total_meshlets", ctypes.c_uint32),
                ("meshlet_mesh_inds", ctypes.POINTER(ctypes.c_uint16)), ("blas_nodes", ctypes.c_void_p),
                ("blas_node_count", ctypes.c_uint32), ("blas_root_ind", ctypes.c_uint32),
                ("lightmaps", ctypes.c_void_p), ("lightmap_count", ctypes.c_uint32),
                ("shadow_proxies", ctypes.c_void_p), ("shadow_proxy_count", ctypes.c_uint32),
                ("bones", ctypes.c_void_p), ("bone_count", ctypes.c_int32),
                ("animations", ctypes.c_void_p), ("anim_count", ctypes.c_int32)]


class RawGeometry(ctypes.Structure):
    """sftx_raw_geometry: per-vertex arrays of a mesh for sftx_mesh_build_attribs."""
    _fields_ = [(n, _fp) for n in ("vx", "vy", "vz", "nx", "ny", "nz", "u", "v")] + [
        (n, ctypes.POINTER(ctypes.c_uint16)) for n in ("j0", "j1", "j2", "j3")] + [
        (n, ctypes.POINTER(ctypes.c_uint8)) for n in ("w0", "w1", "w2", "w3")]


class Tuning(ctypes.Structure):
    """sftx_tuning: the autotuner's settings and timings."""
    _fields_ = [("raster_chunk", ctypes.c_uint32 * 3), ("raster_split", ctypes.c_uint32 * 3),
                ("raster_groups", ctypes.c_uint32),
                ("raster_groups_resident", ctypes.c_uint32), ("tile_px", ctypes.c_uint32),
                ("occlusion_stack", ctypes.c_uint32), ("ms_class", ctypes.c_float * 3), ("ms_batch", ctypes.c_float),
                ("ms_occlusion", ctypes.c_float)]

    def as_dict(self):
        return {"raster_chunk_by_class": list(self.raster_chunk), "raster_split_by_class": list(self.raster_split),
                "raster_groups": self.raster_groups,
                "raster_groups_resident": self.raster_groups_resident, "tile_px": self.tile_px,
                "occlusion_stack": self.occlusion_stack,
                "visibility_ms_by_class": [round(x, 4) for x in self.ms_class],
                "visibility_ms_batch": round(self.ms_batch, 4), "occlusion_ms": round(self.ms_occlusion, 4)}


class Request(ctypes.Structure):
    """sftx_request: a render request against a scene template."""
    _fields_ = [("scene", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("samples", ctypes.c_int32), ("camera_pos", ctypes.c_float * 3), ("view", ctypes.c_float * 16),
                ("proj", ctypes.c_float * 16), ("color", ctypes.c_void_p), ("vis64", ctypes.c_void_p)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "submitted", "completed", "batches", "views_in_batches", "max_batch", "admission_waits", "resident_bytes",
        "pending", "in_flight")] + [(n, ctypes.c_double) for n in ("host_submit_ms", "host_complete_ms", "gpu_ms")] + [
        ("shadow_cache_views", ctypes.c_uint64)]


def _bind(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


class Library:
    """ctypes bindings of libsofren2_gpu.so and the scene generator."""

    def __init__(self):
        for path in (GPU_LIB, SCENES_LIB):
            if not os.path.exists(path):
                raise RuntimeError(
                    f"{path} is missing: build it with __graft_entry__.build() or python sofren2_amd/build.py")
        self.gpu = ctypes.CDLL(GPU_LIB)
        self.scenes = ctypes.CDLL(SCENES_LIB)
        g, s = self.gpu, self.scenes
        i32, u32, u8, vp = ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint8, ctypes.c_void_p
        self.sft_init = _bind(g, "sft_init", None, [i32, i32, i32])
        self.sft_release = _bind(g, "sft_release", None, [])
        self.sft_resize = _bind(g, "sft_resize", None, [i32, i32])
        self.set_capture = _bind(g, "sftx_set_capture", None, [u32])
        self.cap_grid = _bind(g, "sftx_cap_grid", None, [ctypes.POINTER(i32)] * 3)
        self.cap_visibility = _bind(g, "sftx_cap_visibility", None, [vp, vp])
        self.cap_ss_count = _bind(g, "sftx_cap_ss_count", u32, [])
        self.cap_ss_cache = _bind(g, "sftx_cap_ss_cache", None, [vp])
        self.cap_ss_visible = _bind(g, "sftx_cap_ss_visible", None, [vp])
        self.cap_clip_count = _bind(g, "sftx_cap_clip_count", u32, [])
        self.cap_clip_records = _bind(g, "sftx_cap_clip_records", None, [vp])
        self.cap_proxy_count = _bind(g, "sftx_cap_proxy_count", u32, [])
        self.cap_proxy_visibility = _bind(g, "sftx_cap_proxy_visibility", None, [vp])
        self.cap_tlas_count = _bind(g, "sftx_cap_tlas_count", u32, [])
        self.cap_tlas = _bind(g, "sftx_cap_tlas", None, [vp])
        self.set_gpu_tlas = _bind(g, "sftx_set_gpu_tlas", None, [i32])
        self.get_stats = _bind(g, "sftx_get_stats", None, [ctypes.POINTER(FrameStats)])
        self.export_frame = _bind(g, "sftx_export_frame", None, [vp])
        self.sft_present = _bind(g, "sft_present", None, [])
        self.sft_frame_begin = _bind(g, "sft_frame_begin", None, [])
        self.set_color_buffer = _bind(g, "sft_set_color_buffer", None, [vp, i32])
        self.sft_frame_end = _bind(g, "sft_frame_end", None, [])
        self.scene_record = _bind(s, "sftx_scene_record", None, [i32, vp, i32])
        self.set_output = _bind(g, "sftx_set_output", None, [i32])
        self.set_msaa = _bind(g, "sftx_set_msaa", None, [i32])
        self.set_id_split = _bind(g, "sftx_set_id_split", None, [i32])
        self.get_id_split = _bind(g, "sftx_get_id_split", i32, [])
        self.set_graphs = _bind(g, "sftx_set_graphs", None, [i32])
        self.set_stream_download = _bind(g, "sftx_set_stream_download", None, [i32])
        self._stream_stats = _bind(g, "sftx_stream_stats", None, [vp, vp])
        self.graph_captures = _bind(g, "sftx_graph_captures", ctypes.c_uint64, [])
        self.synchronize = _bind(g, "sftx_synchronize", None, [])
        self.copy_calibration = _bind(g, "sftx_copy_calibration", ctypes.c_float, [ctypes.c_uint64])
        self.scene_create = _bind(s, "sftx_scene_create", i32, [i32, u32, i32, i32])
        self.scene_render = _bind(s, "sftx_scene_render", None, [i32, vp, i32])
        self.scene_destroy = _bind(s, "sftx_scene_destroy", None, [])
        self.scene_tris = _bind(s, "sftx_scene_triangle_count", u32, [])
        self.scene_new = _bind(s, "sftx_scene_new", i32, [i32, u32, i32, i32])
        self.scene_use = _bind(s, "sftx_scene_use", i32, [i32])
        self.scene_draw = _bind(s, "sftx_scene_draw", None, [i32, i32, i32, i32])
        self.scene_free = _bind(s, "sftx_scene_free", None, [i32])
        self.scene_characters = _bind(s, "sftx_scene_characters", i32, [i32])
        self.scene_shadow_lights = _bind(s, "sftx_scene_shadow_lights", i32, [i32])
        fp = ctypes.POINTER(ctypes.c_float)
        self.scene_camera = _bind(s, "sftx_scene_camera", None, [i32, i32, i32, i32, fp, fp, fp])
        self.server_template = _bind(g, "sftx_server_template", i32, [])
        self.server_start = _bind(g, "sftx_server_start", i32, [ctypes.c_uint64, i32])
        self.server_submit = _bind(g, "sftx_server_submit", ctypes.c_uint64, [ctypes.POINTER(Request)])
        self.server_poll = _bind(g, "sftx_server_poll", i32, [ctypes.c_uint64])
        self.server_wait = _bind(g, "sftx_server_wait", i32, [ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)])
        self.server_get_stats = _bind(g, "sftx_server_get_stats", None, [ctypes.POINTER(ServerStats)])
        self.server_stop = _bind(g, "sftx_server_stop", None, [])
        self.server_release = _bind(g, "sftx_server_release", None, [])
        self.batch_begin = _bind(g, "sftx_batch_begin", None, [])
        self.view_begin = _bind(g, "sftx_view_begin", None, [i32, i32, i32])
        self.view_end = _bind(g, "sftx_view_end", i32, [vp, i32, vp])
        self.batch_submit = _bind(g, "sftx_batch_submit", None, [])
        self.batch_end = _bind(g, "sftx_batch_end", None, [])
        self.batch_submit_async = _bind(g, "sftx_batch_submit_async", ctypes.c_int32, [])
        self.device_color = _bind(g, "sftx_device_color", vp, [])
        self.texture_create = _bind(g, "sft_texture_create", u32, [i32, i32, vp])
        self.texture_release = _bind(g, "sft_texture_release", None, [u32])
        self.cap_texture = _bind(g, "sftx_cap_texture", ctypes.c_int64, [u32, vp, ctypes.c_uint64])
        self.batch_wait = _bind(g, "sftx_batch_wait", ctypes.c_int32, [ctypes.c_int32])
        self.batch_tune = _bind(g, "sftx_batch_tune", i32, [i32, ctypes.POINTER(Tuning)])
        self.get_tuning = _bind(g, "sftx_get_tuning", None, [ctypes.POINTER(Tuning)])
        self.set_tuning = _bind(g, "sftx_set_tuning", None, [ctypes.POINTER(Tuning)])
        self.scene_meshlets = _bind(s, "sftx_scene_meshlet_count", u32, [])
        self.pose_check = _bind(s, "sftx_pose_check", i32, [u32, i32, vp, vp])
        self.set_initial_caps = _bind(g, "sftx_set_initial_caps", None, [u32, u32])
        self.overflow_reruns = _bind(g, "sftx_overflow_reruns", ctypes.c_uint64, [])
        self.set_clear = _bind(g, "sft_set_clear", None, [u8])
        self.raycast_check = _bind(s, "sftx_raycast_check", i32, [u32, i32, vp])
        self.builder_check = _bind(s, "sftx_builder_check", ctypes.c_int64, [u32, i32, vp, ctypes.c_uint64])
        self.mesh_build = _bind(g, "sftx_mesh_build", None, [vp, vp, vp, vp, vp, vp, i32])
        self.mesh_build_attribs = _bind(g, "sftx_mesh_build_attribs", u32, [vp, vp, vp, u32, vp, i32, u32])
        self.model_link = _bind(g, "sftx_model_link", None, [vp, vp])
        self.blas_build = _bind(g, "sftx_blas_build", None, [vp, vp])
        self.shadow_proxies_build = _bind(g, "sftx_shadow_proxies_build", None, [vp, vp, i32])
        self.arena_init = _bind(g, "sft_arena_init", None, [vp, ctypes.c_uint64])
        self.scene_edit_room = _bind(s, "sftx_scene_edit_room", i32, [i32, ctypes.c_float])
        self.scene_room_model = _bind(s, "sftx_scene_room_model", vp, [])
        self.model_changed = _bind(g, "sftx_model_changed", i32, [vp])
        self.model_edits = _bind(g, "sftx_model_edits", ctypes.c_uint64, [])

    def stream_stats(self):
        """(batches delivered with the streamed colour download, fallbacks)"""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        self._stream_stats(ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value


_LIB = None


def library():
    global _LIB
    if _LIB is None:
        _LIB = Library()
    return _LIB


class View:
    """One view of a batch: a frame of a scene's camera path at a resolution
    and MSAA setting, with its ARGB frame and 64-bit visibility planes."""

    def __init__(self, scene_handle, frame, width, height, samples=1, vis64=True):
        self.scene, self.frame, self.width, self.height, self.samples = scene_handle, frame, width, height, samples
        self.color = np.zeros((height, width), dtype=np.uint32)
        cols, rows = (width + 63) // 64, (height + 63) // 64
        self.vis = np.zeros(cols * 64 * rows * 64 * samples, dtype=np.uint64) if vis64 else None

    def vid(self):
        return (self.vis & 0xFFFFFFFF).astype(np.uint32)

    def depth(self):
        return (self.vis >> 32).astype(np.uint32).view(np.float32)


def render_batch(lib, views):
    """Records every view and renders them with one submission."""
    lib.batch_begin()
    try:
        for v in views:
            lib.view_begin(v.width, v.height, v.samples)
            lib.scene_draw(v.scene, v.frame, v.width, v.height)
            lib.view_end(v.color.ctypes.data, v.width * 4, v.vis.ctypes.data if v.vis is not None else None)
        lib.batch_submit()
    finally:
        lib.batch_end()
    return views


def tune_batch(lib, views, repeats=3):
    """Records every view and lets the autotuner choose its settings on them
    (sftx_batch_tune); the views are rendered as by render_batch.  Returns
    the chosen settings."""
    t = Tuning()
    lib.batch_begin()
    try:
        for v in views:
            lib.view_begin(v.width, v.height, v.samples)
            lib.scene_draw(v.scene, v.frame, v.width, v.height)
            lib.view_end(v.color.ctypes.data, v.width * 4, v.vis.ctypes.data if v.vis is not None else None)
        if lib.batch_tune(repeats, ctypes.byref(t)) != 0:
            raise RuntimeError("autotuning failed")
    finally:
        lib.batch_end()
    return t


def submit_batch_async(lib, views):
    """Records every view and launches them as one batch without waiting;
    returns the ticket for lib.batch_wait.  The views' buffers must stay
    alive until then."""
    lib.batch_begin()
    try:
        for v in views:
            lib.view_begin(v.width, v.height, v.samples)
            lib.scene_draw(v.scene, v.frame, v.width, v.height)
            lib.view_end(v.color.ctypes.data, v.width * 4, v.vis.ctypes.data if v.vis is not None else None)
        ticket = lib.batch_submit_async()
    finally:
        lib.batch_end()
    if ticket <= 0:
        raise RuntimeError("batch submission failed")
    return ticket


class SceneRenderer:
    """Renders a seeded workload through the public API on the GPU."""

    def __init__(self, scene, seed=0, width=None, height=None, capture=0, samples=1, id_bits=6):
        self.lib = library()
        w, h = SCENE_SIZE.get(scene, (640, 360))
        self.width = width or w
        self.height = height or h
        self.samples = samples
        self.lib.sft_init(self.width, self.height, 0)
        self.lib.set_capture(capture)
        self.lib.set_msaa(samples)
        self.lib.set_id_split(id_bits)
        if self.lib.scene_create(scene, seed, self.width, self.height) != 0:
            raise ValueError(f"unknown scene {scene}")
        self.color = np.zeros((self.height, self.width), dtype=np.uint32)

    def render(self, frame=0):
        self.lib.scene_render(frame, self.color.ctypes.data, self.width * 4)
        return self.color

    def stats(self):
        st = FrameStats()
        self.lib.get_stats(ctypes.byref(st))
        return st

    def visibility(self):
        cols, rows, samples = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self.lib.cap_grid(ctypes.byref(cols), ctypes.byref(rows), ctypes.byref(samples))
        n = cols.value * 64 * rows.value * 64 * samples.value
        vid = np.zeros(n, dtype=np.uint32)
        depth = np.zeros(n, dtype=np.float32)
        self.lib.cap_visibility(vid.ctypes.data, depth.ctypes.data)
        return vid, depth

    def ss_cache(self):
        n = self.lib.cap_ss_count()
        ss = np.zeros(n * SS_MESHLET_BYTES, dtype=np.uint8)
        vis = np.zeros(n, dtype=np.uint8)
        if n:
            self.lib.cap_ss_cache(ss.ctypes.data)
            self.lib.cap_ss_visible(vis.ctypes.data)
        return ss.view(np.uint32).reshape(n, 9, 64), vis

    def proxy_visibility(self):
        n = self.lib.cap_proxy_count()
        vis = np.zeros((n, 64), dtype=np.uint8)
        if n:
            self.lib.cap_proxy_visibility(vis.ctypes.data)
        return vis

    def clips(self):
        n = self.lib.cap_clip_count()
        rec = np.zeros(n * CLIP_RECORD_BYTES, dtype=np.uint8)
        if n:
            self.lib.cap_clip_records(rec.ctypes.data)
        return rec.view(np.uint32).reshape(n, 41)

    def close(self):
        self.lib.scene_destroy()
        self.lib.sft_release()

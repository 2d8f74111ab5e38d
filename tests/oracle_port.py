"""Runs the C restatement (oracle/_ref/liboracle.so) on a frame recorded by the
front-end of libsofren2_gpu.so.  Recording needs no GPU."""
import ctypes
import os

import numpy as np

import sofren2_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PORT_LIB = os.path.join(ROOT, "oracle", "_ref", "liboracle.so")
DESC_BYTES = 4096  # upper bound of sizeof(sftx_frame_desc)


class PortOut(ctypes.Structure):
    _fields_ = [("color", ctypes.c_void_p), ("vid", ctypes.c_void_p), ("depth", ctypes.c_void_p),
                ("ss", ctypes.c_void_p), ("ss_vis", ctypes.c_void_p), ("clips", ctypes.c_void_p),
                ("clip_cap", ctypes.c_uint32), ("clip_count", ctypes.c_uint32), ("samples", ctypes.c_int32),
                ("proxy_vis", ctypes.c_void_p), ("invalid_packets", ctypes.c_int32)]


_PORT = None


def port_lib():
    global _PORT
    if _PORT is None:
        if not os.path.exists(PORT_LIB):
            raise RuntimeError(f"{PORT_LIB} missing: build it with make -C oracle oracle")
        lib = ctypes.CDLL(PORT_LIB)
        lib.sftx_port_render.restype = ctypes.c_int
        lib.sftx_port_render.argtypes = [ctypes.c_void_p, ctypes.POINTER(PortOut)]
        lib.sftx_port_ss_count.restype = ctypes.c_uint32
        lib.sftx_port_ss_count.argtypes = [ctypes.c_void_p]
        lib.sftx_port_proxy_count.restype = ctypes.c_int32
        lib.sftx_port_proxy_count.argtypes = [ctypes.c_void_p]
        lib.sftx_port_samples.restype = ctypes.c_int32
        lib.sftx_port_samples.argtypes = [ctypes.c_void_p]
        lib.sftx_port_grid.restype = None
        lib.sftx_port_grid.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int32)] * 4
        lib.sftx_port_build_texture.restype = ctypes.c_uint32
        lib.sftx_port_build_texture.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
        _PORT = lib
    return _PORT


def port_render_desc(desc, clip_cap=1 << 16, prefill=0):
    """Renders an exported frame description; returns the dump dict.  A
    non-zero `prefill` starts the colour buffer as the reference dump's
    pre-filled pattern (refrun.prefill_pattern)."""
    lib = port_lib()
    cols, rows, w, h = (ctypes.c_int32() for _ in range(4))
    lib.sftx_port_grid(desc, ctypes.byref(cols), ctypes.byref(rows), ctypes.byref(w), ctypes.byref(h))
    n_ss = lib.sftx_port_ss_count(desc)
    n_proxies = lib.sftx_port_proxy_count(desc)
    samples = lib.sftx_port_samples(desc)
    grid = cols.value * 64 * rows.value * 64 * samples
    color = np.zeros((h.value, w.value), dtype=np.uint32)
    if prefill:
        from tests.refrun import prefill_pattern
        color[:] = prefill_pattern(h.value, w.value, prefill)
    vid = np.zeros(grid, dtype=np.uint32)
    depth = np.zeros(grid, dtype=np.float32)
    ss = np.zeros(max(n_ss, 1) * S.SS_MESHLET_BYTES, dtype=np.uint8)
    ssvis = np.zeros(max(n_ss, 1), dtype=np.uint8)
    clips = np.zeros(clip_cap * S.CLIP_RECORD_BYTES, dtype=np.uint8)
    proxy_vis = np.zeros((max(n_proxies, 1), 64), dtype=np.uint8)
    out = PortOut(color.ctypes.data, vid.ctypes.data, depth.ctypes.data, ss.ctypes.data, ssvis.ctypes.data,
                  clips.ctypes.data, clip_cap, 0, samples, proxy_vis.ctypes.data, 0)
    rc = lib.sftx_port_render(desc, ctypes.byref(out))
    if rc != 0:
        raise RuntimeError("clip buffer too small")
    n_clip = out.clip_count
    return {"color": color, "vid": vid, "depth": depth,
            "ss": ss[:n_ss * S.SS_MESHLET_BYTES].view(np.uint32).reshape(n_ss, 9, 64), "ssvis": ssvis[:n_ss],
            "clip": clips[:n_clip * S.CLIP_RECORD_BYTES].view(np.uint32).reshape(n_clip, 41),
            "proxy_vis": proxy_vis[:n_proxies], "invalid_packets": out.invalid_packets}


def port_render_scene(scene, seed, width, height, frame=0, samples=1, prefill=0):
    """Records a frame with the product front-end and renders it with the restatement (no GPU)."""
    lib = S.library()
    lib.sft_init(width, height, 0)
    lib.set_msaa(samples)
    try:
        if lib.scene_create(scene, seed, width, height) != 0:
            raise ValueError(scene)
        color = np.zeros((height, width), dtype=np.uint32)
        lib.scene_record(frame, color.ctypes.data, width * 4)
        desc = ctypes.create_string_buffer(DESC_BYTES)
        lib.export_frame(desc)
        out = port_render_desc(desc, prefill=prefill)
        lib.sft_frame_end()
        lib.scene_destroy()
    finally:
        lib.sft_release()
    return out


def port_mip_chain(level0):
    """The oracle's swizzled mip chain of a (h, w) uint32 image."""
    lib = port_lib()
    h, w = level0.shape
    src = np.ascontiguousarray(level0, dtype=np.uint32)
    n = lib.sftx_port_build_texture(src.ctypes.data, w, h, None)
    out = np.zeros(n, dtype=np.uint32)
    lib.sftx_port_build_texture(src.ctypes.data, w, h, out.ctypes.data)
    return out

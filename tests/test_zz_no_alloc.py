"""sftx_present_noalloc (no device allocation, intermediates in the library's
static scratch array) against sft_present on the same recorded frames.

Calls in a row, each with a smaller input than the one before (the scratch
keeps the previous call's contents): the 1920x1080 box room, the box room at
960x540 with 2x MSAA, the plane of spheres at 320x180, each bit-identical to
sft_present's colour; then the
refusals: a frame too large for the scratch, a scene not yet resident, and
frames still in flight.
"""
import ctypes

import numpy as np
import pytest

import sofren2_amd as S

pytestmark = pytest.mark.gpu

FULL = (S.SCENE_BOX_ROOM, 3, 1920, 1080, 1)      # kind, frame, width, height, MSAA samples
BIG = (S.SCENE_BOX_ROOM, 5, 960, 540, 2)
SMALL = (S.SCENE_PLANE_SPHERES, 0, 320, 180, 1)


def _lib():
    lib = S.library()
    g = lib.gpu
    g.sftx_present_noalloc.restype = ctypes.c_int32
    g.sftx_present_noalloc.argtypes = []
    g.sftx_noalloc_scratch_bytes.restype = ctypes.c_uint64
    return lib, g


def _record(lib, handle, frame, w, h, samples, buf):
    lib.sft_resize(w, h)
    lib.set_msaa(samples)
    lib.set_color_buffer(buf.ctypes.data, w * 4)
    lib.sft_frame_begin()
    lib.scene_draw(handle, frame, w, h)


def test_noalloc_matches_present_calls_with_shrinking_inputs():
    lib, g = _lib()
    lib.sft_init(FULL[2], FULL[3], 0)
    try:
        handles = {k: lib.scene_new(k, 0, 640, 360) for k in (FULL[0], SMALL[0])}
        assert min(handles.values()) >= 0
        expected = []
        for kind, frame, w, h, ns in (FULL, BIG, SMALL):
            buf = np.zeros((h, w), dtype=np.uint32)
            _record(lib, handles[kind], frame, w, h, ns, buf)
            lib.sft_present()
            lib.sft_frame_end()
            expected.append(buf)
        print("scratch bytes", g.sftx_noalloc_scratch_bytes(), flush=True)
        got = []
        for kind, frame, w, h, ns in (FULL, BIG, SMALL):
            buf = np.full((h, w), 0xDEADBEEF, dtype=np.uint32)
            _record(lib, handles[kind], frame, w, h, ns, buf)
            print("noalloc", kind, w, h, ns, flush=True)
            rc = g.sftx_present_noalloc()
            st = S.FrameStats()
            lib.get_stats(ctypes.byref(st))
            lib.sft_frame_end()
            assert rc == 0, rc
            print("noalloc done", kind, "shaded", st.shaded_pixels, "tile payloads", st.tile_payloads, flush=True)
            assert st.shaded_pixels > 0
            got.append(buf)
        for e, x in zip(expected, got):
            np.testing.assert_array_equal(x, e)

        # refusals: nothing is rendered and the buffer is left alone
        buf = np.full((2160, 3840), 7, dtype=np.uint32)
        _record(lib, handles[BIG[0]], 5, 3840, 2160, 4, buf)
        assert g.sftx_present_noalloc() == -4      # TOO_LARGE
        lib.sft_frame_end()
        assert (buf == 7).all()
        hf = lib.scene_new(S.SCENE_FEATURES, 0, 640, 360)  # new textures, not resident yet
        buf = np.full((360, 640), 7, dtype=np.uint32)
        _record(lib, hf, 0, 640, 360, 1, buf)
        assert g.sftx_present_noalloc() == -3      # NOT_RESIDENT
        lib.sft_frame_end()
        assert (buf == 7).all()
        lib.set_clear(1)                           # the features scene turned clearing off
        lib.set_output(1)                          # device output: frames stay in flight
        _record(lib, handles[SMALL[0]], 0, 320, 180, 1, np.zeros((180, 320), dtype=np.uint32))
        lib.sft_present()
        lib.sft_frame_end()
        _record(lib, handles[SMALL[0]], 0, 320, 180, 1, buf)
        assert g.sftx_present_noalloc() == -2      # IN_FLIGHT
        lib.sft_frame_end()
        lib.synchronize()
        lib.set_output(0)

        # after the refusals the path still renders, again equal to sft_present
        kind, frame, w, h, ns = SMALL
        buf = np.full((h, w), 0xDEADBEEF, dtype=np.uint32)
        _record(lib, handles[kind], frame, w, h, ns, buf)
        assert g.sftx_present_noalloc() == 0
        lib.sft_frame_end()
        np.testing.assert_array_equal(buf, expected[2])
    finally:
        lib.scene_destroy()
        lib.sft_release()

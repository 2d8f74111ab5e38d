"""Host colour streamed band by band while the frame shades (the default)
against the download after the frame: the same bits in the caller's buffer,
for blocking presents (1x and MSAA, padded pitch, no clearing over a
pre-filled buffer), asynchronous host-output batches and served requests;
and every streamed batch delivered without falling back."""
import numpy as np
import pytest

import sofren2_amd as S
from sofren2_amd import workloads
from tests.refrun import prefill_pattern

pytestmark = pytest.mark.gpu


def _presents(stream, scene, w, h, frames, samples=1, pad=0, prefill=0):
    lib = S.library()
    lib.sft_init(w, h, 0)
    try:
        lib.set_stream_download(stream)
        assert lib.scene_create(scene, 0, w, h) == 0
        lib.set_msaa(samples)
        s0 = lib.stream_stats()
        outs = []
        for f in frames:
            buf = prefill_pattern(h, w + pad // 4, prefill) if prefill else np.zeros((h, w + pad // 4), np.uint32)
            lib.scene_render(f, buf.ctypes.data, w * 4 + pad)
            outs.append(buf)
        s1 = lib.stream_stats()
        lib.set_msaa(1)
        lib.scene_destroy()
    finally:
        lib.set_stream_download(1)
        lib.sft_release()
    return outs, (s1[0] - s0[0], s1[1] - s0[1])


@pytest.mark.parametrize("scene,w,h,samples,pad,prefill", [
    (S.SCENE_BOX_ROOM, 1920, 1080, 1, 0, 0),
    (S.SCENE_CROWD_SHADOWS, 1280, 720, 4, 0, 0),
    (S.SCENE_FEATURES, 640, 360, 1, 64, 7),
    (S.SCENE_FEATURES, 700, 333, 2, 32, 9),
])
def test_streamed_presents_equal_download_after_frame(scene, w, h, samples, pad, prefill):
    frames = [0, 13, 77]
    on, st_on = _presents(1, scene, w, h, frames, samples, pad, prefill)
    off, st_off = _presents(0, scene, w, h, frames, samples, pad, prefill)
    print("stream stats", st_on, st_off, flush=True)
    assert st_on == (len(frames), 0), st_on
    assert st_off == (0, 0), st_off
    for a, b in zip(on, off):
        np.testing.assert_array_equal(a, b)


def test_streamed_async_batches_and_serving():
    lib = S.library()
    lib.sft_init(640, 360, 0)
    try:
        results = {}
        for stream in (1, 0):
            lib.set_stream_download(stream)
            mb = workloads.MixedBatch(lib, views=workloads.mixed_views(16, seed=3), vis64=True)
            s0 = lib.stream_stats()
            mb.render_pipelined(3)
            s1 = lib.stream_stats()
            results[stream] = ([v.color.copy() for v in mb.views], [v.vis.copy() for v in mb.views],
                               (s1[0] - s0[0], s1[1] - s0[1]))
            mb.close()
        print("batch stream stats", results[1][2], results[0][2], flush=True)
        assert results[1][2][0] >= 3 and results[1][2][1] == 0
        for a, b in zip(results[1][0], results[0][0]):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(results[1][1], results[0][1]):
            np.testing.assert_array_equal(a, b)

        specs = [(1, 21, 5), (2, 22, 40)]
        reqs = [(0, 17, 640, 360, 1), (1, 40, 480, 300, 2), (0, 99, 1024, 576, 4), (1, 3, 320, 180, 1)]
        served = {}
        for stream in (1, 0):
            lib.set_stream_download(stream)
            srv = workloads.Server(lib, specs, budget_bytes=4 << 30, max_views=4)
            try:
                outs = [np.zeros((h, w), np.uint32) for _, _, w, h, _ in reqs]
                ids = [srv.submit(srv.request(t, cf, w, h, ns, c, None)) for (t, cf, w, h, ns), c in zip(reqs, outs)]
                for i in ids:
                    srv.wait(i)
            finally:
                srv.close()
            served[stream] = outs
        for a, b in zip(served[1], served[0]):
            np.testing.assert_array_equal(a, b)
    finally:
        lib.set_stream_download(1)
        lib.scene_destroy()
        lib.sft_release()

"""Batches whose clip pool or tile lists overflow are run again with larger
intermediates; the outputs must be those of a batch that never overflowed.

The frame sets start with small capacities (sftx_set_initial_caps) so that
every batch overflows once.  Cases: a blocking present; an asynchronous
batch of views that do not clear, into pre-filled buffers at a padded pitch
(its re-run must start again from the caller's pixels, which the first
attempt must not have overwritten), while the next batch makes a new model
resident and so moves the first batch's transient rows; the request server.
"""
import numpy as np
import pytest

import sofren2_amd as S
from sofren2_amd import workloads
from tests.refrun import prefill_pattern

pytestmark = pytest.mark.gpu

W, H, PAD = 640, 360, 64
SMALL = (64, 256)


def _features_buffer(seed):
    return prefill_pattern(H, W + PAD // 4, seed)


def _render_present(small):
    lib = S.library()
    lib.sft_init(W, H, 0)
    try:
        assert lib.scene_create(S.SCENE_FEATURES, 0, W, H) == 0
        if small:
            lib.set_initial_caps(*SMALL)
        before = lib.overflow_reruns()
        buf = _features_buffer(3)
        lib.scene_render(0, buf.ctypes.data, W * 4 + PAD)
        reruns = lib.overflow_reruns() - before
        lib.scene_destroy()
    finally:
        lib.sft_release()
    return buf, reruns


def test_blocking_present_rerun_equals_plain_present():
    plain, r0 = _render_present(False)
    small, r1 = _render_present(True)
    print("blocking reruns", r0, r1, flush=True)
    assert r0 == 0 and r1 >= 1
    np.testing.assert_array_equal(small, plain)


def _async_pair(small):
    """Batch A: two features views (no clearing, padded pre-filled buffers)
    submitted asynchronously; batch B: a box-room view, whose preparation
    makes the box-room model resident and grows the pools, submitted before
    A is waited for.  Returns A's and B's buffers and the re-runs."""
    lib = S.library()
    lib.sft_init(W, H, 0)
    try:
        hf = lib.scene_new(S.SCENE_FEATURES, 0, W, H)
        hb = lib.scene_new(S.SCENE_BOX_ROOM, 0, W, H)
        assert hf >= 0 and hb >= 0
        if small:
            lib.set_initial_caps(*SMALL)
        before = lib.overflow_reruns()
        a_bufs = [_features_buffer(11), _features_buffer(12)]
        lib.batch_begin()
        lib.set_clear(0)
        for frame, buf in zip((0, 37), a_bufs):
            lib.view_begin(W, H, 1)
            lib.scene_draw(hf, frame, W, H)
            lib.view_end(buf.ctypes.data, W * 4 + PAD, None)
        ta = lib.batch_submit_async()
        lib.batch_end()
        b_buf = np.zeros((H, W), dtype=np.uint32)
        lib.batch_begin()
        lib.set_clear(1)
        lib.view_begin(W, H, 1)
        lib.scene_draw(hb, 5, W, H)
        lib.view_end(b_buf.ctypes.data, W * 4, None)
        tb = lib.batch_submit_async()
        lib.batch_end()
        assert ta > 0 and tb > 0
        assert lib.batch_wait(ta) == 0 and lib.batch_wait(tb) == 0
        reruns = lib.overflow_reruns() - before
        lib.scene_destroy()
    finally:
        lib.sft_release()
    return a_bufs, b_buf, reruns


def test_async_rerun_after_layout_move_equals_plain_batches():
    a0, b0, r0 = _async_pair(False)
    a1, b1, r1 = _async_pair(True)
    print("async reruns", r0, r1, flush=True)
    assert r0 == 0 and r1 >= 2
    for x, y in zip(a0, a1):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(b0, b1)
    # the caller's pixels show through where the views draw nothing
    assert (a1[0][:, :W] == _features_buffer(11)[:, :W]).mean() > 0.3


SPECS = [(5, 0, 0), (1, 21, 5), (0, 23, 0)]
REQUESTS = [(0, 10, 640, 360, 1), (1, 40, 480, 300, 2), (2, 0, 256, 256, 4), (0, 20, 320, 200, 4),
            (1, 3, 320, 180, 1), (2, 0, 700, 400, 2)]


def _served(small):
    lib = S.library()
    lib.sft_init(W, H, 0)
    srv = None
    try:
        if small:
            lib.set_initial_caps(*SMALL)
        before = lib.overflow_reruns()
        srv = workloads.Server(lib, SPECS, budget_bytes=8 << 30, max_views=4)
        outs = [np.zeros((h, w), dtype=np.uint32) for _, _, w, h, _ in REQUESTS]
        ids = [srv.submit(srv.request(t, cf, w, h, ns, c, None)) for (t, cf, w, h, ns), c in zip(REQUESTS, outs)]
        for i in ids:
            srv.wait(i)
        srv.close()
        srv = None
        reruns = lib.overflow_reruns() - before
    finally:
        if srv is not None:
            srv.close()
        lib.scene_destroy()
        lib.sft_release()
    return outs, reruns


def test_served_reruns_equal_plain_serving():
    plain, r0 = _served(False)
    small, r1 = _served(True)
    print("served reruns", r0, r1, flush=True)
    assert r0 == 0 and r1 >= 1
    for x, y in zip(plain, small):
        np.testing.assert_array_equal(x, y)

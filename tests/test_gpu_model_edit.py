"""Edits of a resident model after its first draw.

The reference reads a model's arrays at every present (sofren2.c:8139-8289);
this renderer keeps drawn models in HBM.  An edit must give the frame that a
fresh renderer, which never saw the old model, gives for the edited model:
moving every meshlet is seen by the draw-time check alone; moving only the
meshlets between each mesh's first and last is not, and is declared with
sftx_model_changed.
"""
import numpy as np
import pytest

import sofren2_amd as S

pytestmark = pytest.mark.gpu

W, H, FRAME, DY = 640, 360, 7, 0.5


def _render(lib):
    color = np.zeros((H, W), dtype=np.uint32)
    lib.scene_render(FRAME, color.ctypes.data, W * 4)
    return color


def _fresh(mode):
    lib = S.library()
    lib.sft_init(W, H, 0)
    try:
        assert lib.scene_create(S.SCENE_BOX_ROOM, 0, W, H) == 0
        assert lib.scene_edit_room(mode, DY) > 0
        return _render(lib)
    finally:
        lib.scene_destroy()
        lib.sft_release()


def test_rewritten_model_is_detected_at_draw():
    lib = S.library()
    lib.sft_init(W, H, 0)
    try:
        assert lib.scene_create(S.SCENE_BOX_ROOM, 0, W, H) == 0
        before = _render(lib)
        e0 = lib.model_edits()
        moved = lib.scene_edit_room(1, DY)
        assert moved > 0
        after = _render(lib)
        print("meshlets moved", moved, "edits", lib.model_edits() - e0, flush=True)
        assert lib.model_edits() - e0 == 1
        again = _render(lib)                # no further upload
        assert lib.model_edits() - e0 == 1
    finally:
        lib.scene_destroy()
        lib.sft_release()
    fresh = _fresh(1)
    assert (after != before).mean() > 0.01
    np.testing.assert_array_equal(after, fresh)
    np.testing.assert_array_equal(again, fresh)


def test_declared_edit_is_uploaded_again():
    lib = S.library()
    lib.sft_init(W, H, 0)
    try:
        assert lib.scene_create(S.SCENE_BOX_ROOM, 0, W, H) == 0
        before = _render(lib)
        e0 = lib.model_edits()
        assert lib.scene_edit_room(2, DY) > 0
        undeclared = _render(lib)           # the draw-time check does not see it
        assert lib.model_edits() == e0
        assert lib.model_changed(lib.scene_room_model()) == 0
        assert lib.model_edits() - e0 == 1
        declared = _render(lib)
        assert lib.model_changed(None) == -1          # SFTX_MODEL_UNKNOWN
    finally:
        lib.scene_destroy()
        lib.sft_release()
    fresh = _fresh(2)
    np.testing.assert_array_equal(undeclared, before)
    assert (declared != before).mean() > 0.01
    np.testing.assert_array_equal(declared, fresh)

"""GPU render path vs the reference renderer on the same seeded scenes.

Bar: screen-space meshlet cache, cull result, clipped-triangle records,
per-pixel visibility IDs and depth bit-exact; colour within 1 LSB on at least
99.9% of pixels (the test also reports the exact-match fraction).
"""
import numpy as np
import pytest

import sofren2_amd as S
from tests import parity
from tests.refrun import ref_dump

pytestmark = pytest.mark.gpu


def render_gpu(scene, seed, w, h, frame=0):
    r = S.SceneRenderer(scene, seed, w, h, capture=S.CAPTURE_VISIBILITY | S.CAPTURE_SS_CACHE | S.CAPTURE_CLIPS)
    try:
        color = r.render(frame).copy()
        vid, depth = r.visibility()
        ss, ssvis = r.ss_cache()
        clip = r.clips()
        st = r.stats()
    finally:
        r.close()
    return {"color": color, "vid": vid, "depth": depth, "ss": ss, "ssvis": ssvis, "clip": clip}, st


@pytest.mark.parametrize("seed", [0, 1])
def test_plane_spheres_parity(seed):
    w, h = S.SCENE_SIZE[S.SCENE_PLANE_SPHERES]
    gpu, st = render_gpu(S.SCENE_PLANE_SPHERES, seed, w, h)
    ref = ref_dump(S.SCENE_PLANE_SPHERES, seed, w, h)
    rep = parity.compare_frame(gpu, ref)
    print("plane_spheres seed", seed, rep, "visible", st.meshlets_visible, "ms", st.ms_total, flush=True)
    parity.assert_parity(rep)


@pytest.mark.parametrize("frame", [0, 37])
def test_box_room_parity(frame):
    w, h = S.SCENE_SIZE[S.SCENE_BOX_ROOM]
    gpu, st = render_gpu(S.SCENE_BOX_ROOM, 0, w, h, frame)
    ref = ref_dump(S.SCENE_BOX_ROOM, 0, w, h, frame)
    rep = parity.compare_frame(gpu, ref)
    print("box_room frame", frame, rep, "visible", st.meshlets_visible, "clips", st.clip_triangles,
          "ms", st.ms_total, flush=True)
    parity.assert_parity(rep)

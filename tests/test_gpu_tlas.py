"""The frame's top-level BVH built on the device (k_tlas) against the host
builder (tlas.cpp, the reference's grouping, sofren2.c:2895-3062 and
3619-3703): identical node bytes, the same shadow-proxy light visibility
and the same frame, on crowd frames, on served-scene variants with shadows
(box lattices of 16^3 and 25^3, 5 to 43 humanoids, 2 to 16 lights) and on
the instanced scene (41 instances with a BLAS, so the grouping's three
halving levels and the zeroed items of slots drawn without a BLAS are
exercised).  The instanced scene is also compared with the reference
renderer: per-proxy light visibility exact, the frame to the parity bar."""
import ctypes

import numpy as np
import pytest

import sofren2_amd as S
from sofren2_amd import workloads
from tests import parity
from tests.refrun import ref_dump
from tests.test_gpu_parity import render_gpu

pytestmark = pytest.mark.gpu

NODE_BYTES = 256
# (scene kind, seed, frame); served frames take the template pose and a camera
CASES = [(S.SCENE_CROWD_SHADOWS, 0, f) for f in (0, 17, 34, 51, 68, 85, 102, 119)] + [
    (S.SCENE_SERVED, seed, workloads.reference_frame_code(pose, cam))
    for seed, pose, cam in ((35, 10, 3), (45, 60, 77), (82, 30, 17), (71, 5, 40), (11, 90, 0), (146, 33, 66))] + [
    (S.SCENE_INSTANCED, seed, f) for seed, f in ((0, 0), (0, 45), (1, 90), (2, 7))]


def _frame(lib, scene, seed, frame, gpu_tlas):
    w, h = 960, 540
    color = np.zeros((h, w), dtype=np.uint32)
    lib.sft_init(w, h, 0)
    try:
        lib.set_gpu_tlas(gpu_tlas)
        lib.set_capture(S.CAPTURE_SHADOWS)
        assert lib.scene_create(scene, seed, w, h) == 0
        lib.scene_render(frame, color.ctypes.data, w * 4)
        n = lib.cap_tlas_count()
        nodes = np.zeros(n * NODE_BYTES, dtype=np.uint8)
        if n:
            lib.cap_tlas(nodes.ctypes.data)
        npx = lib.cap_proxy_count()
        vis = np.zeros((npx, 64), dtype=np.uint8)
        if npx:
            lib.cap_proxy_visibility(vis.ctypes.data)
        st = S.FrameStats()
        lib.get_stats(ctypes.byref(st))
        lib.set_capture(0)
        lib.scene_destroy()
    finally:
        lib.sft_release()
    return nodes, vis, color, st


@pytest.mark.parametrize("scene,seed,frame", CASES)
def test_device_tlas_bytes_equal_host_builder(scene, seed, frame):
    lib = S.library()
    host = _frame(lib, scene, seed, frame, 0)
    dev = _frame(lib, scene, seed, frame, 1)
    print("case", scene, seed, frame, "nodes", host[0].size // NODE_BYTES, "proxies", host[1].shape[0],
          "rays", dev[3].occlusion_rays, flush=True)
    assert host[0].size >= NODE_BYTES and dev[3].occlusion_rays > 0
    if scene == S.SCENE_INSTANCED:
        assert host[0].size // NODE_BYTES > 1, "the grouping levels run"
    np.testing.assert_array_equal(dev[0], host[0])
    np.testing.assert_array_equal(dev[1], host[1])
    np.testing.assert_array_equal(dev[2], host[2])


@pytest.mark.parametrize("frame", [0, 60])
def test_instanced_scene_matches_reference(frame):
    w, h = S.SCENE_SIZE[S.SCENE_INSTANCED]
    gpu, st = render_gpu(S.SCENE_INSTANCED, 0, w, h, frame)
    ref = ref_dump(S.SCENE_INSTANCED, 0, w, h, frame)
    print("instanced frame", frame, "rays", st.occlusion_rays, "nodes", st.occlusion_nodes, "tri tests",
          st.occlusion_tri_tests, flush=True)
    assert st.occlusion_invalid == 0 and st.occlusion_overflow == 0 and st.occlusion_rays > 0
    np.testing.assert_array_equal(gpu["proxy_vis"], ref["proxy_vis"])
    rep = parity.compare_frame(gpu, ref)
    print("instanced frame", frame, rep, flush=True)
    parity.assert_parity(rep)

"""Host-side queries of the API on the CPU: skeletal pose evaluation
(sft_pose_create / sft_pose_update, sofren2.c:8940-9107) and the pick ray
sft_raycast_instance (sofren2.c:3733-3838), word for word against the
reference build running the same seeded checks (sftx_pose_check,
sftx_raycast_check in the scene generator), plus properties that hold
without the reference."""
import ctypes

import numpy as np
import pytest

import sofren2_amd as S
from tests.refrun import have_ref, ref_pose, ref_rays

TIMES = np.concatenate([np.float32(-0.3) + np.float32(0.137) * np.arange(40, dtype=np.float32),
                        np.array([0.0, 1e-6, 2.5, 7.25, 100.0], dtype=np.float32)]).astype(np.float32)


def our_pose(seed, times):
    lib = S.library()
    times = np.ascontiguousarray(times, dtype=np.float32)
    out = np.zeros(len(times) * 2 * 12 * 16, dtype=np.float32)
    n = lib.pose_check(seed, len(times), times.ctypes.data, out.ctypes.data)
    assert n == out.size
    return out.reshape(-1, 2, 12, 16)


def our_rays(scene, seed, n):
    lib = S.library()
    lib.sft_init(64, 64, 0)
    try:
        assert lib.scene_create(scene, 0, 64, 64) == 0
        out = np.zeros(n * 8, dtype=np.float32)
        got = lib.raycast_check(seed, n, out.ctypes.data)
        lib.scene_destroy()
    finally:
        lib.sft_release()
    assert got == n
    return out.reshape(n, 8)


def test_pose_properties():
    """Without the reference: posed = inverse_bind * global, and a bone's
    global matrix is its local one times its parent's (row vectors), so
    global(child) * inverse(global(parent)) is a product of rotation, scale
    and translation only: its last column is (0, 0, 0, 1)."""
    p = our_pose(3, TIMES[:6])
    assert np.isfinite(p).all()
    g = p[:, 0].reshape(-1, 12, 4, 4)
    assert np.allclose(g[..., :, 3], [0, 0, 0, 1], atol=1e-5)


@pytest.mark.skipif(not have_ref(), reason="reference checker build missing")
@pytest.mark.parametrize("seed", [0, 1, 7])
def test_pose_matches_reference(seed):
    ours = our_pose(seed, TIMES)
    ref = ref_pose(seed, TIMES)
    assert ours.shape == ref.shape
    diff = ours.view(np.uint32) != ref.view(np.uint32)
    assert not diff.any(), np.argwhere(diff)[:5]


@pytest.mark.skipif(not have_ref(), reason="reference checker build missing")
def test_raycast_matches_reference():
    """2000 seeded rays at the crowd room's BLAS through a rotated,
    translated instance: hit flag, distance, position and normal equal word
    for word; some rays hit and some miss or stop at max_dist."""
    ours = our_rays(S.SCENE_CROWD_SHADOWS, 5, 2000)
    ref = ref_rays(S.SCENE_CROWD_SHADOWS, 5, 2000)
    assert ours.shape == ref.shape
    diff = ours.view(np.uint32) != ref.view(np.uint32)
    assert not diff.any(), (np.argwhere(diff)[:5], ours[np.argwhere(diff)[0][0]], ref[np.argwhere(diff)[0][0]])
    hits = ours[:, 0] == 1
    assert 0.2 < hits.mean() < 0.999


def test_raycast_brute_force_distance():
    """Without the reference: every reported hit lies on the ray at the
    reported distance, and its normal is unit length."""
    rays = our_rays(S.SCENE_CROWD_SHADOWS, 9, 300)
    hits = rays[rays[:, 0] == 1]
    assert len(hits) > 30
    n = np.linalg.norm(hits[:, 5:8], axis=1)
    assert np.allclose(n, 1.0, atol=1e-5)
    assert (hits[:, 1] > 0).all()


def test_raycast_without_blas_misses():
    """A model without a BLAS (the box-room workload keeps none) reports no
    hit at max_dist, as the reference does."""
    lib = S.library()
    lib.sft_init(64, 64, 0)
    try:
        assert lib.scene_create(S.SCENE_BOX_ROOM, 0, 64, 64) == 0
        out = np.zeros(8 * 8, dtype=np.float32)
        lib.raycast_check(2, 8, out.ctypes.data)
        lib.scene_destroy()
    finally:
        lib.sft_release()
    r = out.reshape(8, 8)
    assert (r[:, 0] == 0).all()

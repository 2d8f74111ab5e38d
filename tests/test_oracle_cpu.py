"""The C restatement (oracle) against the reference renderer, on the CPU.

The restatement renders the frame the product's front-end records; the
reference renders the same seeded scene through its own front-end.  Every
intermediate must agree bit for bit: this pins the oracle (and the
front-end's host math) to the reference.
"""
import numpy as np
import pytest

import sofren2_amd as S
from tests import parity
from tests.oracle_port import port_render_scene
from tests.refrun import ref_dump, have_ref, prefill_pattern

pytestmark = pytest.mark.skipif(not have_ref(), reason="reference checker build missing")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_port_matches_reference_plane_spheres(seed):
    w, h = S.SCENE_SIZE[S.SCENE_PLANE_SPHERES]
    port = port_render_scene(S.SCENE_PLANE_SPHERES, seed, w, h)
    ref = ref_dump(S.SCENE_PLANE_SPHERES, seed, w, h)
    rep = parity.compare_frame(port, ref)
    print(rep)
    parity.assert_parity(rep)
    assert rep["color_exact_frac"] == 1.0, rep


def test_reference_dump_is_sane():
    w, h = S.SCENE_SIZE[S.SCENE_PLANE_SPHERES]
    ref = ref_dump(S.SCENE_PLANE_SPHERES, 0, w, h)
    assert ref["meta"]["ss_count"] == 569
    assert ref["ssvis"].sum() > 0
    covered = (ref["vid"] != 0xFFFFFFFF).mean()
    assert 0.2 < covered < 0.95
    assert (ref["color"] >> 24 == 0xFF).all()


@pytest.mark.parametrize("frame", [0, 37])
def test_port_matches_reference_box_room(frame):
    w, h = S.SCENE_SIZE[S.SCENE_BOX_ROOM]
    port = port_render_scene(S.SCENE_BOX_ROOM, 0, w, h, frame)
    ref = ref_dump(S.SCENE_BOX_ROOM, 0, w, h, frame)
    rep = parity.compare_frame(port, ref)
    print(rep, "clips", rep["clip_count"])
    parity.assert_parity(rep)
    assert rep["color_exact_frac"] == 1.0, rep


@pytest.mark.parametrize("frame", [0, 50])
def test_port_matches_reference_crowd_shadows(frame):
    """Skinned crowd with capsule shadows: skinned screen-space cache, per-proxy
    light visibility from the occlusion packets, and the shadowed colour."""
    w, h = S.SCENE_SIZE[S.SCENE_CROWD_SHADOWS]
    port = port_render_scene(S.SCENE_CROWD_SHADOWS, 0, w, h, frame)
    ref = ref_dump(S.SCENE_CROWD_SHADOWS, 0, w, h, frame)
    assert port["invalid_packets"] == 0
    assert ref["proxy_vis"].shape == port["proxy_vis"].shape
    np.testing.assert_array_equal(port["proxy_vis"], ref["proxy_vis"])
    vis = ref["proxy_vis"][:, :8]
    assert (vis == 3).any() and (vis < 3).any(), "occlusion pass must see both lit and occluded proxies"
    rep = parity.compare_frame(port, ref)
    print(rep, "clips", rep["clip_count"])
    parity.assert_parity(rep)
    assert rep["color_exact_frac"] == 1.0, rep


@pytest.mark.parametrize("samples", [2, 4])
@pytest.mark.parametrize("scene,frame", [(S.SCENE_PLANE_SPHERES, 0), (S.SCENE_BOX_ROOM, 37)])
def test_port_matches_reference_msaa(scene, frame, samples):
    """Rotated-grid MSAA: per-sample IDs and depth, sample-weighted resolve."""
    if not have_ref(samples):
        pytest.skip("MSAA reference build missing")
    w, h = S.SCENE_SIZE[scene]
    port = port_render_scene(scene, 0, w, h, frame, samples=samples)
    ref = ref_dump(scene, 0, w, h, frame, samples=samples)
    rep = parity.compare_frame(port, ref)
    print(rep)
    parity.assert_parity(rep)
    assert rep["color_exact_frac"] == 1.0, rep


@pytest.mark.parametrize("seed", [10, 12, 14])
def test_port_matches_reference_served_variants(seed):
    """Served-scene variants (box lattice size, humanoid and shadow-light
    counts drawn from the seed) at a small resolution."""
    w, h = 480, 272
    port = port_render_scene(S.SCENE_SERVED, seed, w, h, 1000 * 8 + 21)
    ref = ref_dump(S.SCENE_SERVED, seed, w, h, 1000 * 8 + 21)
    rep = parity.compare_frame(port, ref)
    print(rep)
    parity.assert_parity(rep)
    assert rep["color_exact_frac"] == 1.0, rep
    assert np.array_equal(port["proxy_vis"], ref["proxy_vis"])


@pytest.mark.parametrize("frame", [0, 40])
def test_port_matches_reference_map_file(frame):
    """The reference's own lightmapped map (res/simple.qsf), lit by its
    game-data lights."""
    from tests.mapfile import map_path
    map_path()
    w, h = S.SCENE_SIZE[S.SCENE_MAP]
    port = port_render_scene(S.SCENE_MAP, 0, w, h, frame)
    ref = ref_dump(S.SCENE_MAP, 0, w, h, frame)
    rep = parity.compare_frame(port, ref)
    print(rep)
    parity.assert_parity(rep)
    assert rep["color_exact_frac"] == 1.0, rep
    covered = (ref["vid"] != 0xFFFFFFFF).mean()
    assert covered > 0.3, covered


@pytest.mark.parametrize("samples", [1, 2, 4])
@pytest.mark.parametrize("frame", [0, 37])
def test_port_matches_reference_features(frame, samples):
    """The features workload: non-power-of-two textures (edge clamp,
    bilinear and nearest), unlit materials, directional lights, more than 16
    lights reaching a tile, no clearing over a pre-filled buffer with 64
    bytes of row padding (the reference leaves the padding alone)."""
    if not have_ref(samples):
        pytest.skip("MSAA reference build missing")
    w, h = S.SCENE_SIZE[S.SCENE_FEATURES]
    port = port_render_scene(S.SCENE_FEATURES, 0, w, h, frame, samples=samples, prefill=7)
    ref = ref_dump(S.SCENE_FEATURES, 0, w, h, frame, samples=samples, pad=64, prefill=7)
    rep = parity.compare_frame(port, ref)
    print(rep)
    parity.assert_parity(rep)
    assert rep["color_exact_frac"] == 1.0, rep
    assert np.array_equal(ref["color_buffer"][:, w:], prefill_pattern(h, w + 16, 7)[:, w:])
    covered = (ref["vid"] != 0xFFFFFFFF).mean()
    assert 0.2 < covered < 0.6, covered
    # the caller's buffer shows through where nothing is drawn
    assert (ref["color"] == prefill_pattern(h, w, 7)).mean() > 0.3


def test_port_drops_instances_beyond_64():
    """Frames 500-999 of the features workload draw six more instances after
    the 64 the visibility-ID split holds; the front-end drops them, so the
    frame equals the same camera frame without them (the reference itself
    cannot render such a frame: its TLAS build reads the dropped slots)."""
    w, h = S.SCENE_SIZE[S.SCENE_FEATURES]
    a = port_render_scene(S.SCENE_FEATURES, 0, w, h, 537, prefill=7)
    b = port_render_scene(S.SCENE_FEATURES, 0, w, h, 37, prefill=7)
    for k in ("color", "vid", "depth"):
        assert np.array_equal(a[k], b[k]), k

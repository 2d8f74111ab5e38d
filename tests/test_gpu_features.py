"""GPU render path vs the reference on the features workload: the shading
branches the other workloads do not reach (sofren2.c:5594-5599 edge clamp of
non-power-of-two textures, 6416-6423 unlit materials, 6440-6470 directional
lights, 5929-5931 the 16-light cut of a tile's list, 6067-6124 and
6732-6743 rendering without clearing over the caller's buffer at a padded
pitch, 8412-8417 dropping draws beyond 64 instances).  Each branch is shown
to have run by a frame counter.

Bar: visibility IDs and depth bit-exact; colour within 1 LSB on at least
99.9% of pixels; the row padding of the caller's buffer untouched.
"""
import numpy as np
import pytest

import sofren2_amd as S
from tests import parity
from tests.refrun import prefill_pattern, ref_dump

pytestmark = pytest.mark.gpu

W, H = S.SCENE_SIZE[S.SCENE_FEATURES]
PAD = 64           # bytes of row padding in the caller's buffer
PREFILL = 7


def render_features(frame, samples):
    r = S.SceneRenderer(S.SCENE_FEATURES, 0, W, H, capture=S.CAPTURE_VISIBILITY | S.CAPTURE_SS_CACHE | S.CAPTURE_CLIPS,
                        samples=samples)
    try:
        buf = prefill_pattern(H, W + PAD // 4, PREFILL)
        r.lib.scene_render(frame, buf.ctypes.data, W * 4 + PAD)
        vid, depth = r.visibility()
        ss, ssvis = r.ss_cache()
        clip = r.clips()
        st = r.stats()
    finally:
        r.close()
    return {"color": np.ascontiguousarray(buf[:, :W]), "color_buffer": buf, "vid": vid, "depth": depth, "ss": ss,
            "ssvis": ssvis, "clip": clip}, st


@pytest.mark.parametrize("samples", [1, 4])
@pytest.mark.parametrize("frame", [0, 37])
def test_features_parity(frame, samples):
    print("features frame", frame, "samples", samples, flush=True)
    gpu, st = render_features(frame, samples)
    ref = ref_dump(S.SCENE_FEATURES, 0, W, H, frame, samples=samples, pad=PAD, prefill=PREFILL)
    rep = parity.compare_frame(gpu, ref)
    print("features", frame, samples, rep, "npot", st.tex_npot, "nearest", st.tex_nearest, "bilinear",
          st.tex_bilinear, "unlit", st.unlit_shadings, "dirlit", st.dirlit_shadings, "light cut tiles",
          st.light_cut_tiles, "bg blend", st.bg_blend_pixels, flush=True)
    parity.assert_parity(rep)
    assert np.array_equal(gpu["color_buffer"][:, W:], prefill_pattern(H, W + PAD // 4, PREFILL)[:, W:])
    # every branch ran
    assert st.tex_npot > 0 and st.tex_nearest > 0 and st.tex_bilinear > 0
    assert st.unlit_shadings > 0
    assert st.dirlit_shadings > 0
    assert st.light_cut_tiles > 0
    assert st.instances_dropped == 0
    if samples > 1:
        assert st.bg_blend_pixels > 0


@pytest.mark.parametrize("samples", [1, 4])
def test_features_drop_instances_beyond_64(samples):
    """Frame 537 draws the scene of camera frame 37 plus six instances past
    the 64 the visibility-ID split holds: they are dropped, so the frame is
    the reference's frame 37."""
    gpu, st = render_features(537, samples)
    assert st.instances_dropped == 6
    ref = ref_dump(S.SCENE_FEATURES, 0, W, H, 37, samples=samples, pad=PAD, prefill=PREFILL)
    rep = parity.compare_frame(gpu, ref)
    print("features dropped", samples, rep, flush=True)
    parity.assert_parity(rep)

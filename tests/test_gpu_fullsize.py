"""Full-size checks against the reference renderer.

The whole 64-view mixed-resolution batch, rendered as one submission with
64-bit visibility written (as the batched throughput leg times it), with
ten of its views compared against the reference: the largest view, the
largest 4x MSAA view, every crowd-with-shadows view among the first ones and
views of each other scene kind.  And one 3840x2160 4x MSAA request served on
a template of 1.39 million triangles (box lattice L = 25 plus 64 skinned
humanoids under 6 shadow-casting lights: 65 instances, so the 7/17
visibility-ID split), compared against the reference's 7/17 build rendering
the template's pose with the request's camera.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import sofren2_amd as S
from sofren2_amd import workloads
from tests import parity
from tests.refrun import ref_dump

pytestmark = pytest.mark.gpu


def _picks(spec):
    order = sorted(range(len(spec)), key=lambda i: -spec[i][2] * spec[i][3])
    picks = [order[0], next(i for i in order if spec[i][4] == 4)]
    crowd = [i for i in range(len(spec)) if workloads.MIXED_SCENES[spec[i][0]][0] == S.SCENE_CROWD_SHADOWS]
    picks += crowd[:4]
    for si in range(len(workloads.MIXED_SCENES)):
        picks += [i for i in order if spec[i][0] == si and i not in picks][:1]
    picks += [i for i in order[1:] if i not in picks][:10 - len(set(picks))]
    out = []
    for i in picks:
        if i not in out:
            out.append(i)
    return out


def test_mixed_batch_of_64_views_matches_reference():
    lib = S.library()
    lib.sft_init(640, 360, 0)
    mb = None
    try:
        mb = workloads.MixedBatch(lib, vis64=True)
        assert len(mb.views) == 64
        print("mixed batch", mb.pixels() / 1e6, "Mpix", flush=True)
        mb.render_batched()
        st = S.FrameStats()
        lib.get_stats(st)
        assert st.occlusion_invalid == 0 and st.occlusion_overflow == 0
        picks = _picks(mb.spec)
        assert len(picks) >= 8
        kinds = {workloads.MIXED_SCENES[mb.spec[i][0]][0] for i in picks}
        assert kinds == {S.SCENE_PLANE_SPHERES, S.SCENE_BOX_ROOM, S.SCENE_CROWD_SHADOWS}
        assert any(mb.spec[i][4] == 4 for i in picks)

        def ref(i):
            si, frame, w, h, ns = mb.spec[i]
            kind, seed = workloads.MIXED_SCENES[si]
            return ref_dump(kind, seed, w, h, frame, samples=ns)

        with ThreadPoolExecutor(max_workers=8) as ex:
            refs = list(ex.map(ref, picks))
        for i, r in zip(picks, refs):
            v = mb.views[i]
            rep = parity.compare_frame({"color": v.color, "vid": v.vid(), "depth": v.depth()}, r)
            print("view", i, mb.spec[i], rep, flush=True)
            parity.assert_parity(rep)
    finally:
        if mb is not None:
            mb.close()
        lib.sft_release()


SERVED_SPEC = (S.SCENE_SERVED, 568, 30)     # L = 25 lattice, 64 humanoids, 6 lights


def test_served_4k_msaa4_request_on_a_million_triangle_template():
    lib = S.library()
    lib.sft_init(640, 360, 0)
    srv = None
    try:
        srv = workloads.Server(lib, [SERVED_SPEC], budget_bytes=16 << 30, max_views=4)
        tris = lib.scene_tris()
        print("template triangles", tris, "ID split", srv.id_splits, flush=True)
        assert tris >= 1_000_000 and srv.id_splits == [7]
        w, h, ns, cf = 3840, 2160, 4, 17
        color = np.zeros((h, w), dtype=np.uint32)
        vis = np.zeros(((w + 63) // 64) * 64 * ((h + 63) // 64) * 64 * ns, dtype=np.uint64)
        srv.wait(srv.submit(srv.request(0, cf, w, h, ns, color, vis)))
        print("served", srv.stats(), flush=True)
        kind, seed, pose = SERVED_SPEC
        r = ref_dump(kind, seed, w, h, workloads.reference_frame_code(pose, cf), samples=ns, id_bits=7)
        got = {"color": color, "vid": (vis & 0xFFFFFFFF).astype(np.uint32),
               "depth": (vis >> np.uint64(32)).astype(np.uint32).view(np.float32)}
        rep = parity.compare_frame(got, r)
        print("4k request vs reference", rep, flush=True)
        parity.assert_parity(rep)
        assert (got["vid"] != 0xFFFFFFFF).mean() > 0.5
    finally:
        if srv is not None:
            srv.close()
        lib.scene_destroy()
        lib.sft_release()

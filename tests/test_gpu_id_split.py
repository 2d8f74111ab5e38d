"""The 7/17 visibility-ID split (the reference built with SFT_INSTANCE_BITS=7,
SFT_MESHLET_BITS=17, sofren2.c:154-165): a served-scene variant with 64
skinned humanoids and the room (65 instances, capsule shadows) rendered with
sftx_set_id_split(7) against the reference's 7/17 build, per frame and served
in a mix with a default-split template; with the default split the 65th draw
is dropped as in the reference (sofren2.c:8412-8417)."""
import numpy as np
import pytest

import sofren2_amd as S
from sofren2_amd import workloads
from tests import parity
from tests.refrun import ref_dump

pytestmark = pytest.mark.gpu

SEED, POSE = 102, 30    # served variant: 64 humanoids, ~990k triangles, shadow lights


def _render(frame, samples, id_bits, w=640, h=360):
    r = S.SceneRenderer(S.SCENE_SERVED, SEED, w, h, capture=S.CAPTURE_VISIBILITY | S.CAPTURE_SS_CACHE
                        | S.CAPTURE_CLIPS | S.CAPTURE_SHADOWS, samples=samples, id_bits=id_bits)
    try:
        assert r.lib.scene_characters(0) == 64
        color = r.render(frame).copy()
        vid, depth = r.visibility()
        ss, ssvis = r.ss_cache()
        out = {"color": color, "vid": vid, "depth": depth, "ss": ss, "ssvis": ssvis, "clip": r.clips(),
               "proxy_vis": r.proxy_visibility()}
        return out, r.stats()
    finally:
        r.close()


@pytest.mark.parametrize("cam,samples", [(0, 1), (20, 4)])
def test_65_instances_with_the_7_17_split_match_the_reference(cam, samples):
    frame = workloads.reference_frame_code(POSE, cam)
    gpu, st = _render(frame, samples, 7)
    ref = ref_dump(S.SCENE_SERVED, SEED, 640, 360, frame, samples=samples, id_bits=7)
    rep = parity.compare_frame(gpu, ref)
    v = gpu["vid"]
    drawn = v[(v != 0xFFFFFFFF) & ((v & 0x80000000) == 0)]
    n65 = int(np.count_nonzero((drawn >> 24) == 64))
    print("7/17 split cam", cam, "samples", samples, rep, "dropped", st.instances_dropped, "pixels of instance 64",
          n65, "proxies", gpu["proxy_vis"].shape[0], flush=True)
    parity.assert_parity(rep)
    assert st.instances_dropped == 0
    assert n65 > 0
    assert gpu["proxy_vis"].shape[0] == 640


def test_default_split_drops_the_65th_instance():
    """Under the default split the 65th draw is dropped (sofren2.c:8412-8417).
    The reference itself cannot render this frame: with shadows on, its TLAS
    build walks the claimed slot count and dereferences the dropped slot's
    never-written instance (sofren2.c:3630-3634), a NULL model, and the
    checker build dies with SIGSEGV.  So the frame is checked against the
    7/17 render of the same scene: identical wherever neither shows the 65th
    humanoid nor can see its capsule shadows, i.e. no pixel carries its ID,
    and the drawn instances' IDs are the same under both splits."""
    frame = workloads.reference_frame_code(POSE, 0)
    gpu6, st6 = _render(frame, 1, 6)
    gpu7, st7 = _render(frame, 1, 7)
    v6, v7 = gpu6["vid"], gpu7["vid"]
    own6 = (v6 != 0xFFFFFFFF) & ((v6 & 0x80000000) == 0)
    own7 = (v7 != 0xFFFFFFFF) & ((v7 & 0x80000000) == 0)
    print("6/18 split dropped", st6.instances_dropped, "7/17 dropped", st7.instances_dropped,
          "proxies", gpu6["proxy_vis"].shape[0], gpu7["proxy_vis"].shape[0], flush=True)
    assert st6.instances_dropped == 1 and st7.instances_dropped == 0
    assert not np.any((v6[own6] >> 25) >= 64)
    assert gpu6["proxy_vis"].shape[0] == 630 and gpu7["proxy_vis"].shape[0] == 640
    # the same triangles win wherever the 7/17 frame shows one of the first 64 instances
    keep = own7 & ((v7 >> 24) < 64)
    inst7, rest7 = v7[keep] >> 24, v7[keep] & 0xFFFFFF
    assert np.array_equal(v6[keep], (inst7 << 25) | rest7)
    assert np.array_equal(gpu6["depth"][keep], gpu7["depth"][keep])


def test_served_mix_of_splits_matches_alone_and_reference():
    """Templates recorded under different splits keep them: a 65-instance
    template (7/17) and a box room (6/18) served in one mix equal each request
    alone and the reference build of its split."""
    specs = [(3, SEED, POSE), (1, 21, 5)]
    reqs = [(0, 0, 640, 360, 1), (1, 17, 480, 300, 2), (0, 20, 512, 288, 4), (1, 40, 320, 200, 1),
            (0, 60, 800, 450, 2)]
    lib = S.library()
    lib.sft_init(640, 360, 0)
    srv = None
    try:
        srv = workloads.Server(lib, specs, budget_bytes=16 << 30, max_views=8)
        assert srv.id_splits == [7, 6]
        bufs = []
        for _, _, w, h, ns in reqs:
            g = ((w + 63) // 64) * 64 * ((h + 63) // 64) * 64 * ns
            bufs.append((np.zeros((h, w), dtype=np.uint32), np.zeros(g, dtype=np.uint64)))
        ids = [srv.submit(srv.request(t, cf, w, h, ns, c, v)) for (t, cf, w, h, ns), (c, v) in zip(reqs, bufs)]
        for i in ids:
            srv.wait(i)
        assert srv.stats()["max_batch"] > 1
        for k, ((t, cf, w, h, ns), (c_mix, v_mix)) in enumerate(zip(reqs, bufs)):
            c = np.zeros_like(c_mix)
            v = np.zeros_like(v_mix)
            srv.wait(srv.submit(srv.request(t, cf, w, h, ns, c, v)))
            np.testing.assert_array_equal(c, c_mix)
            np.testing.assert_array_equal(v, v_mix)
            kind, seed, pose = specs[t]
            ref = ref_dump(kind, seed, w, h, workloads.reference_frame_code(pose, cf), samples=ns,
                           id_bits=srv.id_splits[t])
            got = {"color": c_mix, "vid": (v_mix & 0xFFFFFFFF).astype(np.uint32),
                   "depth": (v_mix >> 32).astype(np.uint32).view(np.float32)}
            rep = parity.compare_frame(got, ref)
            print("served request", k, "split", srv.id_splits[t], rep, flush=True)
            parity.assert_parity(rep)
    finally:
        if srv is not None:
            srv.close()
        lib.scene_destroy()
        lib.sft_release()

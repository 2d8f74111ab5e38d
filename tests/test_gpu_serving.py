"""Request server on the GPU: continuous batching gives each request the same
frame it gets alone, and that frame is the reference's.  Requests of a
template with capsule shadows after its first batch render from the
template's traced shadow records: the same frames."""
import numpy as np
import pytest

import sofren2_amd as S
from sofren2_amd import workloads
from tests import parity
from tests.refrun import ref_dump

pytestmark = pytest.mark.gpu

SPECS = [(1, 21, 5), (2, 22, 40), (0, 23, 0), (3, 12, 30)]          # (kind, seed, pose frame)
REQUESTS = [  # (template, camera frame, width, height, samples)
    (0, 17, 640, 360, 1), (1, 40, 480, 300, 2), (2, 0, 256, 256, 4), (0, 99, 1024, 576, 1),
    (1, 3, 320, 180, 1), (2, 0, 700, 400, 2), (1, 40, 512, 512, 4), (0, 60, 300, 200, 2),
    (3, 8, 800, 448, 1), (3, 51, 400, 240, 4),
]


def _buffers(w, h, ns):
    return np.zeros((h, w), dtype=np.uint32), np.zeros(((w + 63) // 64) * 64 * ((h + 63) // 64) * 64 * ns,
                                                       dtype=np.uint64)


def test_served_requests_same_alone_and_in_the_mix():
    lib = S.library()
    lib.sft_init(640, 360, 0)
    srv = None
    try:
        srv = workloads.Server(lib, SPECS, budget_bytes=8 << 30, max_views=8)
        # the mix: every request submitted at once, batched as they are admitted
        mix = [_buffers(w, h, ns) for _, _, w, h, ns in REQUESTS]
        qs = [srv.request(t, cf, w, h, ns, c, v) for (t, cf, w, h, ns), (c, v) in zip(REQUESTS, mix)]
        ids = [srv.submit(q) for q in qs]
        lat = [srv.wait(i) for i in ids]
        st = srv.stats()
        print("mix latencies ms", [round(x, 2) for x in lat], "stats", st, flush=True)
        assert st["completed"] == len(REQUESTS)
        # each request alone
        for k, ((t, cf, w, h, ns), (c_mix, v_mix)) in enumerate(zip(REQUESTS, mix)):
            c, v = _buffers(w, h, ns)
            q = srv.request(t, cf, w, h, ns, c, v)
            srv.wait(srv.submit(q))
            print("alone vs mix", k, flush=True)
            np.testing.assert_array_equal(c, c_mix)
            np.testing.assert_array_equal(v, v_mix)
        # the shadowed templates' requests alone reused the records the mix traced
        st = srv.stats()
        print("shadow cache views", st["shadow_cache_views"], flush=True)
        assert st["shadow_cache_views"] >= sum(1 for t, *_ in REQUESTS if t in (1, 3))
        # against the reference: the template's pose with the request's camera
        for k in (1, 2, 3, 8, 9):
            t, cf, w, h, ns = REQUESTS[k]
            kind, seed, pose = SPECS[t]
            ref = ref_dump(kind, seed, w, h, workloads.reference_frame_code(pose, cf), samples=ns)
            c, v = mix[k]
            got = {"color": c, "vid": (v & 0xFFFFFFFF).astype(np.uint32),
                   "depth": (v >> 32).astype(np.uint32).view(np.float32)}
            rep = parity.compare_frame(got, ref)
            print("served vs reference", k, rep, flush=True)
            parity.assert_parity(rep)
    finally:
        if srv is not None:
            srv.close()
        lib.scene_destroy()
        lib.sft_release()


def test_admission_under_a_small_budget():
    """A budget too small for two requests at once still serves them all,
    one batch at a time, and records the waits."""
    lib = S.library()
    lib.sft_init(640, 360, 0)
    srv = None
    try:
        srv = workloads.Server(lib, SPECS[:1], budget_bytes=1, max_views=8)
        outs = [_buffers(320, 200, 1) for _ in range(4)]
        ids = [srv.submit(srv.request(0, 10 * k, 320, 200, 1, c, v)) for k, (c, v) in enumerate(outs)]
        for i in ids:
            srv.wait(i)
        st = srv.stats()
        print("small budget stats", st, flush=True)
        assert st["completed"] == 4 and st["max_batch"] == 1
    finally:
        if srv is not None:
            srv.close()
        lib.scene_destroy()
        lib.sft_release()


def test_visibility_only_for_the_requests_that_ask():
    """In a mix where every other request asks for 64-bit visibility, those
    requests get the planes they get alone, the others get their colour only
    (their visibility is neither written nor downloaded), and every colour
    equals the request rendered alone."""
    lib = S.library()
    lib.sft_init(640, 360, 0)
    srv = None
    try:
        srv = workloads.Server(lib, SPECS, budget_bytes=8 << 30, max_views=8)
        mix = []
        for k, (_, _, w, h, ns) in enumerate(REQUESTS):
            c, v = _buffers(w, h, ns)
            mix.append((c, v if k % 2 == 0 else None))
        ids = [srv.submit(srv.request(t, cf, w, h, ns, c, v)) for (t, cf, w, h, ns), (c, v) in zip(REQUESTS, mix)]
        for i in ids:
            srv.wait(i)
        st = srv.stats()
        print("mix stats", st, flush=True)
        assert st["max_batch"] > 1
        for k, ((t, cf, w, h, ns), (c_mix, v_mix)) in enumerate(zip(REQUESTS, mix)):
            c, v = _buffers(w, h, ns)
            srv.wait(srv.submit(srv.request(t, cf, w, h, ns, c, v)))
            print("alone vs mix", k, "visibility" if v_mix is not None else "colour only", flush=True)
            np.testing.assert_array_equal(c, c_mix)
            if v_mix is not None:
                np.testing.assert_array_equal(v, v_mix)
    finally:
        if srv is not None:
            srv.close()
        lib.scene_destroy()
        lib.sft_release()

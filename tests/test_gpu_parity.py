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
    r = S.SceneRenderer(scene, seed, w, h, capture=S.CAPTURE_VISIBILITY | S.CAPTURE_SS_CACHE | S.CAPTURE_CLIPS
                        | S.CAPTURE_SHADOWS)
    try:
        color = r.render(frame).copy()
        vid, depth = r.visibility()
        ss, ssvis = r.ss_cache()
        clip = r.clips()
        proxy_vis = r.proxy_visibility()
        st = r.stats()
    finally:
        r.close()
    return {"color": color, "vid": vid, "depth": depth, "ss": ss, "ssvis": ssvis, "clip": clip,
            "proxy_vis": proxy_vis}, st


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


@pytest.mark.parametrize("frame,samples", [(0, 1), (40, 1), (40, 2)])
def test_map_file_parity(frame, samples):
    """The reference's own lightmapped map (res/simple.qsf): meshlets, BLAS and
    lightmaps built by the reference's offline tools, lit by the file's
    game-data lights."""
    from tests.mapfile import map_path
    map_path()
    w, h = S.SCENE_SIZE[S.SCENE_MAP]
    print("map frame", frame, "samples", samples, flush=True)
    if samples > 1:
        gpu, st = render_gpu_ms(S.SCENE_MAP, 0, w, h, frame, samples)
    else:
        gpu, st = render_gpu(S.SCENE_MAP, 0, w, h, frame)
    ref = ref_dump(S.SCENE_MAP, 0, w, h, frame, samples=samples)
    rep = parity.compare_frame(gpu, ref)
    print("map frame", frame, samples, rep, "clips", st.clip_triangles, "ms", st.ms_total, flush=True)
    parity.assert_parity(rep)


@pytest.mark.parametrize("frame", [0, 50])
def test_crowd_shadows_parity(frame):
    """Skinned crowd in the box room with 16 shadow lights: skinned screen-space
    cache bit-exact, per-proxy light visibility from the occlusion packets
    exact, and the capsule-shadowed colour."""
    w, h = S.SCENE_SIZE[S.SCENE_CROWD_SHADOWS]
    print("crowd_shadows frame", frame, flush=True)
    gpu, st = render_gpu(S.SCENE_CROWD_SHADOWS, 0, w, h, frame)
    ref = ref_dump(S.SCENE_CROWD_SHADOWS, 0, w, h, frame)
    assert st.occlusion_invalid == 0 and st.occlusion_overflow == 0
    assert st.occlusion_rays == 16 * 320 * 3
    np.testing.assert_array_equal(gpu["proxy_vis"], ref["proxy_vis"])
    rep = parity.compare_frame(gpu, ref)
    print("crowd_shadows frame", frame, rep, "visible", st.meshlets_visible, "skinned", st.meshlets_skinned,
          "occ nodes", st.occlusion_nodes, "tri tests", st.occlusion_tri_tests, "ms", st.ms_total,
          "shadows ms", st.ms_shadows, flush=True)
    parity.assert_parity(rep)


def render_gpu_ms(scene, seed, w, h, frame, samples):
    r = S.SceneRenderer(scene, seed, w, h, capture=S.CAPTURE_VISIBILITY | S.CAPTURE_SS_CACHE | S.CAPTURE_CLIPS,
                        samples=samples)
    try:
        color = r.render(frame).copy()
        vid, depth = r.visibility()
        ss, ssvis = r.ss_cache()
        clip = r.clips()
        st = r.stats()
    finally:
        r.close()
    return {"color": color, "vid": vid, "depth": depth, "ss": ss, "ssvis": ssvis, "clip": clip}, st


@pytest.mark.parametrize("samples", [2, 4])
@pytest.mark.parametrize("scene,frame", [(S.SCENE_PLANE_SPHERES, 0), (S.SCENE_BOX_ROOM, 37),
                                         (S.SCENE_CROWD_SHADOWS, 0)])
def test_msaa_parity(scene, frame, samples):
    """Rotated-grid MSAA: per-sample visibility IDs and depth bit-exact, the
    sample-weighted resolve within the colour bar."""
    w, h = S.SCENE_SIZE[scene]
    print("msaa scene", scene, "frame", frame, "samples", samples, flush=True)
    gpu, st = render_gpu_ms(scene, 0, w, h, frame, samples)
    ref = ref_dump(scene, 0, w, h, frame, samples=samples)
    assert gpu["vid"].size == ref["vid"].size
    rep = parity.compare_frame(gpu, ref)
    print("msaa", scene, frame, samples, rep, "groups", st.shading_groups, "ms", st.ms_total, flush=True)
    parity.assert_parity(rep)


BATCH_VIEWS = [  # (scene, seed, width, height, frame, samples)
    (S.SCENE_BOX_ROOM, 1, 640, 400, 5, 1),
    (S.SCENE_PLANE_SPHERES, 2, 320, 256, 0, 4),
    (S.SCENE_CROWD_SHADOWS, 3, 800, 450, 10, 2),
    (S.SCENE_BOX_ROOM, 1, 1024, 512, 77, 1),
    (S.SCENE_PLANE_SPHERES, 2, 300, 300, 0, 1),
    (S.SCENE_CROWD_SHADOWS, 3, 512, 288, 64, 4),
]


def test_batched_views_match_reference():
    """Views of several scenes at mixed resolutions and MSAA rendered in one
    submission: each is the reference's frame for that scene, size and frame
    (visibility IDs and depth bit-exact, colour within the bar)."""
    lib = S.library()
    lib.sft_init(640, 360, 0)
    try:
        handles = {}
        for scene, seed, w, h, frame, samples in BATCH_VIEWS:
            if (scene, seed) not in handles:
                handles[(scene, seed)] = lib.scene_new(scene, seed, w, h)
                assert handles[(scene, seed)] >= 0
        views = [S.View(handles[(sc, sd)], fr, w, h, ns) for sc, sd, w, h, fr, ns in BATCH_VIEWS]
        print("batch of", len(views), flush=True)
        S.render_batch(lib, views)
        st = S.FrameStats()
        lib.get_stats(st)
        print("batch stats visible", st.meshlets_visible, "tiles", st.tiles, "ms", st.ms_total, flush=True)
        assert st.occlusion_invalid == 0 and st.occlusion_overflow == 0
        for (scene, seed, w, h, frame, samples), v in zip(BATCH_VIEWS, views):
            ref = ref_dump(scene, seed, w, h, frame, samples=samples)
            got = {"color": v.color, "vid": v.vid(), "depth": v.depth()}
            rep = parity.compare_frame(got, ref)
            print("view", scene, seed, w, h, frame, samples, rep, flush=True)
            parity.assert_parity(rep)
    finally:
        lib.scene_destroy()
        lib.sft_release()


def test_tuned_settings_change_no_output_bit():
    """The autotuner's candidates only re-cut passes into work: the tuned
    batch, and batches under extreme settings (one payload per work item,
    every payload split into row bands or none, half an occupancy-sized grid,
    the smallest occlusion stack), give the default settings' colour,
    visibility and shadow-proxy visibility bit for bit, in every resolution
    class."""
    lib = S.library()
    lib.sft_init(640, 360, 0)
    try:
        handles = {}
        spec = BATCH_VIEWS + [(S.SCENE_BOX_ROOM, 1, 1280, 832, 40, 1), (S.SCENE_CROWD_SHADOWS, 3, 1408, 768, 3, 2)]
        for scene, seed, w, h, frame, samples in spec:
            if (scene, seed) not in handles:
                handles[(scene, seed)] = lib.scene_new(scene, seed, w, h)
                assert handles[(scene, seed)] >= 0

        def views():
            return [S.View(handles[(sc, sd)], fr, w, h, ns) for sc, sd, w, h, fr, ns in spec]

        base = views()
        S.render_batch(lib, base)
        default = S.Tuning()
        lib.get_tuning(default)
        assert list(default.raster_chunk) == [128, 128, 128] and default.tile_px == 64
        assert list(default.raster_split) == [2048, 2048, 2048] and default.occlusion_stack == 1024
        tuned = views()
        t = S.tune_batch(lib, tuned, repeats=2)
        print("tuned", t.as_dict(), flush=True)
        assert all(c in (64, 96, 128, 256, 384, 512) for c in t.raster_chunk)
        assert all(c in (1024, 2048, 4096, 8192, 0xFFFFFFFF) for c in t.raster_split)
        assert t.occlusion_stack in (256, 512, 1024) and t.ms_occlusion > 0
        assert all(ms > 0 for ms in t.ms_class), "every class has views here"
        g = t.raster_groups_resident
        assert t.raster_groups in (g // 2, g, 2 * g)
        got = S.Tuning()
        lib.get_tuning(got)
        assert list(got.raster_chunk) == list(t.raster_chunk) and got.raster_groups == t.raster_groups
        assert list(got.raster_split) == list(t.raster_split) and got.occlusion_stack == t.occlusion_stack
        for v, b in zip(tuned, base):
            np.testing.assert_array_equal(v.color, b.color)
            np.testing.assert_array_equal(v.vis, b.vis)
        # the last two: items of up to 1024 payloads, and every triangle split
        # into 8 row parts, more than a staged chunk's unit list holds (each
        # triangle then swept as one unit)
        for chunks, split, stack in (([1, 16, 1], [1, 1, 1], 256), ([128, 1, 32], [0xFFFFFFFF] * 3, 512),
                                     ([1024, 300, 129], [2048] * 3, 1024), ([128, 512, 1024], [1, 1, 1], 1024)):
            extreme = S.Tuning()
            extreme.raster_chunk[:] = chunks
            extreme.raster_split[:] = split
            extreme.raster_groups = max(1, g // 2)
            extreme.occlusion_stack = stack
            lib.set_tuning(extreme)
            r0 = lib.overflow_reruns()
            again = views()
            S.render_batch(lib, again)
            st = S.FrameStats()
            lib.get_stats(st)
            print("extreme settings", chunks, split, stack, ": raster ms", st.ms_raster, "occlusion ms", st.ms_shadows,
                  "re-runs", lib.overflow_reruns() - r0, flush=True)
            assert st.occlusion_overflow == 0
            for v, b in zip(again, base):
                np.testing.assert_array_equal(v.color, b.color)
                np.testing.assert_array_equal(v.vis, b.vis)
        lib.set_tuning(default)
    finally:
        lib.scene_destroy()
        lib.sft_release()


def test_pipelined_batches_match_blocking_batches():
    """Batches submitted without waiting (three in flight, a fourth forcing the
    oldest to complete, transient dmesh rows in every frame set) give the same
    colour and visibility as blocking submissions of the same views."""
    lib = S.library()
    lib.sft_init(640, 360, 0)
    try:
        handles = {}
        for scene, seed, w, h, frame, samples in BATCH_VIEWS:
            if (scene, seed) not in handles:
                handles[(scene, seed)] = lib.scene_new(scene, seed, w, h)
                assert handles[(scene, seed)] >= 0

        def views_for(shift):
            return [S.View(handles[(sc, sd)], (fr + shift) % 120, w + 16 * shift, h, ns)
                    for sc, sd, w, h, fr, ns in BATCH_VIEWS]

        blocking = {}
        for shift in (0, 1):
            vs = views_for(shift)
            S.render_batch(lib, vs)
            blocking[shift] = [(v.color.copy(), v.vis.copy()) for v in vs]
        order = [0, 1, 0, 1]
        sets = [views_for(k) for k in order]
        tickets = [S.submit_batch_async(lib, vs) for vs in sets]
        print("tickets", tickets, flush=True)
        assert lib.batch_wait(tickets[0]) == -1  # completed by the fourth submission
        assert lib.batch_wait(tickets[2]) == 0
        assert lib.batch_wait(tickets[3]) == 0
        assert lib.batch_wait(tickets[1]) == -1  # completed by the wait for a later ticket
        for k, vs in zip(order, sets):
            for (c, vis), v in zip(blocking[k], vs):
                np.testing.assert_array_equal(v.color, c)
                np.testing.assert_array_equal(v.vis, vis)
    finally:
        lib.scene_destroy()
        lib.sft_release()


@pytest.mark.parametrize("scene", [S.SCENE_BOX_ROOM, S.SCENE_CROWD_SHADOWS])
def test_graph_replays_match_eager_launches(scene):
    """Frames replayed from the captured HIP graph equal frames launched
    eagerly, replay after replay, with eager frames in between."""
    w, h = S.SCENE_SIZE[scene]
    r = S.SceneRenderer(scene, 0, w, h)
    try:
        frames = [0, 7, 19, 44, 90]
        r.lib.set_graphs(0)
        eager = [r.render(f).copy() for f in frames]
        r.lib.set_graphs(1)
        c0 = r.lib.graph_captures()
        for k, f in enumerate(frames):
            got = r.render(f)
            print("graph replay", scene, f, flush=True)
            np.testing.assert_array_equal(got, eager[k])
            r.lib.set_graphs(0)
            np.testing.assert_array_equal(r.render(frames[-1 - k]), eager[-1 - k])
            r.lib.set_graphs(1)
        captures = r.lib.graph_captures() - c0
        print("captures", captures, flush=True)
        assert 1 <= captures <= 2  # the second frame of a shape captures it
    finally:
        r.close()


@pytest.mark.parametrize("scene", [S.SCENE_PLANE_SPHERES, S.SCENE_CROWD_SHADOWS])
def test_device_output_frames_in_flight(scene):
    """With device output sft_present returns before the frame completes (two
    in flight, dmesh rows in both frame sets); the last frame in HBM equals
    the same frame rendered with host output."""
    import ctypes
    w, h = S.SCENE_SIZE[scene]
    r = S.SceneRenderer(scene, 0, w, h)
    try:
        frames = [3, 17, 29, 41, 53, 60]
        want = r.render(frames[-1]).copy()
        r.lib.set_output(1)
        for f in frames:
            r.render(f)
        ptr = r.lib.device_color()
        got = np.zeros((h, w), dtype=np.uint32)
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        assert hip.hipMemcpy(got.ctypes.data, ptr, got.nbytes, 2) == 0  # device to host
        print("device output frame", frames[-1], flush=True)
        np.testing.assert_array_equal(got, want)
        r.lib.set_output(0)
    finally:
        r.close()

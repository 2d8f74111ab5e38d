"""Seeded view sets for the batched and served workloads.

The mixed-resolution batch: 64 camera views over 4 scene instances (the
textured interior, the crowd with shadows and the plane with spheres, scene
seeds 1-4), resolution log-uniform between 256x256 and 1920x1080 pixels with
aspect U[1, 2], and 2x or 4x MSAA with probability 0.25 each.  Sizes are whole
pixels; each view takes a frame of its scene's camera path.
"""
import math

import numpy as np

# scene kind per seed 1..4: interior, crowd with shadows, interior, plane + spheres
MIXED_SCENES = [(1, 1), (2, 2), (1, 3), (0, 4)]


def mixed_views(n_views=64, seed=0, min_px=256 * 256, max_px=1920 * 1080):
    """Returns [(scene_index, frame, width, height, samples)] for the mixed batch."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_views):
        px = math.exp(rng.uniform(math.log(min_px), math.log(max_px)))
        aspect = rng.uniform(1.0, 2.0)
        h = max(16, int(round(math.sqrt(px / aspect))))
        w = max(16, int(round(h * aspect)))
        r = rng.uniform()
        samples = 2 if r < 0.25 else (4 if r < 0.5 else 1)
        out.append((i % len(MIXED_SCENES), int(rng.integers(0, 120)), w, h, samples))
    return out


class MixedBatch:
    """The mixed-resolution view set over resident scene instances, rendered
    either as one batch per step or view by view through sft_present."""

    def __init__(self, lib, views=None, vis64=False):
        import sofren2_amd as S
        self.S = S
        self.lib = lib
        self.vis64 = vis64
        self.spec = views if views is not None else mixed_views()
        self.handles = []
        for kind, seed in MIXED_SCENES:
            h = lib.scene_new(kind, seed, 640, 360)
            if h < 0:
                raise RuntimeError("scene creation failed")
            self.handles.append(h)
        self.views = [S.View(self.handles[si], fr, w, h, ns, vis64=vis64) for si, fr, w, h, ns in self.spec]

    def pixels(self):
        return sum(v.width * v.height for v in self.views)

    def render_batched(self):
        self.S.render_batch(self.lib, self.views)

    def tune(self, repeats=3):
        """The autotuner's settings for this view set (kept by the renderer)."""
        return self.S.tune_batch(self.lib, self.views, repeats)

    def render_pipelined(self, steps):
        """`steps` batches with three in flight: batches k + 1 and k + 2 are
        recorded and launched while batch k runs; outputs cycle through three
        buffer sets.  Returns each batch's latency (record start to
        completion, s)."""
        import time
        S = self.S
        if not hasattr(self, "views_b"):
            self.views_b = [S.View(v.scene, v.frame, v.width, v.height, v.samples, vis64=self.vis64) for v in self.views]
            self.views_c = [S.View(v.scene, v.frame, v.width, v.height, v.samples, vis64=self.vis64) for v in self.views]
        sets = [self.views, self.views_b, self.views_c]
        pending = []
        lat = []
        for k in range(steps):
            t0 = time.perf_counter()
            ticket = S.submit_batch_async(self.lib, sets[k % 3])
            pending.append((ticket, t0))
            if len(pending) == 3:
                t, ts = pending.pop(0)
                self.lib.batch_wait(t)
                lat.append(time.perf_counter() - ts)
        for t, ts in pending:
            self.lib.batch_wait(t)
            lat.append(time.perf_counter() - ts)
        return lat

    def render_per_frame(self):
        """Every view through sft_present; with vis64 views each frame also
        delivers its visibility IDs and depth (the capture of the last
        present: the same 64 bits per sample)."""
        lib = self.lib
        if self.vis64:
            lib.set_capture(self.S.CAPTURE_VISIBILITY)
        for v in self.views:
            lib.sft_resize(v.width, v.height)
            lib.set_msaa(v.samples)
            lib.sft_frame_begin()
            lib.set_color_buffer(v.color.ctypes.data, v.width * 4)
            lib.scene_draw(v.scene, v.frame, v.width, v.height)
            lib.sft_present()
            if self.vis64:
                n = v.vis.size
                if not hasattr(v, "_vid"):
                    v._vid = np.zeros(n, np.uint32)
                    v._depth = np.zeros(n, np.float32)
                lib.cap_visibility(v._vid.ctypes.data, v._depth.ctypes.data)
            lib.sft_frame_end()
        lib.set_capture(0)
        lib.set_msaa(1)

    def close(self):
        for h in self.handles:
            self.lib.scene_free(h)
        self.handles = []


# ---------------------------------------------------------------- serving
def served_scene_specs(n_scenes=32):
    """Resident scenes of the served workload: (kind, seed, pose_frame) of
    served-scene variants (50k-1M triangles, 0-64 skinned humanoids, 0-16
    shadow-casting lights, drawn from the seed).  The last one is seed 102,
    a 990k-triangle variant with 64 humanoids (65 instances: the 7/17
    visibility-ID split), so the workload reaches the top of both ranges."""
    seeds = [10 + i for i in range(n_scenes - 1)] + [102] if n_scenes > 1 else [10]
    return [(3, seed, (7 * i) % 120) for i, seed in enumerate(seeds)]


def served_requests(n, rate_hz, n_scenes, seed=1, min_px=128 * 128, max_px=3840 * 2160):
    """Poisson arrivals: [(t_arrival_s, template_index, camera_frame, width, height, samples)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate_hz)
        px = math.exp(rng.uniform(math.log(min_px), math.log(max_px)))
        aspect = rng.uniform(1.0, 2.0)
        h = max(16, int(round(math.sqrt(px / aspect))))
        w = max(16, int(round(h * aspect)))
        r = rng.uniform()
        samples = 1 if r < 1 / 3 else (2 if r < 2 / 3 else 4)
        out.append((t, int(rng.integers(0, n_scenes)), int(rng.integers(0, 120)), w, h, samples))
    return out


class Server:
    """Resident scene templates behind the request server."""

    def __init__(self, lib, specs, budget_bytes=200 << 30, max_views=64):
        import ctypes
        import sofren2_amd as S
        self.S, self.lib, self.ct = S, lib, ctypes
        self.specs = specs
        self.handles, self.templates, self.id_splits = [], [], []
        for kind, seed, pose in specs:
            h = lib.scene_new(kind, seed, 640, 360)
            if h < 0:
                raise RuntimeError("scene creation failed")
            self.handles.append(h)
            # the room and 64 humanoids are 65 instances: the 7/17 ID split
            split = id_split_for(lib.scene_characters(h) + 1)
            self.id_splits.append(split)
            lib.set_id_split(split)
            lib.sft_frame_begin()
            lib.scene_draw(h, pose, 640, 360)
            self.templates.append(lib.server_template())
            lib.sft_frame_end()
            lib.set_id_split(6)
        if lib.server_start(budget_bytes, max_views) != 0:
            raise RuntimeError("server start failed")

    def request(self, template_index, camera_frame, width, height, samples, color=None, vis=None):
        ct = self.ct
        kind, seed, pose = self.specs[template_index]
        q = self.S.Request()
        q.scene = self.templates[template_index]
        q.width, q.height, q.samples = width, height, samples
        pos, view, proj = (ct.c_float * 3)(), (ct.c_float * 16)(), (ct.c_float * 16)()
        self.lib.scene_camera(self.handles[template_index], camera_frame, width, height, pos, view, proj)
        q.camera_pos, q.view, q.proj = pos, view, proj
        q.color = color.ctypes.data if color is not None else None
        q.vis64 = vis.ctypes.data if vis is not None else None
        return q

    def submit(self, q):
        rid = self.lib.server_submit(self.ct.byref(q))
        if rid == 0:
            raise RuntimeError("request rejected")
        return rid

    def wait(self, rid):
        lat = self.ct.c_double()
        if self.lib.server_wait(rid, self.ct.byref(lat)) != 0:
            raise RuntimeError("unknown request")
        return lat.value

    def present_alone(self, template_index, camera_frame, width, height, samples, keep=False):
        """The request rendered through sft_present (the per-frame path, the
        server stopped): the template's scene at its pose with the request's
        camera, colour delivered to host memory.  With keep, returns colour
        and the 64-bit visibility planes ((1/w bits << 32) | ID)."""
        ct, lib, S = self.ct, self.lib, self.S
        kind, seed, pose = self.specs[template_index]
        lib.set_output(0)
        lib.sft_resize(width, height)
        lib.set_msaa(samples)
        lib.set_id_split(self.id_splits[template_index])
        lib.set_capture(S.CAPTURE_VISIBILITY if keep else 0)
        color = np.zeros((height, width), dtype=np.uint32)
        lib.sft_frame_begin()
        lib.set_color_buffer(color.ctypes.data, width * 4)
        lib.scene_draw(self.handles[template_index], reference_frame_code(pose, camera_frame), width, height)
        lib.sft_present()
        lib.sft_frame_end()
        out = None
        if keep:
            cols, rows, ns = ct.c_int32(), ct.c_int32(), ct.c_int32()
            lib.cap_grid(ct.byref(cols), ct.byref(rows), ct.byref(ns))
            n = cols.value * 64 * rows.value * 64 * ns.value
            vid = np.zeros(n, dtype=np.uint32)
            depth = np.zeros(n, dtype=np.float32)
            lib.cap_visibility(vid.ctypes.data, depth.ctypes.data)
            out = (color, (depth.view(np.uint32).astype(np.uint64) << np.uint64(32)) | vid.astype(np.uint64))
            lib.set_capture(0)
        lib.set_msaa(1)
        lib.set_id_split(6)
        return out

    def stats(self):
        st = self.S.ServerStats()
        self.lib.server_get_stats(self.ct.byref(st))
        return {n: (getattr(st, n) if t is self.ct.c_double else int(getattr(st, n)))
                for n, t in self.S.ServerStats._fields_}

    def close(self):
        self.lib.server_release()
        for h in self.handles:
            self.lib.scene_free(h)


def id_split_for(instances):
    """Instance bits of the visibility-ID split a frame of `instances` draws
    needs: the default 6 up to 64 instances, else the reference's 7/17 build
    (up to 128)."""
    return 6 if instances <= 64 else 7


def reference_frame_code(pose_frame, camera_frame):
    """Frame number that makes a scene take the template's pose and the
    request's camera (what the reference checker is given)."""
    return 1000 * (pose_frame + 1) + camera_frame


def run_served(server, requests, keep=None):
    """Replays Poisson arrivals against the server from a submitting thread,
    each request delivering its frame into a host buffer of its own; returns
    per-request latencies (ms), wall time and, for the indices in `keep`, the
    colour and visibility outputs."""
    import threading
    import time
    kept = {}
    ids = [None] * len(requests)
    bufs = [None] * len(requests)
    keep = set(keep or ())

    def submitter():
        t0 = time.perf_counter()
        for i, (ta, si, cf, w, h, ns) in enumerate(requests):
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            color = np.empty((h, w), dtype=np.uint32)
            vis = None
            if i in keep:
                vis = np.zeros(((w + 63) // 64) * 64 * ((h + 63) // 64) * 64 * ns, dtype=np.uint64)
                kept[i] = (color, vis)
            q = server.request(si, cf, w, h, ns, color, vis)
            bufs[i] = (q, color, vis)
            ids[i] = server.submit(q)

    t_start = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    lat = []
    for i in range(len(requests)):
        while ids[i] is None:
            time.sleep(0.0002)
        lat.append(server.wait(ids[i]))
        bufs[i] = None
    th.join()
    wall = time.perf_counter() - t_start
    return np.array(lat), wall, kept

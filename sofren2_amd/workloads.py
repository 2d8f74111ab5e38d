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

    def __init__(self, lib, views=None):
        import sofren2_amd as S
        self.S = S
        self.lib = lib
        self.spec = views if views is not None else mixed_views()
        self.handles = []
        for kind, seed in MIXED_SCENES:
            h = lib.scene_new(kind, seed, 640, 360)
            if h < 0:
                raise RuntimeError("scene creation failed")
            self.handles.append(h)
        self.views = [S.View(self.handles[si], fr, w, h, ns, vis64=False) for si, fr, w, h, ns in self.spec]

    def pixels(self):
        return sum(v.width * v.height for v in self.views)

    def render_batched(self):
        self.S.render_batch(self.lib, self.views)

    def render_per_frame(self):
        lib = self.lib
        for v in self.views:
            lib.sft_resize(v.width, v.height)
            lib.set_msaa(v.samples)
            lib.sft_frame_begin()
            lib.set_color_buffer(v.color.ctypes.data, v.width * 4)
            lib.scene_draw(v.scene, v.frame, v.width, v.height)
            lib.sft_present()
            lib.sft_frame_end()
        lib.set_msaa(1)

    def close(self):
        for h in self.handles:
            self.lib.scene_free(h)
        self.handles = []

"""Writes the golden fixtures of tests/test_oracle_golden.py.

Each case is rendered by the reference renderer compiled from its own source
(oracle/_ref, one worker thread) and stored as data: SHA-256 digests of the
full colour, visibility-ID and depth planes, a 32x32 crop of each at the
frame centre, and the cull and clip counts.  The rcp/rsqrt estimates behind
these values are the build host's, so the CPU model is recorded too.

    python3 tests/golden/make_golden.py        (needs make -C oracle ref)
"""
import hashlib
import json
import os
import platform
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

# (name, scene, seed, width, height, frame, samples): scene 0 plane + spheres,
# 1 box room, 2 crowd with capsule shadows
CASES = [
    ("plane_spheres_s0", 0, 0, 640, 360, 0, 1),
    ("plane_spheres_s1_aa4", 0, 1, 320, 180, 0, 4),
    ("box_room_f5", 1, 0, 320, 192, 5, 1),
    ("box_room_f37_aa2", 1, 0, 256, 160, 37, 2),
    ("crowd_f10", 2, 0, 320, 192, 10, 1),
]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def centre_crop(plane, width, height, n=32):
    y0, x0 = max(0, height // 2 - n // 2), max(0, width // 2 - n // 2)
    return plane[y0:y0 + n, x0:x0 + n].copy()


def frame_summary(d, width, height, samples):
    """Digests and crops of a dump dict (color [h][w]; vid/depth over the tile
    grid, per sample)."""
    cols, rows = (width + 63) // 64, (height + 63) // 64
    vid0 = d["vid"].reshape(samples, rows * 64, cols * 64)[0]
    depth0 = d["depth"].reshape(samples, rows * 64, cols * 64)[0]
    return {
        "color_sha256": digest(d["color"]), "vid_sha256": digest(d["vid"]), "depth_sha256": digest(d["depth"]),
        "color_crop": centre_crop(d["color"], width, height),
        "vid_crop": centre_crop(vid0, width, height),
        "depth_crop": centre_crop(depth0, width, height),
    }


def main():
    from tests.refrun import ref_dump
    index = {"cpu_model": cpu_model(), "cases": []}
    for name, scene, seed, w, h, frame, samples in CASES:
        d = ref_dump(scene, seed, w, h, frame, samples=samples)
        s = frame_summary(d, w, h, samples)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), color_crop=s["color_crop"], vid_crop=s["vid_crop"],
                            depth_crop=s["depth_crop"])
        index["cases"].append({"name": name, "scene": scene, "seed": seed, "width": w, "height": h, "frame": frame,
                               "samples": samples, "color_sha256": s["color_sha256"], "vid_sha256": s["vid_sha256"],
                               "depth_sha256": s["depth_sha256"], "visible_meshlets": int(d["ssvis"].sum()),
                               "clip_records": int(d["clip"].shape[0])})
        print(name, s["color_sha256"][:16], flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(index, f, indent=1)


if __name__ == "__main__":
    main()

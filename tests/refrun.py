"""Runs the reference checker build (oracle/_ref) and loads what it dumps.

The checker is the reference renderer compiled from its own source with
capture hooks (oracle/ref_harness.c).  Parity runs use one worker thread: with
more, the reference's tile payload order and clip-slot numbering depend on
thread scheduling.
"""
import functools
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = {1: "sofren2_ref", 2: "sofren2_ref_aa2", 4: "sofren2_ref_aa4"}


def ref_path(samples=1):
    return os.path.join(ROOT, "oracle", "_ref", REF_BIN[samples])


def have_ref(samples=1):
    return os.path.exists(ref_path(samples))


@functools.lru_cache(maxsize=16)
def ref_dump(scene, seed, width, height, frame=0, samples=1, threads=1):
    exe = ref_path(samples)
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} missing: build it with make -C oracle ref")
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([exe, "dump", str(scene), str(seed), str(width), str(height), str(threads), str(frame), d],
                       check=True, timeout=600)
        meta = {}
        for line in open(os.path.join(d, "meta.txt")):
            k, v = line.split()
            meta[k] = int(v)
        out = {"meta": meta}
        out["color"] = np.fromfile(os.path.join(d, "color.bin"), dtype=np.uint32).reshape(height, width)
        out["vid"] = np.fromfile(os.path.join(d, "vid.bin"), dtype=np.uint32)
        out["depth"] = np.fromfile(os.path.join(d, "depth.bin"), dtype=np.float32)
        out["ss"] = np.fromfile(os.path.join(d, "ss.bin"), dtype=np.uint32).reshape(-1, 9, 64)
        out["ssvis"] = np.fromfile(os.path.join(d, "ssvis.bin"), dtype=np.uint8)
        out["clip"] = np.fromfile(os.path.join(d, "clip.bin"), dtype=np.uint32).reshape(-1, 41)
        # world shadow proxies: posa, posb (3 f32 each), radius, light_vis[64] (92 bytes)
        px = np.fromfile(os.path.join(d, "proxies.bin"), dtype=np.uint8).reshape(-1, 92)
        out["proxy_vis"] = px[:, 28:92].copy()
        out["proxy_geom"] = px[:, :28].copy().view(np.float32)
        return out


def ref_bench(scene, seed, width, height, threads=0, warmup=10, frames=120, samples=1):
    exe = ref_path(samples)
    proc = subprocess.run([exe, "bench", str(scene), str(seed), str(width), str(height), str(threads),
                           str(warmup), str(frames)], check=True, stdout=subprocess.PIPE, text=True, timeout=1200)
    import json
    return json.loads(proc.stdout.strip().splitlines()[-1])

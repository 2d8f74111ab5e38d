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


def ref_path(samples=1, id_bits=6):
    """The checker build for `samples` per pixel and the visibility-ID split
    with `id_bits` instance bits (6: the default; 7: SFT_INSTANCE_BITS=7)."""
    suffix = {6: "", 7: "_i7"}[id_bits]
    return os.path.join(ROOT, "oracle", "_ref", REF_BIN[samples] + suffix)


def have_ref(samples=1, id_bits=6):
    return os.path.exists(ref_path(samples, id_bits))


def prefill_pattern(rows, cols, seed):
    """The colour buffer a pre-filled dump starts from (ref_harness.c
    sfth_prefill): a hash of (x, y, seed) per 32-bit word."""
    y, x = np.meshgrid(np.arange(rows, dtype=np.uint32), np.arange(cols, dtype=np.uint32), indexing="ij")
    with np.errstate(over="ignore"):
        v = (x * np.uint32(73856093)) ^ (y * np.uint32(19349663)) ^ np.uint32((seed * 83492791) & 0xFFFFFFFF)
        return (v * np.uint32(2654435761)).astype(np.uint32)


@functools.lru_cache(maxsize=16)
def ref_dump(scene, seed, width, height, frame=0, samples=1, threads=1, pad=0, prefill=0, id_bits=6):
    """The reference's render of a scene frame.  `pad` bytes of row padding
    and a non-zero `prefill` seed make the colour buffer start as
    prefill_pattern over (width + pad/4) words per row; out["color_buffer"]
    is then the whole padded buffer after the render.  `id_bits` selects the
    build with that many instance bits in its visibility IDs."""
    exe = ref_path(samples, id_bits)
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} missing: build it with make -C oracle ref")
    with tempfile.TemporaryDirectory() as d:
        args = [exe, "dump", str(scene), str(seed), str(width), str(height), str(threads), str(frame), d]
        if pad or prefill:
            args += [str(pad), str(prefill)]
        subprocess.run(args, check=True, timeout=600)
        meta = {}
        for line in open(os.path.join(d, "meta.txt")):
            k, v = line.split()
            meta[k] = int(v)
        out = {"meta": meta}
        buf = np.fromfile(os.path.join(d, "color.bin"), dtype=np.uint32).reshape(height, width + pad // 4)
        out["color_buffer"] = buf
        out["color"] = np.ascontiguousarray(buf[:, :width])
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


def ref_pose(seed, times):
    """sftx_pose_check of the reference build: (n, 2, 12, 16) float32."""
    times = np.ascontiguousarray(times, dtype=np.float32)
    with tempfile.TemporaryDirectory() as d:
        tp, op = os.path.join(d, "t.bin"), os.path.join(d, "o.bin")
        times.tofile(tp)
        subprocess.run([ref_path(), "pose", str(seed), tp, op], check=True, timeout=120)
        return np.fromfile(op, dtype=np.float32).reshape(-1, 2, 12, 16)


def ref_rays(scene, seed, n):
    """sftx_raycast_check of the reference build: (n, 8) float32."""
    with tempfile.TemporaryDirectory() as d:
        op = os.path.join(d, "o.bin")
        subprocess.run([ref_path(), "ray", str(scene), str(seed), str(n), op], check=True, timeout=300)
        return np.fromfile(op, dtype=np.float32).reshape(-1, 8)


def ref_build(seed, kind):
    """sftx_builder_check of the reference build: the raw output bytes."""
    with tempfile.TemporaryDirectory() as d:
        op = os.path.join(d, "o.bin")
        subprocess.run([ref_path(), "build", str(seed), str(kind), op], check=True, timeout=300)
        return np.fromfile(op, dtype=np.uint8)

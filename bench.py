"""Frame-rate benchmark of the per-frame GPU render path on one MI355X.

Workload: the 1920x1080 textured box-room interior (about 260k triangles in
4096 boxes, 16 value-noise textures with mip chains, a lightmap page, 8 point
lights), rendered along its fixed 120-frame camera path.  One step is one
sft_present of one frame.  The scene is resident in HBM after warm-up and the
frame stays in HBM (device output target), so the timed value is GPU-resident
throughput; the PCIe-inclusive rate with the drop-in host output is reported
beside it.

Prints one JSON line: frames/s, Mpix/s, latency percentiles, the roofline of
the dominant kernel (the pass with the longest average time) from the
algorithmic-byte formulas and the frame's own counters, and the reference
renderer timed on this host's cores.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "1080p frames/s & Mpix/s; p50/p99 request latency, served vs per-frame GPU"
HBM_PEAK_GBS = 8000.0
SCENES = {"plane_spheres": 0, "box_room": 1, "crowd_shadows": 2}
KERNEL_OF_PASS = {"geometry": "k_geom", "binning": "k_clip+k_bin", "visibility": "k_raster<1>", "shading": "k_shade"}
# the shading kernel is specialised for views with and without capsule shadows
SHADOW_SCENES = {"crowd_shadows"}


def pass_kernel(pass_name, scene):
    k = KERNEL_OF_PASS[pass_name]
    if k == "k_shade":
        k += "<true>" if scene in SHADOW_SCENES else "<false>"
    return k
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_{workload}.json")


def summarize_pmc(fetch_csv, write_csv, copy_bytes, workload):
    """Per-kernel HBM bytes per dispatch from two rocprofv3 --pmc passes
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by the factor
    that makes the streaming copy kernel over `copy_bytes` read and write
    exactly copy_bytes."""
    import csv

    def per_kernel(path, counter):
        acc = {}
        for row in csv.DictReader(open(path)):
            if row["Counter_Name"] != counter:
                continue
            name = row["Kernel_Name"].split("(")[0].replace("void ", "").strip()
            acc.setdefault(name, []).append(float(row["Counter_Value"]) * 1024.0)
        return {k: (float(np.mean(v)), len(v)) for k, v in acc.items()}

    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    f_factor = copy_bytes / fetch["sft::k_copy"][0]
    w_factor = copy_bytes / write["sft::k_copy"][0]
    kernels = {}
    for k in sorted(set(fetch) & set(write)):
        if not k.startswith("sft::"):
            continue
        kernels[k.replace("sft::", "")] = {
            "dispatches": fetch[k][1], "fetch_bytes": round(fetch[k][0] * f_factor),
            "write_bytes": round(write[k][0] * w_factor),
            "hbm_bytes": round(fetch[k][0] * f_factor + write[k][0] * w_factor)}
    return {"workload": workload, "copy_calibration_bytes": copy_bytes, "fetch_factor": round(f_factor, 4),
            "write_factor": round(w_factor, 4), "kernels": kernels}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary of this
    workload, or None."""
    path = TRAFFIC_FILE.format(workload=workload)
    if not os.path.exists(path):
        return None
    k = json.load(open(path)).get("kernels", {}).get(kernel)
    return int(k["hbm_bytes"]) if k else None


PASS_KERNELS = {"geometry": ["k_geom", "k_scan_sums", "k_scan_apply"],
                "binning": ["k_clip", "k_bin_count", "k_bin_scan", "k_bin"],
                "visibility": ["k_raster_plan", "k_raster<1>", "k_raster<2>", "k_tile_lights"],
                "shading": ["k_shade<false>", "k_shade<true>", "k_shade_ms<false>", "k_shade_ms<true>"]}


def measured_pass_gbs(workload, pass_ms):
    """Each pass's rocprof-measured HBM bytes per frame (the committed PMC
    summary of this workload: FETCH_SIZE + WRITE_SIZE of the pass's kernels
    per dispatch, calibrated) over the pass's event-timed duration, or None
    without a summary."""
    path = TRAFFIC_FILE.format(workload=workload)
    if not os.path.exists(path):
        return None
    ks = json.load(open(path)).get("kernels", {})
    out = {}
    for name, kernels in PASS_KERNELS.items():
        b = sum(ks[k]["hbm_bytes"] for k in kernels if k in ks)
        out[name] = round(b / (pass_ms[name] * 1e-3) / 1e9, 1) if pass_ms[name] > 0 else None
    return out


def algorithmic_bytes(st):
    """Per-frame algorithmic HBM bytes of each pass (counts from the frame's
    own counters): compulsory traffic plus the reference's materialised
    intermediates, each counted once written and once read."""
    geometry = (36 * st.meshlets_tested + st.meshlets_visible * (768 + 768 + 384) + 768 * st.meshlets_skinned
                + 2304 * st.meshlets_visible)
    binning = 8 * st.tile_payloads + 164 * st.clip_triangles
    visibility = 8 * st.tile_payloads + 288 * st.tile_payloads + 4 * st.samples
    shading = (4 * st.samples + 156 * st.shading_groups + 16 * st.tex_bilinear + 4 * st.tex_nearest
               + 16 * st.lightmapped_pixels + 4 * st.pixels)
    return {"geometry": geometry, "binning": binning, "visibility": visibility, "shading": shading}


def physical_cores():
    """Physical cores of this host (lscpu: cores per socket x sockets), or
    None."""
    try:
        txt = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, timeout=30).stdout
        info = dict(line.split(":", 1) for line in txt.splitlines() if ":" in line)
        return int(info["Core(s) per socket"].strip()) * int(info["Socket(s)"].strip())
    except (OSError, subprocess.SubprocessError, KeyError, ValueError):
        return None


def cpu_baseline(scene, name, width, height, frames):
    exe = os.path.join(ROOT, "oracle", "_ref", "sofren2_ref")
    if not os.path.exists(exe):
        return None
    cores = os.cpu_count() or 1
    try:
        proc = subprocess.run([exe, "bench", str(scene), "0", str(width), str(height), str(cores), "5", str(frames)],
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=600, check=True)
        r = json.loads(proc.stdout.strip().splitlines()[-1])
    except (subprocess.SubprocessError, ValueError) as e:
        print(f"cpu baseline failed: {e}", file=sys.stderr)
        return None
    return {"value": round(r["fps"], 3), "unit": "frames/s", "cores": r["threads"], "kind": "reference",
            "logical_cpus": cores, "physical_cores": physical_cores(),
            "sample": f"{name} {width}x{height}, {frames} frames of the camera path after 5 warm-up frames, "
                      f"reference AVX2 build with {r['threads']} worker threads (logical CPUs); "
                      f"median {r['ms_median']:.2f} ms, p99 {r['ms_p99']:.2f} ms"}


def tuning_table(lib, repeats=3):
    """Every autotuner candidate timed on the 64-view mixed batch (eager
    launches, best of `repeats`): the visibility pass per payloads-per-item and
    per split threshold for each resolution class's views, the visibility
    grid and the occlusion stack on the whole batch.  One JSON object."""
    import ctypes
    import sofren2_amd as S
    from sofren2_amd import workloads
    lib.sft_init(640, 360, 0)
    lib.set_graphs(0)
    mb = workloads.MixedBatch(lib)

    def res_class(w, h):
        return 0 if w * h < (1 << 18) else (1 if w * h < (1 << 20) else 2)

    classes = {k: [mb.views[i] for i, (_, _, w, h, _) in enumerate(mb.spec) if res_class(w, h) == k] for k in range(3)}
    base = S.Tuning()
    lib.get_tuning(base)

    def measure(views, field):
        best = 1e9
        for _ in range(repeats):
            S.render_batch(lib, views)
            st = S.FrameStats()
            lib.get_stats(ctypes.byref(st))
            best = min(best, getattr(st, field))
        return round(best, 4)

    def tuned(**kw):
        t = S.Tuning()
        lib.get_tuning(t)
        for k, v in kw.items():
            if k in ("raster_chunk", "raster_split"):
                getattr(t, k)[:] = v
            else:
                setattr(t, k, v)
        lib.set_tuning(t)

    out = {"views_per_class": {k: len(v) for k, v in classes.items()}, "classes": {}}
    for k, views in classes.items():
        row = {"payloads_per_item": {}, "split_threshold": {}}
        for ch in (64, 96, 128, 256, 384, 512):
            tuned(raster_chunk=[ch] * 3, raster_split=list(base.raster_split))
            row["payloads_per_item"][ch] = measure(views, "ms_raster")
        for sp in (1024, 2048, 4096, 8192, 0xFFFFFFFF):
            tuned(raster_chunk=list(base.raster_chunk), raster_split=[sp] * 3)
            row["split_threshold"]["never" if sp == 0xFFFFFFFF else sp] = measure(views, "ms_raster")
        out["classes"][k] = row
        lib.set_tuning(base)
    out["visibility_grid"] = {}
    for g in (base.raster_groups_resident // 2, base.raster_groups_resident, 2 * base.raster_groups_resident):
        tuned(raster_groups=g)
        out["visibility_grid"][g] = measure(mb.views, "ms_raster")
    lib.set_tuning(base)
    out["occlusion_stack"] = {}
    for stack in (256, 512, 1024):
        tuned(occlusion_stack=stack)
        out["occlusion_stack"][stack] = measure(mb.views, "ms_shadows")
    lib.set_tuning(base)
    mb.close()
    lib.sft_release()
    return out


def mixed_batch_leg(lib, steps):
    """64 views over 4 scene instances at mixed resolution and MSAA: one
    batched submission per step, the next batch recorded while the previous
    one runs (three in flight), against the same views presented one by one
    through sft_present.  With the frames left in HBM (device output:
    presents also keep four frames in flight), with every frame copied to its
    host buffer (the drop-in output), and with every view's ARGB frame and
    64-bit visibility delivered to host memory (the full output of the
    workload; per frame, the visibility capture of each present)."""
    from sofren2_amd.workloads import MixedBatch
    out = _mixed_modes(lib, steps, MixedBatch(lib), ((1, "device"), (0, "host")), tune=True)
    out["host_vis64"] = _mixed_modes(lib, steps, MixedBatch(lib, vis64=True), ((0, "host_vis64"),))["host_vis64"]
    return out


def _mixed_modes(lib, steps, mb, modes, tune=False):
    out = {"views": len(mb.views), "megapixels_per_batch": round(mb.pixels() / 1e6, 3)}
    try:
        lib.set_output(0)
        if tune:
            # the autotuner picks the visibility work split per resolution
            # class on this view set before anything is timed; every leg uses
            # its choice
            before = mb.S.Tuning()
            lib.get_tuning(before)
            out["tuning_default"] = before.as_dict()
            t0 = time.perf_counter()
            out["tuning"] = mb.tune().as_dict()
            out["tuning"]["tune_s"] = round(time.perf_counter() - t0, 2)
        for mode, tag in modes:
            lib.set_output(mode)
            mb.render_batched()
            mb.render_per_frame()
            lib.synchronize()
            # two batches per frame set: the second submission of a set with
            # the same launch shape captures its HIP graph
            mb.render_pipelined(6)
            t0 = time.perf_counter()
            lat = np.array([x * 1e3 for x in mb.render_pipelined(steps)])
            batched_s = (time.perf_counter() - t0) / steps
            t0 = time.perf_counter()
            per_steps = max(1, steps // 2)
            for _ in range(per_steps):
                mb.render_per_frame()
            lib.synchronize()
            per_frame_s = (time.perf_counter() - t0) / per_steps
            n, px = len(mb.views), mb.pixels()
            out[tag] = {"batched_views_s": round(n / batched_s, 2), "per_frame_views_s": round(n / per_frame_s, 2),
                        "batched_mpix_s": round(px / batched_s / 1e6, 1),
                        "per_frame_mpix_s": round(px / per_frame_s / 1e6, 1),
                        "speedup": round(per_frame_s / batched_s, 3),
                        "batch_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                             "p99": round(float(np.percentile(lat, 99)), 3)}}
        lib.set_output(1)
        return out
    finally:
        mb.close()


def crowd_leg(lib, steps, warmup):
    """The skinned crowd with capsule shadows at 1920x1080 (the box room, 32
    humanoids of 6144 triangles, 16 shadow-casting lights): frames/s with
    the frame kept in HBM as in the headline leg, then per-frame counters
    from eager launches: BVH proxy-occlusion rays (Mrays/s over the
    occlusion pass and at the frame rate) and the capsule ray tests of the
    reference's shading (per shaded pixel and light reaching it, every
    capsule cached for the light; the kernel itself first drops the capsules
    no ray of the wave passes near, so it evaluates fewer)."""
    import sofren2_amd as S
    scene, w, h = SCENES["crowd_shadows"], 1920, 1080
    lib.sft_resize(w, h)
    if lib.scene_create(scene, 0, w, h) != 0:
        raise SystemExit("crowd scene creation failed")
    tris = lib.scene_tris()
    color = np.zeros((h, w), dtype=np.uint32)
    lib.set_output(1)
    lib.set_graphs(1)
    for i in range(warmup):
        lib.scene_render(i % 120, color.ctypes.data, w * 4)
    lib.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        lib.scene_render((warmup + i) % 120, color.ctypes.data, w * 4)
    lib.synchronize()
    fps = steps / (time.perf_counter() - t0)
    lib.set_graphs(0)
    st = S.FrameStats()
    rays, caps, occ_ms, shade_ms, n = 0, 0, [], [], min(steps, 40)
    for i in range(n):
        lib.scene_render((warmup + i) % 120, color.ctypes.data, w * 4)
        lib.get_stats(st)
        rays += st.occlusion_rays
        caps += st.capsule_tests
        occ_ms.append(st.ms_shadows)
        shade_ms.append(st.ms_shade)
    lib.set_graphs(1)
    lib.scene_destroy()
    rays_f, caps_f = rays / n, caps / n
    return {"workload": f"crowd_shadows_{w}x{h}", "triangles": int(tris), "frames_s": round(fps, 2),
            "mpix_s": round(fps * w * h / 1e6, 1), "mtri_s": round(fps * tris / 1e6, 1),
            "occlusion_rays_per_frame": round(rays_f, 1),
            "occlusion_mrays_s": round(rays_f * fps / 1e6, 2),
            "occlusion_pass_ms": round(float(np.mean(occ_ms)), 4),
            "occlusion_mrays_s_in_pass": round(rays_f / (float(np.mean(occ_ms)) * 1e3), 2),
            "capsule_tests_per_frame": round(caps_f, 1),
            "capsule_mtests_s": round(caps_f * fps / 1e6, 1),
            "shade_pass_ms": round(float(np.mean(shade_ms)), 4),
            "capsule_mtests_s_in_shade_pass": round(caps_f / (float(np.mean(shade_ms)) * 1e3), 1)}


def served_accuracy(sample, specs, id_splits, max_checks=48, workers=8):
    """Served frames of the 1% sample against the reference renderer
    (oracle/_ref, one worker thread, run on this host) rendering the same
    template pose and request camera: visibility IDs and depth bit-exact,
    colour within 1 LSB on at least 99.9% of pixels.  Part of the
    CPU-baseline leg: the reference is the checker here."""
    from concurrent.futures import ThreadPoolExecutor
    from sofren2_amd import workloads
    sys.path.insert(0, ROOT)
    from tests import parity
    from tests.refrun import have_ref, ref_dump
    picks = sample[:max_checks]
    if not picks or not all(have_ref(ns, bits) for ns in (1, 2, 4) for bits in set(id_splits)):
        return None

    def check(item):
        (_, si, cf, w, h, ns), (color, vis) = item
        kind, seed, pose = specs[si]
        ref = ref_dump(kind, seed, w, h, workloads.reference_frame_code(pose, cf), samples=ns,
                       id_bits=id_splits[si])
        got = {"color": color, "vid": (vis & 0xFFFFFFFF).astype(np.uint32),
               "depth": (vis >> np.uint64(32)).astype(np.uint32).view(np.float32)}
        rep = parity.compare_frame(got, ref)
        ids_equal = rep["vid_mismatch"] == 0 and rep["depth_mismatch"] == 0
        return ids_equal, rep["color_within1_frac"], rep["color_exact_frac"], w * h

    with ThreadPoolExecutor(max_workers=workers) as ex:
        res = list(ex.map(check, picks))
    return {"checked": len(res), "visibility_bit_exact": sum(r[0] for r in res),
            "color_within_1lsb_999": sum(r[1] >= 0.999 for r in res),
            "min_within_1lsb_frac": round(min(r[1] for r in res), 6),
            "mean_exact_frac": round(float(np.mean([r[2] for r in res])), 6),
            "largest_request_px": max(r[3] for r in res)}


def served_leg(lib, rates, n_requests, n_scenes, sample_frac=0.01):
    """Poisson arrivals over resident scene templates at each rate; each
    request a random template, camera frame, resolution (128x128 to 3840x2160
    pixels, log-uniform) and MSAA.  Returns throughput and latency per rate;
    the requests of the last rate rendered one by one through sft_present
    (the per-frame path) for comparison; and a 1% sample of the served frames
    checked against the same request presented alone (the per-frame path
    itself equals the reference in the parity tests)."""
    from sofren2_amd import workloads
    specs = workloads.served_scene_specs(n_scenes)
    srv = workloads.Server(lib, specs, budget_bytes=200 << 30, max_views=64)
    out = {"scenes": n_scenes, "budget_gb": 200, "rates": []}
    try:
        # warm-up: make every template resident, then one burst of full
        # batches so that the frame sets' buffers reach their working size
        warm = [srv.submit(srv.request(i, 0, 256, 256, 1)) for i in range(n_scenes)]
        for r in warm:
            srv.wait(r)
        burst = [(0.0,) + q[1:] for q in workloads.served_requests(3 * 64, 1000.0, n_scenes, seed=99)]
        workloads.run_served(srv, burst)
        sample, last_reqs = [], []
        for k, rate in enumerate(rates):
            # at least two seconds of arrivals, so the first batches (small,
            # while the queue builds up) and the last ones (the queue
            # draining) do not dominate the wall time
            n = max(n_requests, int(2 * rate))
            reqs = workloads.served_requests(n, rate, n_scenes, seed=100 + k)
            keep = range(0, n, max(1, int(round(1 / sample_frac))))
            st0 = srv.stats()
            lat, wall, kept = workloads.run_served(srv, reqs, keep=keep)
            st1 = srv.stats()
            batches = st1["batches"] - st0["batches"]
            out["rates"].append({
                "offered_req_s": rate, "requests": n,
                "throughput_req_s": round(n / wall, 2),
                "mpix_s": round(sum(w * h for _, _, _, w, h, _ in reqs) / wall / 1e6, 1),
                "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                               "p99": round(float(np.percentile(lat, 99)), 3)},
                "mean_batch": round((st1["views_in_batches"] - st0["views_in_batches"]) / max(batches, 1), 2),
                "batches": batches,
                # the server thread's time per request: preparing and
                # launching, completing (waits included), and GPU spans
                "host_submit_ms_per_req": round((st1["host_submit_ms"] - st0["host_submit_ms"]) / n, 4),
                "host_complete_ms_per_req": round((st1["host_complete_ms"] - st0["host_complete_ms"]) / n, 4),
                "gpu_ms_per_req": round((st1["gpu_ms"] - st0["gpu_ms"]) / n, 4)})
            sample += [(reqs[i], kept[i]) for i in kept]
            last_reqs = reqs
        out["resident_gb"] = round(srv.stats()["resident_bytes"] / 2**30, 3)
        lib.server_stop()
        # the per-frame path over the same requests, back to back
        per = []
        t0 = time.perf_counter()
        for _, si, cf, w, h, ns in last_reqs:
            a = time.perf_counter()
            srv.present_alone(si, cf, w, h, ns)
            per.append((time.perf_counter() - a) * 1e3)
        per_wall = time.perf_counter() - t0
        best = max(out["rates"], key=lambda r: r["throughput_req_s"])
        out["per_frame"] = {"requests": len(last_reqs), "throughput_req_s": round(len(last_reqs) / per_wall, 2),
                            "latency_ms": {"p50": round(float(np.percentile(per, 50)), 3),
                                           "p99": round(float(np.percentile(per, 99)), 3)},
                            "served_speedup": round(best["throughput_req_s"] * per_wall / len(last_reqs), 3)}
        checked = same = 0
        for (_, si, cf, w, h, ns), (color, vis) in sample:
            c, v = srv.present_alone(si, cf, w, h, ns, keep=True)
            checked += 1
            same += int(np.array_equal(c, color) and np.array_equal(v, vis))
        out["accuracy_sample"] = {"checked": checked, "equal_to_per_frame": same}
        out["_sample"] = sample
        out["_specs"] = specs
        out["_id_splits"] = list(srv.id_splits)
        out["id_split_7_templates"] = sum(1 for b in srv.id_splits if b == 7)
    finally:
        srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=120)
    ap.add_argument("--scene", choices=sorted(SCENES), default="box_room",
                    help="workload; box_room is the headline 1080p textured interior")
    ap.add_argument("--batch-steps", type=int, default=20, help="steps of the 64-view mixed batch (0: skip)")
    ap.add_argument("--tuning-table", action="store_true",
                    help="time every autotuner candidate on the 64-view mixed batch and exit")
    ap.add_argument("--crowd-steps", type=int, default=120,
                    help="frames of the crowd-with-shadows 1080p leg (0: skip)")
    ap.add_argument("--served-requests", type=int, default=200,
                    help="requests per offered rate, at least two seconds of arrivals (0: skip serving)")
    ap.add_argument("--served-rates", default="100,500,2000,5000", help="offered request rates (req/s)")
    ap.add_argument("--served-scenes", type=int, default=32)
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run the streaming copy kernel over this many MiB first (HBM counter calibration)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarise two rocprofv3 --pmc counter_collection.csv files (FETCH_SIZE, WRITE_SIZE) of a "
                         "run with --copy-calibration-mb into per-kernel HBM bytes (JSON on stdout) and exit")
    args = ap.parse_args()
    if args.pmc_summary:
        if args.copy_calibration_mb <= 0:
            raise SystemExit("--pmc-summary needs the --copy-calibration-mb of the profiled run")
        workload = "{}_{}x{}".format(args.scene, *{0: (640, 360), 1: (1920, 1080), 2: (1920, 1080)}[SCENES[args.scene]])
        print(json.dumps(summarize_pmc(args.pmc_summary[0], args.pmc_summary[1], args.copy_calibration_mb << 20,
                                       workload), indent=1))
        return
    if args.tuning_table:
        import sofren2_amd as S
        print(json.dumps({"tuning_table": tuning_table(S.library())}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build renders on a single MI355X and has no multi-GPU path "
              "(views and requests would shard across GPUs with no exchange).", file=sys.stderr)
        sys.exit(2)

    import sofren2_amd as S
    lib = S.library()
    scene = SCENES[args.scene]
    w, h = S.SCENE_SIZE[scene]
    lib.sft_init(w, h, 0)
    if lib.scene_create(scene, 0, w, h) != 0:
        raise SystemExit("scene creation failed")
    tris = lib.scene_tris()
    color = np.zeros((h, w), dtype=np.uint32)
    if args.copy_calibration_mb > 0:
        for _ in range(3):
            ms = lib.copy_calibration(args.copy_calibration_mb << 20)
        print(f"copy calibration: {args.copy_calibration_mb} MiB in {ms:.4f} ms", file=sys.stderr)

    # drop-in path with the frame copied to host memory (PCIe included)
    lib.set_output(0)
    for i in range(3):
        lib.scene_render(i, color.ctypes.data, w * 4)
    lib.synchronize()
    t0 = time.perf_counter()
    n_host = 30
    for i in range(n_host):
        lib.scene_render(i, color.ctypes.data, w * 4)
    lib.synchronize()
    host_fps = n_host / (time.perf_counter() - t0)

    # GPU-resident output; the timed frames replay the captured HIP graph
    lib.set_output(1)
    lib.set_graphs(1)
    for i in range(args.warmup):
        lib.scene_render(i % 120, color.ctypes.data, w * 4)
    lib.synchronize()
    st = S.FrameStats()
    counter_names = [n for n, _ in S.FrameStats._fields_ if not n.startswith("ms_")]
    counters_sum = {n: 0 for n in counter_names}
    # with device output sft_present returns once the frame is launched (four
    # frames in flight); the clock stops after the last frame completes
    t0 = time.perf_counter()
    for i in range(args.steps):
        lib.scene_render((args.warmup + i) % 120, color.ctypes.data, w * 4)
    lib.synchronize()
    elapsed = time.perf_counter() - t0
    fps = args.steps / elapsed
    # per-frame latency: present to completion, one frame at a time
    lat = []
    for i in range(min(args.steps, 60)):
        a = time.perf_counter()
        lib.scene_render((args.warmup + i) % 120, color.ctypes.data, w * 4)
        lib.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)

    # per-pass kernel durations: the same frames launched eagerly so HIP
    # events can bracket each pass on the render stream
    lib.set_graphs(0)
    n_inst = min(args.steps, 60)
    tile_ms, geom_ms, bin_ms, rast_ms, shade_ms, shadow_ms = [], [], [], [], [], []
    bytes_acc = {"geometry": 0, "binning": 0, "visibility": 0, "shading": 0}
    total_gpu_ms = []
    for i in range(n_inst):
        lib.scene_render((args.warmup + i) % 120, color.ctypes.data, w * 4)
        lib.get_stats(st)
        total_gpu_ms.append(st.ms_total)
        for n in counter_names:
            counters_sum[n] += getattr(st, n)
        tile_ms.append(st.ms_raster_shade)
        geom_ms.append(st.ms_geometry)
        bin_ms.append(st.ms_binning)
        rast_ms.append(st.ms_raster)
        shade_ms.append(st.ms_shade)
        shadow_ms.append(st.ms_shadows)
        for k, v in algorithmic_bytes(st).items():
            bytes_acc[k] += v
    lib.set_graphs(1)
    n_div = n_inst
    tile_avg_ms = float(np.mean(tile_ms))
    per_pass = {
        "geometry": (bytes_acc["geometry"] / n_div, float(np.mean(geom_ms))),
        "binning": (bytes_acc["binning"] / n_div, float(np.mean(bin_ms))),
        "visibility": (bytes_acc["visibility"] / n_div, float(np.mean(rast_ms))),
        "shading": (bytes_acc["shading"] / n_div, float(np.mean(shade_ms))),
    }
    counters_avg = {n: round(v / n_inst, 1) for n, v in counters_sum.items()}
    dominant = max(per_pass, key=lambda k: per_pass[k][1])
    dom_bytes, dom_ms = per_pass[dominant]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    lat = np.array(lat)
    out = {
        "metric": METRIC,
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.scene}_{w}x{h}", "width": w, "height": h, "triangles": int(tris),
                   "frames_in_path": 120, "msaa": 1, "output": "device"},
        "mpix_s": round(fps * w * h / 1e6, 2),
        "mtri_s": round(fps * tris / 1e6, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4)},
        "pcie_inclusive_frames_s": round(host_fps, 3),
        "pass_ms": {"geometry": round(float(np.mean(geom_ms)), 4), "binning": round(float(np.mean(bin_ms)), 4),
                    "raster_shade": round(tile_avg_ms, 4), "raster": round(float(np.mean(rast_ms)), 4),
                    "shade": round(float(np.mean(shade_ms)), 4), "shadow_occlusion": round(float(np.mean(shadow_ms)), 4),
                    "gpu_total": round(float(np.mean(total_gpu_ms)), 4)},
        "occlusion_mrays_s": (round(counters_sum["occlusion_rays"] / n_inst / (float(np.mean(shadow_ms)) * 1e3), 2)
                              if counters_sum["occlusion_rays"] else None),
        "roofline": {"bound": "hbm", "kernel": pass_kernel(dominant, args.scene), "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic(f"{args.scene}_{w}x{h}", pass_kernel(dominant, args.scene)),
                     "algorithmic_bytes_per_launch": int(dom_bytes)},
        "pass_roofline_gbs": {k: round(v[0] / (v[1] * 1e-3) / 1e9, 1) for k, v in per_pass.items()},
        "pass_measured_gbs": measured_pass_gbs(f"{args.scene}_{w}x{h}", {k: v[1] for k, v in per_pass.items()}),
        "counters_per_frame": counters_avg,
        "cpu_baseline": None,
    }
    lib.scene_destroy()
    if args.crowd_steps > 0 and args.scene != "crowd_shadows":
        out["crowd"] = crowd_leg(lib, args.crowd_steps, args.warmup)
    if args.batch_steps > 0:
        out["mixed_batch"] = mixed_batch_leg(lib, args.batch_steps)
    if args.served_requests > 0:
        rates = [float(x) for x in args.served_rates.split(",") if x]
        out["served"] = served_leg(lib, rates, args.served_requests, args.served_scenes)
    lib.sft_release()
    served = out.get("served") or {}
    sample, specs = served.pop("_sample", []), served.pop("_specs", [])
    id_splits = served.pop("_id_splits", [])
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, args.scene, w, h, args.cpu_frames)
        if sample:
            served["accuracy_vs_reference"] = served_accuracy(sample, specs, id_splits)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

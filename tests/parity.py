"""Field-by-field comparison of a GPU frame against a reference dump."""
import numpy as np


def color_lsb_stats(a, b):
    """max per-channel difference and fraction of pixels within 1 LSB."""
    ca = np.stack([(a >> s) & 255 for s in (24, 16, 8, 0)], -1).astype(np.int16)
    cb = np.stack([(b >> s) & 255 for s in (24, 16, 8, 0)], -1).astype(np.int16)
    d = np.abs(ca - cb).max(-1)
    return int(d.max()) if d.size else 0, float((d <= 1).mean()) if d.size else 1.0


def compare_frame(gpu, ref, check_color=True):
    """gpu / ref: dicts with color, vid, depth and optionally ss, ssvis, clip
    (compared when gpu has them). Returns a report dict."""
    rep = {}
    if "ssvis" not in gpu:
        return _compare_images(gpu, ref, rep, check_color)
    rep["ssvis_equal"] = bool(np.array_equal(gpu["ssvis"], ref["ssvis"]))
    vis = ref["ssvis"].astype(bool)
    if rep["ssvis_equal"] and vis.any():
        diff = gpu["ss"][vis] != ref["ss"][vis]
        rep["ss_mismatch_words"] = int(diff.sum())
        if rep["ss_mismatch_words"]:
            m, f, v = np.argwhere(diff)[0]
            rep["ss_first"] = (int(np.flatnonzero(vis)[m]), int(f), int(v),
                               hex(int(gpu["ss"][vis][m, f, v])), hex(int(ref["ss"][vis][m, f, v])))
    else:
        rep["ss_mismatch_words"] = -1
    rep["clip_count"] = (int(gpu["clip"].shape[0]), int(ref["clip"].shape[0]))
    if gpu["clip"].shape == ref["clip"].shape:
        rep["clip_mismatch_words"] = int((gpu["clip"] != ref["clip"]).sum())
    else:
        rep["clip_mismatch_words"] = -1
    return _compare_images(gpu, ref, rep, check_color)


def _compare_images(gpu, ref, rep, check_color):
    vd = gpu["vid"] != ref["vid"]
    rep["vid_mismatch"] = int(vd.sum())
    if rep["vid_mismatch"]:
        i = int(np.flatnonzero(vd)[0])
        rep["vid_first"] = (i, hex(int(gpu["vid"][i])), hex(int(ref["vid"][i])))
    rep["depth_mismatch"] = int((gpu["depth"].view(np.uint32) != ref["depth"].view(np.uint32)).sum())
    if check_color:
        rep["color_exact_frac"] = float((gpu["color"] == ref["color"]).mean())
        rep["color_max_lsb"], rep["color_within1_frac"] = color_lsb_stats(gpu["color"], ref["color"])
    return rep


def assert_parity(rep):
    if "ssvis_equal" in rep:
        assert rep["ssvis_equal"], rep
        assert rep["ss_mismatch_words"] == 0, rep
        assert rep["clip_count"][0] == rep["clip_count"][1] and rep["clip_mismatch_words"] == 0, rep
    assert rep["vid_mismatch"] == 0, rep
    assert rep["depth_mismatch"] == 0, rep
    if "color_within1_frac" in rep:
        assert rep["color_within1_frac"] >= 0.999, rep

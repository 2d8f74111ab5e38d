"""The C restatement (oracle) against golden fixtures of the reference.

tests/golden/ holds, per case, digests of the full colour / visibility-ID /
depth planes the reference renderer produced (compiled from its own source,
one worker thread) and crops of them; tests/golden/make_golden.py wrote them.
These pin the restatement without the reference tree.  The reference's
arithmetic goes through the host's rcpps/rsqrtps estimates, which differ
between CPU vendors: on a host of another CPU model a mismatch is reported as
a skip, on the fixture's own model it fails.
"""
import json
import os

import numpy as np
import pytest

from tests.golden.make_golden import cpu_model, frame_summary
from tests.oracle_port import PORT_LIB, port_render_scene

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INDEX = json.load(open(os.path.join(HERE, "golden.json")))

pytestmark = pytest.mark.skipif(not os.path.exists(PORT_LIB), reason="oracle build missing")


@pytest.mark.parametrize("case", INDEX["cases"], ids=[c["name"] for c in INDEX["cases"]])
def test_port_matches_golden(case):
    w, h, ns = case["width"], case["height"], case["samples"]
    d = port_render_scene(case["scene"], case["seed"], w, h, case["frame"], samples=ns)
    got = frame_summary(d, w, h, ns)
    crops = np.load(os.path.join(HERE, case["name"] + ".npz"))
    same = all(got[k] == case[k] for k in ("color_sha256", "vid_sha256", "depth_sha256"))
    if not same and cpu_model() != INDEX["cpu_model"]:
        pytest.skip(f"fixtures made on {INDEX['cpu_model']!r}; this host's estimates may differ")
    assert int(d["ssvis"].sum()) == case["visible_meshlets"]
    assert int(d["clip"].shape[0]) == case["clip_records"]
    for k in ("color_crop", "vid_crop", "depth_crop"):
        np.testing.assert_array_equal(got[k], crops[k], err_msg=k)
    for k in ("color_sha256", "vid_sha256", "depth_sha256"):
        assert got[k] == case[k], k


def test_planted_mistake_is_caught():
    """A one-ULP change to a single depth value of the restatement's output
    (as a dropped term would cause) no longer matches the fixture."""
    case = INDEX["cases"][0]
    w, h, ns = case["width"], case["height"], case["samples"]
    d = port_render_scene(case["scene"], case["seed"], w, h, case["frame"], samples=ns)
    got = frame_summary(d, w, h, ns)
    if got["depth_sha256"] != case["depth_sha256"]:
        pytest.skip("host estimates differ from the fixture's")
    covered = np.flatnonzero(d["vid"] != 0xFFFFFFFF)
    depth = d["depth"].copy()
    i = covered[len(covered) // 2]
    depth[i] = np.nextafter(depth[i], np.float32(np.inf))
    d2 = dict(d, depth=depth)
    assert frame_summary(d2, w, h, ns)["depth_sha256"] != case["depth_sha256"]

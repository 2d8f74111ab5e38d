"""The shadowed shading pass drops, per wave, the capsules that no lane's ray
passes near (capsule_near over the span of the wave's shading positions).
That is exact only if capsule_near keeps every capsule for which some lane's
cheap ray test (ray_near, sofren2.c:6521-6530) passes.  capsule_prop.cpp
checks that implication on the host, built from the kernels' own header,
over millions of seeded capsules placed at both thresholds; a planted
mistake (the threshold without the span radius and margins) must fail it.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ missing")
    exe = str(tmp_path_factory.mktemp("capsule") / "capsule_prop")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-msse2",
                    "-I" + os.path.join(ROOT, "sofren2_amd", "csrc"), "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "capsule_prop.cpp"), "-o", exe], check=True, timeout=300)
    return exe


def _run(exe, cases, seed, *extra):
    out = subprocess.run([exe, str(cases), str(seed), *extra], capture_output=True, text=True, check=True,
                         timeout=300)
    n, rays, near, violations, culled = map(int, out.stdout.split())
    return {"cases": n, "rays": rays, "near": near, "violations": violations, "culled": culled,
            "stderr": out.stderr}


def test_capsule_near_keeps_every_capsule_a_lane_ray_passes_near(harness):
    total_rays = total_near = total_culled = 0
    for seed in (1, 2):
        r = _run(harness, 1_500_000, seed)
        print(seed, r, flush=True)
        assert r["violations"] == 0, r
        total_rays += r["rays"]
        total_near += r["near"]
        total_culled += r["culled"]
    assert total_rays >= 10_000_000
    # the cases exercise both sides: many rays pass the cheap test, and the
    # wave test still culls capsules
    assert total_near > total_rays // 10
    assert total_culled > 10_000


def test_planted_tight_threshold_is_caught(harness):
    r = _run(harness, 200_000, 3, "tight")
    assert r["violations"] > 1000, r

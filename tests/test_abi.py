"""The C-ABI library exports every function the public headers declare.

Loads libsofren2_gpu.so without calling into it (no GPU needed) and checks the
symbols against the prototypes in include/sofren2.h and include/sofren2_ext.h.
"""
import ctypes
import os
import re

import pytest

import sofren2_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROTO = re.compile(r"^(?!\s*(?:SFT_INLINE|typedef|#|static|return))[A-Za-z_][\w\s\*]*?\b(sftx?_\w+)\s*\([^;{}]*\)\s*;",
                   re.M)


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(PROTO.findall(text)))


@pytest.mark.parametrize("header", ["sofren2.h", "sofren2_ext.h"])
def test_every_declared_function_is_exported(header):
    names = declared(header)
    assert len(names) > 10, names
    lib = ctypes.CDLL(S.GPU_LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_public_api_surface():
    names = set(declared("sofren2.h"))
    for n in ("sft_init", "sft_release", "sft_resize", "sft_present", "sft_frame_begin", "sft_frame_end",
              "sft_set_camera", "sft_draw_instance", "sft_draw_point_light", "sft_texture_create",
              "sft_material_create", "sft_model_load_qsf", "sft_dmesh_begin", "sft_arena_alloc"):
        assert n in names, n


def test_ctypes_structs_match_the_header(tmp_path):
    """The ctypes mirrors used by tests and the bench have the C layout of
    the structs in include/sofren2_ext.h (size and field offsets)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    checks = {
        "sftx_tuning": (S.Tuning, ["raster_chunk", "raster_split", "raster_groups", "raster_groups_resident",
                                   "tile_px", "occlusion_stack", "ms_class", "ms_batch", "ms_occlusion"]),
        "sftx_frame_stats": (S.FrameStats, ["meshlets_tested", "shadow_proxies", "ms_upload", "ms_shadows", "tex_npot",
                                            "instances_dropped"]),
        "sftx_request": (S.Request, ["scene", "camera_pos", "view", "proj", "color", "vis64"]),
        "sftx_server_stats": (S.ServerStats, ["submitted", "resident_bytes", "in_flight", "host_submit_ms", "gpu_ms",
                                              "shadow_cache_views"]),
        "sftx_raw_geometry": (S.RawGeometry, ["vx", "v", "j0", "j3", "w0", "w3"]),
        "sft_mesh": (S.Mesh, ["meshlet_count", "inds", "bounds_max_x", "cone_cutoff", "culling_bone_inds",
                              "lightmap_page"]),
        "sft_model": (S.Model, ["total_meshlets", "blas_nodes", "blas_root_ind", "shadow_proxy_count", "anim_count"]),
        "sft_arena": (S.Arena, ["offset", "commit_lock"]),
    }
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "sofren2_ext.h"', "int main(void) {"]
    for name, (_, fields) in checks.items():
        src.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        src += [f'printf("{name} {f} %zu\\n", offsetof({name}, {f}));' for f in fields]
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=gnu11", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {(a, b): int(v) for a, b, v in (line.split() for line in out if line)}
    for name, (cls, fields) in checks.items():
        assert got[(name, "size")] == ctypes.sizeof(cls), name
        for f in fields:
            assert got[(name, f)] == getattr(cls, f).offset, (name, f)

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

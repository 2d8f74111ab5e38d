"""Error paths of the drop-in API on the CPU (no GPU is touched).

The reference's conventions (sofren2.c:238-247): a fatal error prints a
message and exits with status 1; a recoverable one prints a message and
returns a sentinel (NULL, SFT_INVALID_HANDLE, or nothing).  Covered here:
corrupt QSF2 files (sofren2.c:7816-7833; the reference reads a short or
inconsistent file without checks, this build rejects it), stale texture and
material handles (9360-9370, 9409-9419), arena exhaustion (9480-9496) and
display sizes outside what the tile binning encodes.
"""
import ctypes
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import sofren2_amd as S
from tests.mapfile import map_path

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gpu():
    lib = S.library().gpu
    lib.sft_arena_init.restype = None
    lib.sft_arena_init.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    lib.sft_arena_alloc.restype = ctypes.c_void_p
    lib.sft_arena_alloc.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    lib.sft_model_load_qsf.restype = ctypes.c_void_p
    lib.sft_model_load_qsf.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    lib.sft_texture_get_width.restype = ctypes.c_int32
    lib.sft_texture_get_width.argtypes = [ctypes.c_uint32]
    lib.sft_material_create.restype = ctypes.c_uint32
    lib.sft_material_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float]
    lib.sft_material_release.restype = None
    lib.sft_material_release.argtypes = [ctypes.c_uint32]
    return lib


@pytest.fixture
def ctx():
    lib = S.library()
    lib.sft_init(64, 64, 0)
    try:
        yield lib
    finally:
        lib.sft_release()


def _load(path):
    g = _gpu()
    arena = ctypes.create_string_buffer(64)
    g.sft_arena_init(arena, 1 << 32)
    return g.sft_model_load_qsf(path.encode(), arena)


def test_qsf_valid_file_loads(ctx):
    assert _load(map_path())


@pytest.mark.parametrize("case", ["magic", "version", "cut_header", "cut_texture", "cut_middle", "cut_tail",
                                  "huge_count", "bad_root"])
def test_qsf_corrupt_file_is_rejected(ctx, tmp_path, case):
    data = bytearray(open(map_path(), "rb").read())
    if case == "magic":
        data[0:4] = b"QSF1"
    elif case == "version":
        data[4:8] = struct.pack("<I", 3)
    elif case == "cut_header":
        data = data[:10]
    elif case == "cut_texture":
        data = data[:40]
    elif case == "cut_middle":
        data = data[: len(data) // 2]
    elif case == "cut_tail":
        # the model ends where the game data (not read by the loader) starts
        game = struct.unpack_from("<Q", data, 8)[0]
        data = data[:game - 7]
    elif case == "huge_count":
        data[16:20] = struct.pack("<I", 0xFFFFFFFF)   # texture count
    elif case == "bad_root":
        # the model header follows the textures and lightmap pages; find it
        off = 20
        n_tex = struct.unpack_from("<I", data, 16)[0]
        for _ in range(n_tex):
            w, h = struct.unpack_from("<II", data, off)
            off += 8 + w * h * 4
        n_lm = struct.unpack_from("<I", data, off)[0]
        off += 4
        for _ in range(n_lm):
            w, h = struct.unpack_from("<II", data, off)
            off += 8 + w * h * 4
        n_nodes = struct.unpack_from("<I", data, off + 8)[0]
        struct.pack_into("<I", data, off + 12, n_nodes + 5)
    p = tmp_path / "bad.qsf"
    p.write_bytes(bytes(data))
    assert not _load(str(p)), case


def test_qsf_missing_file(ctx, tmp_path):
    assert not _load(str(tmp_path / "none.qsf"))


def test_stale_texture_handles(ctx):
    lib = ctx
    g = _gpu()
    px = np.arange(16, dtype=np.uint32)
    t1 = lib.texture_create(4, 4, px.ctypes.data)
    lib.texture_release(t1)
    lib.texture_release(t1)                  # stale: nothing happens
    t2 = lib.texture_create(2, 8, px.ctypes.data)
    assert t2 != t1 and (t2 & 0xFFFF) == (t1 & 0xFFFF)   # same slot, new generation
    lib.texture_release(t1)                  # still stale: t2 stays live
    assert g.sft_texture_get_width(t2) == 2
    # the size getters read the slot without a generation check, as the
    # reference's do (sofren2.c:9316-9329)
    assert g.sft_texture_get_width(t1) == 2
    assert g.sft_texture_get_width(0xFFFFFFFF) == 0


def test_stale_material_handles(ctx):
    g = _gpu()
    m1 = g.sft_material_create(0xFFFFFFFF, 0xFFFFFFFF, 4.0)
    g.sft_material_release(m1)
    m2 = g.sft_material_create(0xFFFFFFFF, 0xFF00FF00, 8.0)
    assert m2 != m1 and (m2 & 0xFFFF) == (m1 & 0xFFFF)
    g.sft_material_release(m1)               # stale: must not free m2's slot
    m3 = g.sft_material_create(0xFFFFFFFF, 0xFF0000FF, 2.0)
    assert (m3 & 0xFFFF) != (m2 & 0xFFFF)


def _run(code):
    env = dict(os.environ, PYTHONPATH=ROOT)
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)


def test_arena_exhaustion_is_fatal():
    code = ("import ctypes, sofren2_amd as S\n"
            "g = S.library().gpu\n"
            "g.sft_arena_init.argtypes = [ctypes.c_void_p, ctypes.c_uint64]\n"
            "g.sft_arena_alloc.argtypes = [ctypes.c_void_p, ctypes.c_uint64]\n"
            "g.sft_arena_alloc.restype = ctypes.c_void_p\n"
            "a = ctypes.create_string_buffer(64)\n"
            "g.sft_arena_init(a, 1 << 16)\n"
            "assert g.sft_arena_alloc(a, 1 << 15)\n"
            "g.sft_arena_alloc(a, 1 << 16)\n"
            "print('survived')\n")
    r = _run(code)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert "arena out of memory" in r.stderr and "survived" not in r.stdout


def test_display_beyond_256_tiles_is_fatal_at_init():
    r = _run("import sofren2_amd as S\nS.library().sft_init(20000, 100, 0)\nprint('survived')\n")
    assert r.returncode == 1 and "survived" not in r.stdout, (r.returncode, r.stderr)


def test_oversized_view_and_resize_are_refused(ctx):
    lib = ctx
    lib.batch_begin()
    try:
        lib.view_begin(16385, 64, 1)
        assert lib.view_end(None, 16385 * 4, None) == -1
        lib.view_begin(16384, 64, 1)
        assert lib.view_end(None, 16384 * 4, None) == 0
    finally:
        lib.batch_end()
    lib.sft_resize(0, 10)          # refused: the display stays 16384 x 64 from the last view
    lib.sft_resize(17000, 10)


def test_id_split_range_and_default(ctx):
    """sftx_set_id_split takes 1..16 instance bits (the reference's
    SFT_INSTANCE_BITS + SFT_MESHLET_BITS = 24, sofren2.c:154-165) and leaves
    the split unchanged for anything else."""
    lib = ctx
    assert lib.get_id_split() == 6
    lib.set_id_split(7)
    assert lib.get_id_split() == 7
    for bad in (0, 17, -1):
        lib.set_id_split(bad)
        assert lib.get_id_split() == 7
    lib.set_id_split(6)
    assert lib.get_id_split() == 6

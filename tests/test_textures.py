"""Texture mip chains (sofren2.c:9181-9310): the oracle's builder against
closed forms on the CPU, and the chain the GPU builds in HBM against the
oracle's, bit for bit."""
import numpy as np
import pytest

import sofren2_amd as S
from tests.oracle_port import PORT_LIB, port_mip_chain

SIZES = [(1, 1), (4, 4), (256, 256), (100, 37), (3, 200), (1024, 16), (5, 5), (1, 64)]


def levels(w, h):
    out = []
    while True:
        out.append((w, h))
        if (w, h) == (1, 1) or len(out) == 14:
            return out
        w, h = max(w // 2, 1), max(h // 2, 1)


def unswizzle(chain, off, w, h):
    bw = ((w + 3) & ~3) // 4
    y, x = np.mgrid[0:h, 0:w]
    return chain[off + ((y >> 2) * bw + (x >> 2)) * 16 + ((y & 3) << 2) + (x & 3)]


def tent(level):
    """[1 2 2 1]^2 tent with edge clamp and truncating division, per channel (numpy)."""
    h, w = level.shape
    nh, nw = max(h // 2, 1), max(w // 2, 1)
    ch = np.stack([(level >> s) & 255 for s in (24, 16, 8, 0)]).astype(np.int64)
    acc = np.zeros((4, nh, nw), dtype=np.int64)
    wsum = 0
    for ky, wy in zip((-1, 0, 1, 2), (1, 2, 2, 1)):
        sy = np.clip(np.arange(nh) * 2 + ky, 0, h - 1)
        for kx, wx in zip((-1, 0, 1, 2), (1, 2, 2, 1)):
            sx = np.clip(np.arange(nw) * 2 + kx, 0, w - 1)
            acc += ch[:, sy][:, :, sx] * (wx * wy)
            wsum += wx * wy
    q = acc // wsum
    return ((q[0] << 24) | (q[1] << 16) | (q[2] << 8) | q[3]).astype(np.uint32)


needs_oracle = pytest.mark.skipif(not __import__("os").path.exists(PORT_LIB), reason="oracle build missing")


@needs_oracle
@pytest.mark.parametrize("w,h", SIZES)
def test_oracle_chain_matches_closed_form(w, h):
    """Every level of the oracle's chain equals the numpy tent of the level
    above; padding texels are zero; the layout is the documented one."""
    rng = np.random.default_rng(w * 1000 + h)
    img = rng.integers(0, 2**32, size=(h, w), dtype=np.uint64).astype(np.uint32)
    chain = port_mip_chain(img)
    off, cur = 0, img
    used = np.zeros(chain.size, dtype=bool)
    for k, (lw, lh) in enumerate(levels(w, h)):
        assert cur.shape == (lh, lw)
        np.testing.assert_array_equal(unswizzle(chain, off, lw, lh), cur, err_msg=f"level {k}")
        bw = ((lw + 3) & ~3) // 4
        y, x = np.mgrid[0:lh, 0:lw]
        used[off + ((y >> 2) * bw + (x >> 2)) * 16 + ((y & 3) << 2) + (x & 3)] = True
        off += ((lw + 3) & ~3) * ((lh + 3) & ~3)
        cur = tent(cur)
    assert off == chain.size
    assert not chain[~used].any()


@needs_oracle
def test_constant_texture_stays_constant():
    """A constant image keeps its value on every level (weights sum to 36)."""
    img = np.full((64, 48), 0x80FF4010, dtype=np.uint32)
    chain = port_mip_chain(img)
    off = 0
    for lw, lh in levels(48, 64):
        assert (unswizzle(chain, off, lw, lh) == 0x80FF4010).all()
        off += ((lw + 3) & ~3) * ((lh + 3) & ~3)


@pytest.mark.gpu
@pytest.mark.parametrize("w,h", SIZES)
def test_gpu_chain_matches_oracle(w, h):
    """sft_texture_create on the GPU backend: the chain built in HBM is the
    oracle's, texel for texel (padding included)."""
    lib = S.library()
    lib.sft_init(64, 64, 0)
    try:
        rng = np.random.default_rng(w * 7 + h)
        img = rng.integers(0, 2**32, size=(h, w), dtype=np.uint64).astype(np.uint32)
        tex = lib.texture_create(w, h, img.ctypes.data)
        n = lib.cap_texture(tex, None, 0)
        got = np.zeros(n, dtype=np.uint32)
        assert lib.cap_texture(tex, got.ctypes.data, n) == n
        want = port_mip_chain(img)
        print("texture", w, h, "texels", n, flush=True)
        np.testing.assert_array_equal(got, want)
        lib.texture_release(tex)
    finally:
        lib.sft_release()

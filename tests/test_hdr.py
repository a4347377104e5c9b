"""Radiance HDR skydomes (loaders.decode_hdr / load_skydome), the files
HostSkyDome::Load reads through FreeImage (lib/RenderSystem/
host_skydome.cpp:80-139). CPU only. Pinned by the format's definition: an
encoder written here from the Radiance specification (flat and new-style
run-length scanlines) and the RGBE closed form (r, g, b) * 2^(e - 136)."""
import struct

import numpy as np
import pytest

from lighthouse2_amd import loaders, scenes


def _rgbe(rgb):
    """float RGB -> RGBE bytes (the largest component's frexp exponent)."""
    rgb = np.asarray(rgb, np.float64)
    m = rgb.max(-1)
    out = np.zeros(rgb.shape[:-1] + (4,), np.uint8)
    nz = m > 1e-32
    mant, e = np.frexp(m[nz])
    f = mant * 256.0 / m[nz]
    out[nz, :3] = np.clip(np.floor(rgb[nz] * f[:, None]), 0, 255).astype(np.uint8)
    out[nz, 3] = (e + 128).astype(np.uint8)
    return out


def _rle_channel(v):
    out, x, n = bytearray(), 0, len(v)
    while x < n:
        run = 1
        while x + run < n and run < 127 and v[x + run] == v[x]:
            run += 1
        if run >= 3:
            out += bytes([128 + run, v[x]])
            x += run
        else:
            lit = 1
            while x + lit < n and lit < 128 and not (x + lit + 2 < n and v[x + lit] == v[x + lit + 1] == v[x + lit + 2]):
                lit += 1
            out += bytes([lit]) + bytes(v[x:x + lit])
            x += lit
    return bytes(out)


def _encode(rgbe, rle):
    h, w = rgbe.shape[:2]
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        if rle:
            out += bytes([2, 2, w >> 8, w & 255])
            for c in range(4):
                out += _rle_channel(rgbe[y, :, c].tolist())
        else:
            out += rgbe[y].tobytes()
    return bytes(out)


def _decoded(rgbe):
    e = rgbe[..., 3].astype(np.int32)
    return np.where(e[..., None] > 0, rgbe[..., :3] * np.ldexp(1.0, e - 136)[..., None], 0.0).astype(np.float32)


@pytest.mark.parametrize("rle", [False, True])
def test_rgbe_closed_form_and_orientation(rle):
    rng = np.random.default_rng(4)
    h, w = 9, 40
    rgb = rng.uniform(0, 50, (h, w, 3)) ** 2
    rgb[:, 10:30] = rgb[:, 10:11]             # long runs for the run-length coder
    rgb[3] = 0                                # a black row: e = 0
    rgbe = _rgbe(rgb)
    px = loaders.decode_hdr(_encode(rgbe, rle))
    assert px.shape == (h, w, 3) and px.dtype == np.float32
    assert np.array_equal(px, _decoded(rgbe))
    assert np.all(px[3] == 0)
    # within one mantissa step of the encoded radiance (8-bit mantissa relative to the largest component)
    m = rgb.max(-1, keepdims=True)
    assert np.all(np.abs(px - rgb) <= m * 2.0 ** -7 + 1e-30)
    # top row first: the first scanline of the file is row 0
    one = rgbe.copy()
    one[0] = _rgbe(np.full((w, 3), 7.0))
    assert np.allclose(loaders.decode_hdr(_encode(one, rle))[0], 7.0, rtol=2 ** -7)


def test_skydome_from_hdr_and_bin_cache(tmp_path):
    rng = np.random.default_rng(5)
    rgbe = _rgbe(rng.uniform(0, 4, (6, 12, 3)))
    p = tmp_path / "sky.hdr"
    p.write_bytes(_encode(rgbe, True))
    s = scenes.Scene()
    px = loaders.load_skydome(s, str(p), scale=(2.0, 1.0, 0.5))
    assert np.array_equal(px, (_decoded(rgbe) * np.float32([2.0, 1.0, 0.5])).astype(np.float32))
    assert s.sky[0] is px and np.array_equal(s.sky[1], np.eye(4, dtype=np.float32).reshape(16))
    # the reference's cache: uint32 width, height, then float3 pixels (top row first)
    b = tmp_path / "sky.bin"
    b.write_bytes(struct.pack("<II", 12, 6) + _decoded(rgbe).tobytes())
    assert np.array_equal(loaders.load_skydome(scenes.Scene(), str(b)), _decoded(rgbe))


def test_hdr_errors():
    with pytest.raises(ValueError):
        loaders.decode_hdr(b"P6\n")
    with pytest.raises(ValueError, match="orientation"):
        loaders.decode_hdr(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n+Y 2 +X 2\n" + bytes(16))

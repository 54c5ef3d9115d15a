"""PNG decoding behind vkvg_surface_create_from_image (src/vkvg_surface.c:181-194; the reference
decodes with stb_image, vendored as src/stb_image.h). The library's decoder is compared, byte for
byte, with PIL's independent decoder on images of every colour type, bit depth and filter, plain and
Adam7-interlaced, written by the small encoder below, and on tests/golden/filled.png (a data file of
the reference's tests/data). 16-bit samples keep their high byte (PIL clips them instead, so those
are checked against the written samples)."""
import ctypes
import io
import os
import struct
import zlib

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _lib():
    import vkvg_amd
    L = vkvg_amd.lib()
    L.vkvg_amd_png_decode.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32)]
    L.vkvg_amd_png_decode.restype = ctypes.c_void_p
    L.vkvg_amd_png_free.argtypes = [ctypes.c_void_p]
    return L


def _decode(L, data: bytes):
    w, h = ctypes.c_uint32(), ctypes.c_uint32()
    p = L.vkvg_amd_png_decode(data, len(data), ctypes.byref(w), ctypes.byref(h))
    if not p:
        return None
    try:
        return np.ctypeslib.as_array((ctypes.c_uint8 * (w.value * h.value * 4)).from_address(p)) \
            .reshape(h.value, w.value, 4).copy()
    finally:
        L.vkvg_amd_png_free(p)


def _chunk(t: bytes, d: bytes) -> bytes:
    return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)


def _pack_rows(samples: np.ndarray, depth: int) -> list[bytes]:
    """samples: rows x (pixels * channels) integer samples -> packed row bytes"""
    rows = []
    for r in samples:
        if depth == 16:
            rows.append(r.astype(">u2").tobytes())
        elif depth == 8:
            rows.append(r.astype(np.uint8).tobytes())
        else:
            bits = np.unpackbits(r.astype(np.uint8)[:, None], axis=1)[:, 8 - depth:].reshape(-1)
            rows.append(np.packbits(bits).tobytes())
    return rows


def _filter(row: bytes, prev: bytes, bpp: int, f: int) -> bytes:
    out = bytearray([f])
    for i, x in enumerate(row):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i] if prev else 0
        c = prev[i - bpp] if prev and i >= bpp else 0
        if f == 0:
            p = 0
        elif f == 1:
            p = a
        elif f == 2:
            p = b
        elif f == 3:
            p = (a + b) >> 1
        else:
            pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
            p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
        out.append((x - p) & 255)
    return bytes(out)


def _encode(samples, w, h, depth, ctype, channels, interlace=False, plte=None, trns=None, seed=0):
    """PNG bytes of samples (h x w x channels), every filter type used at random"""
    rng = np.random.default_rng(seed)
    bpp = max(1, depth * channels // 8)
    passes = ([(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
              if interlace else [(0, 0, 1, 1)])
    raw = b""
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        rows = _pack_rows(sub.reshape(sub.shape[0], -1), depth)
        prev = None
        for r in rows:
            raw += _filter(r, prev, bpp, int(rng.integers(0, 5)))
            prev = r
    png = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    if plte is not None:
        png += _chunk(b"PLTE", bytes(plte))
    if trns is not None:
        png += _chunk(b"tRNS", bytes(trns))
    z = zlib.compress(raw, 6)
    png += _chunk(b"tEXt", b"k\x00v") + _chunk(b"IDAT", z[: len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])
    return png + _chunk(b"IEND", b"")


def _pil_rgba(data: bytes) -> np.ndarray:
    from PIL import Image
    return np.asarray(Image.open(io.BytesIO(data)).convert("RGBA"))


CASES = [(0, d) for d in (1, 2, 4, 8)] + [(2, 8), (3, 1), (3, 2), (3, 4), (3, 8), (4, 8), (6, 8)]
CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("ctype,depth", CASES)
def test_decode_equals_pil(ctype, depth, interlace):
    L = _lib()
    rng = np.random.default_rng(ctype * 100 + depth + 7 * interlace)
    for w, h in ((1, 1), (3, 2), (13, 9), (37, 21)):
        ch = CHANNELS[ctype]
        plte = trns = None
        if ctype == 3:
            n_pal = min(1 << depth, 200)
            samples = rng.integers(0, n_pal, (h, w, 1))
            plte = rng.integers(0, 256, 3 * n_pal).astype(np.uint8)
            trns = rng.integers(0, 256, max(1, n_pal // 2)).astype(np.uint8)
        else:
            samples = rng.integers(0, 1 << depth, (h, w, ch))
            if ctype in (0, 2) and w * h > 4:  # a colour key present in the image
                trns = struct.pack(">" + "H" * ch, *[int(v) for v in samples[1, 1]])
        data = _encode(samples, w, h, depth, ctype, ch, interlace, plte, trns, seed=w)
        got = _decode(L, data)
        assert got is not None, (ctype, depth, w, h)
        exp = _pil_rgba(data).copy()
        if ctype == 0 and depth < 8 and trns is not None:
            # PIL compares the key with the scaled grey and keeps these pixels opaque; stb_image scales
            # the key too, so the keyed pixels are transparent
            exp[..., 3] = np.where(samples[..., 0] == samples[1, 1, 0], 0, 255)
        assert np.array_equal(got, exp), (ctype, depth, interlace, w, h)


@pytest.mark.parametrize("ctype", [0, 2, 4, 6])
def test_sixteen_bit_samples_keep_their_high_byte(ctype):
    L = _lib()
    rng = np.random.default_rng(ctype)
    ch = CHANNELS[ctype]
    w, h = 11, 7
    s = rng.integers(0, 65536, (h, w, ch))
    got = _decode(L, _encode(s, w, h, 16, ctype, ch, interlace=ctype == 6))
    hi = (s >> 8).astype(np.uint8)
    exp = np.empty((h, w, 4), np.uint8)
    exp[..., 3] = 255
    if ctype in (0, 4):
        exp[..., :3] = hi[..., :1]
        if ctype == 4:
            exp[..., 3] = hi[..., 1]
    else:
        exp[..., :ch] = hi
    assert np.array_equal(got, exp)


def test_reference_test_image():
    L = _lib()
    data = open(os.path.join(HERE, "golden", "filled.png"), "rb").read()
    assert np.array_equal(_decode(L, data), _pil_rgba(data))


def test_malformed_files_are_rejected():
    L = _lib()
    good = _encode(np.zeros((4, 4, 3), np.int64), 4, 4, 8, 2, 3)
    assert _decode(L, good) is not None
    assert _decode(L, b"") is None
    assert _decode(L, good[:8]) is None  # no IHDR
    assert _decode(L, good[:-40]) is None  # truncated
    bad_depth = bytearray(good)
    bad_depth[24] = 3  # bit depth 3
    assert _decode(L, bytes(bad_depth)) is None
    # a palette index beyond the palette
    pal = _encode(np.full((2, 2, 1), 3), 2, 2, 8, 3, 1, plte=bytes(6))
    assert _decode(L, pal) is None
    # an unknown critical chunk
    crit = good[:33] + _chunk(b"ABCD", b"x") + good[33:]
    assert _decode(L, crit) is None
    # tRNS as stb_image rejects it (src/stb_image.h:5006-5039): before PLTE, longer than the
    # palette, after IDAT, on an image with alpha, or of the wrong length for a colour key
    idx = np.zeros((2, 2, 1), np.int64)
    ok_pal = _encode(idx, 2, 2, 8, 3, 1, plte=bytes(6), trns=b"\x10\x20")
    assert _decode(L, ok_pal) is not None
    hdr_end = 8 + 25
    plte_chunk = _chunk(b"PLTE", bytes(6))
    assert ok_pal[hdr_end:hdr_end + len(plte_chunk)] == plte_chunk
    trns_chunk = _chunk(b"tRNS", b"\x10\x20")
    before_plte = ok_pal[:hdr_end] + trns_chunk + plte_chunk + ok_pal[hdr_end + len(plte_chunk) + len(trns_chunk):]
    assert _decode(L, before_plte) is None
    assert _decode(L, _encode(idx, 2, 2, 8, 3, 1, plte=bytes(6), trns=b"\x01\x02\x03")) is None
    iend = _chunk(b"IEND", b"")
    assert _decode(L, good[:-len(iend)] + _chunk(b"tRNS", bytes(6)) + iend) is None
    rgba = _encode(np.zeros((2, 2, 4), np.int64), 2, 2, 8, 6, 4, trns=bytes(6))
    assert _decode(L, rgba) is None
    assert _decode(L, _encode(np.zeros((2, 2, 3), np.int64), 2, 2, 8, 2, 3, trns=bytes(4))) is None


def test_colour_key_of_an_8_bit_image_keeps_its_low_byte():
    # stb_image compares 8-bit samples with (key & 255): a key of 0x0105 makes samples of 5 transparent
    L = _lib()
    s = np.array([[[5], [6]], [[5], [0x05]]], np.int64)
    got = _decode(L, _encode(s, 2, 2, 8, 0, 1, trns=b"\x01\x05"))
    assert got[..., 3].tolist() == [[0, 255], [0, 0]]
    rgb = np.array([[[1, 2, 3], [1, 2, 4]]], np.int64)
    got = _decode(L, _encode(rgb, 2, 1, 8, 2, 3, trns=b"\xff\x01\x00\x02\x01\x03"))
    assert got[..., 3].tolist() == [[0, 255]]


@pytest.mark.gpu
def test_surface_create_from_image(device, tmp_path):
    import vkvg_amd
    L = _lib()
    path = tmp_path / "img.png"
    data = open(os.path.join(HERE, "golden", "filled.png"), "rb").read()
    path.write_bytes(data)
    s = vkvg_amd.Surface.from_image(device, str(path))
    try:
        assert (s.width, s.height) == (141, 140)
        assert np.array_equal(s.premultiplied(), _decode(L, data))  # the bytes as decoded
    finally:
        s.close()
    for bad in (str(tmp_path / "missing.png"), os.path.join(HERE, "conftest.py")):
        with pytest.raises(vkvg_amd.VkvgError, match="status 2"):  # VKVG_STATUS_NULL_POINTER
            vkvg_amd.Surface.from_image(device, bad)

"""PNG output (host only, no GPU).

The file is read back with Python's own zlib (which checks the stream's
Adler-32) and zlib.crc32 for every chunk, so the container format, the
stored deflate blocks and both checksums are pinned by independent code.
Pixel values: clamp to [0, 1] and round to 8 bits (NaN -> 0), or the sRGB
transfer curve at known points; rows run top to bottom in the file, so
pt_read_image's bottom-first rows are flipped unless PT_PNG_TOP_DOWN.
"""
import struct
import zlib

import numpy as np
import pytest

import ptamd


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, 3 * w + 1)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 3)


def test_linear_values_and_row_order(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.uniform(-0.3, 1.3, (9, 13, 4)).astype(np.float32)
    img[0, 0, 0] = np.nan
    p = tmp_path / "a.png"
    ptamd.write_png(p, img)
    got = read_png(p)
    want = np.floor(np.clip(np.nan_to_num(img[..., :3].astype(np.float64), nan=0.0), 0, 1) * 255 + 0.5)
    np.testing.assert_array_equal(got, want[::-1].astype(np.uint8))  # bottom row first in memory
    ptamd.write_png(p, img, top_down=True)
    np.testing.assert_array_equal(read_png(p), want.astype(np.uint8))


def test_srgb_known_points(tmp_path):
    vals = np.array([0.0, 0.001, 0.0031308, 0.5, 0.2140, 1.0], np.float32)
    img = np.zeros((1, len(vals), 4), np.float32)
    img[0, :, 0] = img[0, :, 1] = img[0, :, 2] = vals
    p = tmp_path / "s.png"
    ptamd.write_png(p, img, srgb=True)
    got = read_png(p)[0, :, 0]
    # 0.001 -> 12.92 * 0.001 * 255 = 3.29; 0.0031308 (the two pieces meet:
    # 10.31 either way); 0.5 -> 255 * 0.73536 = 187.52; 0.2140 -> 127.49
    assert list(got) == [0, 3, 10, 188, 127, 255]


def test_multiple_stored_blocks(tmp_path):
    # 300 x 200 RGB rows = 180,200 bytes: three stored blocks of <= 65,535
    img = np.linspace(0, 1, 300 * 200 * 4, dtype=np.float32).reshape(200, 300, 4)
    p = tmp_path / "big.png"
    ptamd.write_png(p, img, top_down=True)
    got = read_png(p)
    np.testing.assert_array_equal(got, np.floor(img[..., :3].astype(np.float64) * 255 + 0.5).astype(np.uint8))


def test_errors(tmp_path):
    img = np.zeros((2, 2, 4), np.float32)
    with pytest.raises(ptamd.PtError) as e:
        ptamd.write_png(tmp_path / "no" / "such" / "dir.png", img)
    assert "cannot open" in str(e.value)
    with pytest.raises(ValueError):
        ptamd.write_png(tmp_path / "x.png", np.zeros((2, 2, 3), np.float32))

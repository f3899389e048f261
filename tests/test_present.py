"""PNG writer of the presentation step (tpd_write_png; no GPU needed).  The file is decoded here
by an independent reader: chunk CRCs with binascii.crc32, the zlib stream with Python's zlib
(which also checks the Adler-32 trailer), then scanline filters undone.  Every pixel must come
back bit for bit."""
import binascii
import struct
import zlib

import numpy as np
import pytest


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == binascii.crc32(kind + body) & 0xFFFFFFFF, kind
        chunks.append((kind, body))
        pos += 12 + n
    assert pos == len(data)
    assert chunks[0][0] == b"IHDR" and chunks[-1] == (b"IEND", b"")
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for k, b in chunks if k == b"IDAT"))
    row = 1 + 4 * w
    assert len(raw) == row * h
    rows = np.frombuffer(raw, np.uint8).reshape(h, row)
    assert np.all(rows[:, 0] == 0)  # filter type None on every scanline
    return rows[:, 1:].reshape(h, w, 4).copy()


@pytest.mark.parametrize("w,h", [(1, 1), (37, 13), (300, 200), (16383, 1), (1, 70000)])
def test_png_round_trip(tmp_path, w, h):
    from torpedo_amd.engine import write_png
    rng = np.random.default_rng(w * 7 + h)
    img = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    path = tmp_path / "frame.png"
    write_png(img, path)
    assert np.array_equal(read_png(path), img)


def test_png_block_boundary_on_a_row_boundary(tmp_path):
    # scanlines of 1 + 4 * 5461 = 21845 bytes: three fill a 65535-byte stored block exactly
    from torpedo_amd.engine import write_png
    w, h = 5461, 7
    img = (np.arange(h * w * 4) % 251).astype(np.uint8).reshape(h, w, 4)
    path = tmp_path / "edge.png"
    write_png(img, path)
    assert np.array_equal(read_png(path), img)


def test_png_rejects_bad_arguments(tmp_path):
    from torpedo_amd._capi import TpdError, lib
    from torpedo_amd.engine import write_png
    with pytest.raises(ValueError):
        write_png(np.zeros((4, 4, 3), np.uint8), tmp_path / "rgb.png")
    with pytest.raises(TpdError):
        write_png(np.zeros((0, 4, 4), np.uint8), tmp_path / "empty.png")
    with pytest.raises(TpdError):
        write_png(np.zeros((2, 2, 4), np.uint8), tmp_path / "no_such_dir" / "x.png")
    img = np.zeros((2, 2, 4), np.uint8)
    assert lib().tpd_write_png(None, 2, 2, str(tmp_path / "n.png").encode()) != 0
    assert lib().tpd_write_png(img.ctypes.data, 2, 2, None) != 0

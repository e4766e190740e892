"""Frame and settings loading (include/vh_io.h), host only.

Pins: the settings parser against the camera and voxel values of the reference's own
configuration file (scene0220_02/scene0220_02.yaml, read from the reference checkout when it
is present, and a fixture with the same values in the same form); the 16-bit PNG reader
against an independent encoder written here that uses every PNG row filter; the PNG writer
against an independent decoder; pose text parsing; the per-frame directory layout of the
reference's SaveFrame (depth/<id>.png, tcw/<id>.txt).
"""
import os
import struct
import zlib

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_SETTINGS = "/root/reference/scene0220_02/scene0220_02.yaml"

EXPECTED = dict(fx=577.0, fy=577.0, cx=320.0, cy=240.0, width=640, height=480, depth_factor=1000.0,
                max_depth=10.0, origin=(0.5, 0.0, -0.5), vox_size=np.float32(0.036), trunc=np.float32(0.18),
                vox_dim=(200, 200, 200))


def check_settings(vh, path):
    s = vh.load_settings(path)
    assert s["found"] == 0xFFF
    for k, v in EXPECTED.items():
        got = s[k]
        if isinstance(v, tuple):
            assert tuple(np.float32(x) for x in got) == tuple(np.float32(x) for x in v), k
        else:
            assert np.float32(got) == np.float32(v), k


def test_settings_fixture(vh):
    check_settings(vh, os.path.join(GOLDEN, "settings_scannet.yaml"))


@pytest.mark.skipif(not os.path.exists(REF_SETTINGS), reason="reference checkout not present")
def test_settings_of_the_reference_configuration(vh):
    check_settings(vh, REF_SETTINGS)


def test_settings_errors_and_config(vh, tmp_path):
    p = tmp_path / "s.yaml"
    p.write_text("%YAML:1.0\nCamera.fx: abc\n")
    with pytest.raises(vh.VhError) as e:
        vh.load_settings(str(p))
    assert e.value.code == vh.VH_ERR_INVALID_ARG
    p.write_text("Camera.width: 640.5\n")
    with pytest.raises(vh.VhError):
        vh.load_settings(str(p))
    p.write_text("Camera.fx: 500\nVoxel.Origin.x: 1\n")  # partial: origin needs all three
    s = vh.load_settings(str(p))
    assert s["found"] == 1 and "origin" not in s
    with pytest.raises(vh.VhError) as e:
        vh.load_settings(str(tmp_path / "missing.yaml"))
    assert e.value.code == vh.VH_ERR_IO
    with pytest.raises(vh.VhError):  # lacks required keys
        vh.config_from_settings(str(p))
    cfg = vh.config_from_settings(os.path.join(GOLDEN, "settings_scannet.yaml"), log2_buckets=20)
    assert (cfg.width, cfg.height, cfg.fx, cfg.cx, cfg.max_depth) == (640, 480, 577.0, 320.0, 10.0)
    assert cfg.vox_size == np.float32(0.036) and cfg.trunc == np.float32(0.18) and cfg.log2_buckets == 20


# ---------------------------------------------------------------- independent PNG codec
def _chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def encode_png16(img, filters):
    """16-bit greyscale PNG; row y uses filter filters[y % len(filters)] (0..4)."""
    h, w = img.shape
    raw = bytearray()
    prev = bytes(2 * w)
    for y in range(h):
        cur = img[y].astype(">u2").tobytes()
        ft = filters[y % len(filters)]
        out = bytearray()
        for i in range(2 * w):
            a = cur[i - 2] if i >= 2 else 0
            b = prev[i]
            c = prev[i - 2] if i >= 2 else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][ft]
            out.append((cur[i] - pred) & 0xFF)
        raw += bytes([ft]) + out
        prev = cur
    ihdr = struct.pack(">IIBBBBB", w, h, 16, 0, 0, 0, 0)
    return b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"tEXt", b"k\x00v") + \
        _chunk(b"IDAT", zlib.compress(bytes(raw))[:7]) + _chunk(b"IDAT", zlib.compress(bytes(raw))[7:]) + \
        _chunk(b"IEND", b"")


def decode_png16(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    p, idat, w = 8, b"", None
    while p < len(data):
        n = struct.unpack(">I", data[p:p + 4])[0]
        t, body = data[p + 4:p + 8], data[p + 8:p + 8 + n]
        if t == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert (depth, ctype) == (16, 0)
        elif t == b"IDAT":
            idat += body
        p += 12 + n
    raw = zlib.decompress(idat)
    out = np.zeros((h, w), np.uint16)
    prev = bytearray(2 * w)
    for y in range(h):
        row = raw[y * (2 * w + 1):(y + 1) * (2 * w + 1)]
        ft, cur = row[0], bytearray(row[1:])
        for i in range(2 * w):
            a = cur[i - 2] if i >= 2 else 0
            b = prev[i]
            c = prev[i - 2] if i >= 2 else 0
            cur[i] = (cur[i] + [0, a, b, (a + b) >> 1, _paeth(a, b, c)][ft]) & 0xFF
        out[y] = np.frombuffer(bytes(cur), ">u2")
        prev = cur
    return out


def test_png16_reader_all_filters(vh, tmp_path):
    rng = np.random.default_rng(7)
    img = rng.integers(0, 65536, size=(23, 37), dtype=np.uint16)
    img[0, :5] = [0, 65535, 1, 256, 65280]
    img[5:9] = img[4]  # repeated rows (Up filter) and a smooth ramp (Sub / Paeth)
    img[10] = np.arange(37, dtype=np.uint16) * 1000
    path = tmp_path / "d.png"
    path.write_bytes(encode_png16(img, [0, 1, 2, 3, 4]))
    assert np.array_equal(vh.png16_load(str(path)), img)
    path.write_bytes(encode_png16(img, [4, 3]))
    assert np.array_equal(vh.png16_load(str(path)), img)


def test_png16_writer_round_trip(vh, tmp_path):
    rng = np.random.default_rng(8)
    img = rng.integers(0, 65536, size=(48, 64), dtype=np.uint16)
    path = tmp_path / "w.png"
    vh.png16_save(str(path), img)
    assert np.array_equal(decode_png16(path.read_bytes()), img)
    assert np.array_equal(vh.png16_load(str(path)), img)


def test_png16_rejects_other_formats(vh, tmp_path):
    ihdr8 = struct.pack(">IIBBBBB", 4, 4, 8, 0, 0, 0, 0)
    bad8 = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr8) + _chunk(b"IDAT", zlib.compress(bytes(20))) + \
        _chunk(b"IEND", b"")
    (tmp_path / "b8.png").write_bytes(bad8)
    good = bytearray(encode_png16(np.zeros((4, 4), np.uint16), [0]))
    good[-5] ^= 0xFF  # corrupt the IEND CRC
    (tmp_path / "crc.png").write_bytes(bytes(good))
    (tmp_path / "txt.png").write_text("not a png")
    for name in ("b8.png", "crc.png", "txt.png"):
        with pytest.raises(vh.VhError) as e:
            vh.png16_load(str(tmp_path / name))
        assert e.value.code == vh.VH_ERR_INVALID_ARG, name
    with pytest.raises(vh.VhError) as e:
        vh.png16_load(str(tmp_path / "missing.png"))
    assert e.value.code == vh.VH_ERR_IO


def test_pose_text_and_frame_directory(vh, tmp_path):
    T = np.array([[0.25, -1e-3, 3.5e2, 1.0], [0, 1, 0, -2.5], [0, 0, 1, 0.125], [0, 0, 0, 1]], np.float32)
    text = "  0.25 -1e-3\t350.0 1\n0 1 0 -2.5\n0 0 1 0.125\n0 0 0 1\n"
    (tmp_path / "p.txt").write_text(text)
    assert np.array_equal(vh.pose_load(str(tmp_path / "p.txt")), T)
    (tmp_path / "short.txt").write_text("1 2 3")
    with pytest.raises(vh.VhError):
        vh.pose_load(str(tmp_path / "short.txt"))
    # SaveFrame layout: depth/<id>.png + tcw/<id>.txt
    (tmp_path / "depth").mkdir()
    (tmp_path / "tcw").mkdir()
    img = (np.arange(6 * 8, dtype=np.uint16).reshape(6, 8) * 997) % 65535
    vh.png16_save(str(tmp_path / "depth" / "3.png"), img)
    (tmp_path / "tcw" / "3.txt").write_text(text)
    d, P = vh.frame_load(str(tmp_path), 3, 8, 6)
    assert np.array_equal(d, img) and np.array_equal(P, T)
    with pytest.raises(vh.VhError) as e:
        vh.frame_load(str(tmp_path), 4, 8, 6)
    assert e.value.code == vh.VH_ERR_IO
    with pytest.raises(vh.VhError):  # wrong size
        vh.frame_load(str(tmp_path), 3, 9, 6)

"""Headless output formats and the image / exit-direction diff (host code of
libbhtrace.so, no GPU).

The files are decoded here independently of the writer: PPM by its header
and byte layout, PNG chunk by chunk with CRC-32 and the zlib stream checked
by Python's zlib, PFM with its bottom-to-top row order.  The diff statistics
are compared with numpy, and the exit-angle statistics with rotations by
known angles.
"""
import struct
import zlib

import numpy as np
import pytest

import bhtrace as bt


@pytest.fixture(scope="module")
def lib():
    return bt.load()


def rand_rgba(h, w, seed):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 4), dtype=np.uint8)


@pytest.mark.parametrize("h,w", [(1, 1), (3, 7), (54, 96), (300, 256)])
def test_ppm_roundtrip(lib, tmp_path, h, w):
    img = rand_rgba(h, w, h * 1000 + w)
    p = tmp_path / "a.ppm"
    bt.write_image(str(p), img)
    data = p.read_bytes()
    head = f"P6\n{w} {h}\n255\n".encode()
    assert data[:len(head)] == head
    body = np.frombuffer(data[len(head):], np.uint8)
    assert body.size == h * w * 3
    assert np.array_equal(body.reshape(h, w, 3), img[..., :3])


def parse_png(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, w * 4 + 1)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 4)


@pytest.mark.parametrize("h,w", [(1, 1), (5, 3), (270, 480), (200, 123)])
def test_png_roundtrip(lib, tmp_path, h, w):
    # (270, 480) spans several 65535-byte stored blocks
    img = rand_rgba(h, w, h + 7 * w)
    p = tmp_path / "a.png"
    bt.write_image(str(p), img)
    assert np.array_equal(parse_png(p.read_bytes()), img)


def test_pfm_roundtrip(lib, tmp_path):
    h, w = 17, 29
    hdr = np.random.default_rng(3).normal(0, 10, (h, w, 4)).astype(np.float32)
    hdr[0, 0, :3] = [np.inf, -0.0, 1e-40]
    p = tmp_path / "a.pfm"
    bt.write_pfm(str(p), hdr)
    data = p.read_bytes()
    head = f"PF\n{w} {h}\n-1.0\n".encode()
    assert data[:len(head)] == head
    body = np.frombuffer(data[len(head):], "<f4").reshape(h, w, 3)[::-1]
    assert np.array_equal(body.view(np.uint32), hdr[..., :3].view(np.uint32))


def test_bad_arguments(lib, tmp_path):
    img = rand_rgba(4, 4, 0)
    assert lib.bhWritePPM(str(tmp_path / "x.ppm").encode(), img.ctypes.data, 0, 4) == -1
    assert lib.bhWritePNG(None, img.ctypes.data, 4, 4) == -1
    assert lib.bhWritePFM(str(tmp_path / "no_such_dir" / "x.pfm").encode(), img.ctypes.data, 1, 1) == -1
    with pytest.raises(ValueError):
        bt.image_diff(img, rand_rgba(4, 5, 1))


def test_image_diff_against_numpy(lib):
    a = rand_rgba(61, 83, 11)
    rng = np.random.default_rng(12)
    delta = rng.choice([0, 0, 0, 1, -1, 2, 5, 9, 40, 120], size=a.shape)
    b = np.clip(a.astype(int) + delta, 0, 255).astype(np.uint8)
    b[..., 3] = 0                     # alpha is ignored
    st = bt.image_diff(a, b)
    d = np.abs(a[..., :3].astype(int) - b[..., :3].astype(int))
    m = d.max(-1).ravel()
    edges = [0, 1, 2, 3, 4, 8, 16, 64, 256]
    hist = np.histogram(m, bins=edges)[0]
    assert st.pixels == a.shape[0] * a.shape[1]
    assert list(st.maxDiff) == list(d.reshape(-1, 3).max(0))
    assert list(st.hist) == list(hist)
    assert st.fracWithin1 == pytest.approx((m <= 1).mean(), abs=0)
    same = bt.image_diff(a, a)
    assert same.hist[0] == same.pixels and same.fracWithin1 == 1.0


def test_exit_angle_diff_known_rotations(lib):
    n = 4000
    rng = np.random.default_rng(5)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    # rotate each direction by a chosen angle about a perpendicular axis
    ang = 10.0 ** rng.uniform(-6, -1, n)
    k = np.cross(u, rng.normal(size=(n, 3)))
    k /= np.linalg.norm(k, axis=1, keepdims=True)
    v = u * np.cos(ang)[:, None] + np.cross(k, u) * np.sin(ang)[:, None]
    a = np.zeros(n, bt.RECORD_DTYPE)
    b = np.zeros(n, bt.RECORD_DTYPE)
    a["escDir"], b["escDir"] = u * 3.0, v * 0.5       # norms do not matter
    a["flags"], b["flags"] = 1, 1
    a["flags"][:7] = 2                                 # captured in one record only
    a["flags"][7:10], b["flags"][7:10] = 2, 2          # captured in both
    got, st = bt.exit_angle_diff(a, b)
    keep = np.arange(n) >= 10
    assert st.outcomeMismatch == 7 and st.bothEscaped == keep.sum()
    assert (got[~keep] == -1).all()
    ref = ang[keep]
    assert np.allclose(got[keep], ref, rtol=1e-4, atol=2e-7)
    assert st.over1mrad == int((ref > 1e-3).sum())
    assert st.max == pytest.approx(ref.max(), rel=1e-4, abs=2e-7)
    assert st.median == pytest.approx(np.sort(ref)[int(0.5 * (ref.size - 1) + 0.5)], rel=1e-4, abs=2e-7)


# ------------------------------------------------------------------ bhdiff
def test_bhdiff_reads_what_the_library_writes(lib, tmp_path):
    import bhdiff
    img = rand_rgba(37, 53, 21)
    for ext in ("ppm", "png"):
        p = tmp_path / f"a.{ext}"
        bt.write_image(str(p), img)
        got = bhdiff.read_image(str(p))
        want = img.copy()
        if ext == "ppm":
            want[..., 3] = 255
        assert np.array_equal(got, want), ext
    hdr = np.random.default_rng(1).normal(size=(9, 11, 4)).astype(np.float32)
    bt.write_pfm(str(tmp_path / "a.pfm"), hdr)
    assert np.array_equal(bhdiff.read_image(str(tmp_path / "a.pfm")), hdr[..., :3])


@pytest.mark.parametrize("mode", ["RGB", "RGBA"])
def test_bhdiff_png_filters_match_pillow(tmp_path, mode):
    # Pillow's encoder picks adaptive row filters (sub, up, average, paeth)
    Image = pytest.importorskip("PIL.Image")
    import bhdiff
    rng = np.random.default_rng(4)
    y, x = np.mgrid[0:45, 0:70]
    img = np.stack([(x * 3 + y) % 256, (y * 5) % 256, (x * y) % 256, 255 - x], -1).astype(np.uint8)
    img[::7] = rng.integers(0, 256, img[::7].shape, dtype=np.uint8)
    im = Image.fromarray(img[..., :3] if mode == "RGB" else img, mode)
    p = tmp_path / "f.png"
    im.save(p, optimize=True)
    got = bhdiff.read_image(str(p))
    assert np.array_equal(got[..., :3], img[..., :3])
    if mode == "RGBA":
        assert np.array_equal(got[..., 3], img[..., 3])
    # and Pillow decodes the library's own PNG
    q = tmp_path / "g.png"
    bt.write_image(str(q), img)
    assert np.array_equal(np.asarray(Image.open(q).convert("RGBA")), img)


def test_bhdiff_cli(lib, tmp_path, capsys):
    import json
    import bhdiff
    a = rand_rgba(20, 30, 5)
    b = a.copy()
    b[0, :5, 0] ^= 1                  # 5 pixels at 1 LSB
    bt.write_image(str(tmp_path / "a.png"), a)
    bt.write_image(str(tmp_path / "b.ppm"), b)
    assert bhdiff.main([str(tmp_path / "a.png"), str(tmp_path / "b.ppm")]) == 0
    out = json.loads(capsys.readouterr().out)
    assert out["within1"] == 1.0 and out["hist"][1] == 5
    b[1, :2, 2] = a[1, :2, 2] ^ 0x40  # 2 pixels far off: 598/600 within 1 LSB
    bt.write_image(str(tmp_path / "b.ppm"), b)
    assert bhdiff.main([str(tmp_path / "a.png"), str(tmp_path / "b.ppm")]) == 1
    assert bhdiff.main([str(tmp_path / "a.png"), str(tmp_path / "b.ppm"), "--min-within1", "0.99"]) == 0
    capsys.readouterr()
    r = np.zeros(100, bt.RECORD_DTYPE)
    r["flags"] = 1
    r["escDir"] = [0.0, 0.0, 1.0]
    s = r.copy()
    s["escDir"][:3] = [np.sin(2e-3), 0.0, np.cos(2e-3)]
    np.save(tmp_path / "r.npy", r)
    np.save(tmp_path / "s.npy", s)
    assert bhdiff.main(["--records", str(tmp_path / "r.npy"), str(tmp_path / "s.npy")]) == 1
    out = json.loads(capsys.readouterr().out)
    assert out["within"] == pytest.approx(0.97) and out["max_rad"] == pytest.approx(2e-3, rel=1e-3)
    assert bhdiff.main(["--records", str(tmp_path / "r.npy"), str(tmp_path / "s.npy"),
                        "--min-within1", "0.95"]) == 0


def test_bhdiff_png_every_filter_type(tmp_path):
    """Rows filtered with each of the five PNG filter types in turn (encoded
    here from the PNG specification); Pillow checks the encoding."""
    Image = pytest.importorskip("PIL.Image")
    import bhdiff
    h, w, bpp = 25, 19, 4
    img = rand_rgba(h, w, 9)
    rows = img.reshape(h, w * bpp).astype(np.int32)
    out = bytearray()
    prev = np.zeros(w * bpp, np.int32)
    for y in range(h):
        f = y % 5
        cur = rows[y]
        left = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]])
        upleft = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
        if f == 0:
            pred = np.zeros_like(cur)
        elif f == 1:
            pred = left
        elif f == 2:
            pred = prev
        elif f == 3:
            pred = (left + prev) >> 1
        else:
            p = left + prev - upleft
            pa, pb, pc = np.abs(p - left), np.abs(p - prev), np.abs(p - upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
        out.append(f)
        out += ((cur - pred) & 0xFF).astype(np.uint8).tobytes()
        prev = cur

    def chunk(t, b):
        return struct.pack(">I", len(b)) + t + b + struct.pack(">I", zlib.crc32(t + b) & 0xFFFFFFFF)
    png = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
           + chunk(b"IDAT", zlib.compress(bytes(out))) + chunk(b"IEND", b""))
    p = tmp_path / "filters.png"
    p.write_bytes(png)
    assert np.array_equal(np.asarray(Image.open(p)), img)
    assert np.array_equal(bhdiff.read_image(str(p)), img)

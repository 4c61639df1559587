"""Compare two headless frames, or two ray-record dumps, against the parity bar.

    python bhdiff.py A.png B.png          # PPM (P6) or PNG (8-bit RGB/RGBA)
    python bhdiff.py A.pfm B.pfm          # HDR (PFM)
    python bhdiff.py --records A.npy B.npy   # ray records (bhtrace.RECORD_DTYPE)

Frames: per-channel LSB statistics from bhImageDiff; exit status 0 when at
least --min-within1 of the pixels (default 0.999) differ by at most 1 LSB in
every channel.  HDR: largest absolute and relative difference; exit 0 when
the relative difference stays below --hdr-rtol.  Records: exit-direction
angle statistics from bhExitAngleDiff; exit 0 when at least --min-within1 of
the rays escaping in both are within --max-angle rad (default 1e-3) and no
ray escapes in only one.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import struct
import sys
import zlib

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import bhtrace as bt  # noqa: E402


def _tokens(data: bytes, count: int):
    """First `count` whitespace-separated header tokens (PNM/PFM, with
    comments) and the offset just past the single whitespace after them."""
    out, pos = [], 0
    while len(out) < count:
        while data[pos:pos + 1].isspace():
            pos += 1
        if data[pos:pos + 1] == b"#":
            while data[pos:pos + 1] not in (b"\n", b""):
                pos += 1
            continue
        start = pos
        while pos < len(data) and not data[pos:pos + 1].isspace():
            pos += 1
        out.append(data[start:pos].decode())
    return out, pos + 1


def read_ppm(data: bytes) -> np.ndarray:
    (magic, w, h, mx), off = _tokens(data, 4)
    if magic != "P6" or int(mx) != 255:
        raise ValueError("only 8-bit binary PPM (P6) is read")
    w, h = int(w), int(h)
    rgb = np.frombuffer(data, np.uint8, w * h * 3, off).reshape(h, w, 3)
    out = np.full((h, w, 4), 255, np.uint8)
    out[..., :3] = rgb
    return out


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def read_png(data: bytes) -> np.ndarray:
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, hdr = 8, [], None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if typ == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat.append(body)
        pos += 12 + n
    w, h, depth, ctype, _, _, inter = hdr
    if depth != 8 or ctype not in (2, 6) or inter != 0:
        raise ValueError("only 8-bit non-interlaced RGB/RGBA PNG is read")
    bpp = 4 if ctype == 6 else 3
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8).reshape(h, w * bpp + 1)
    img = np.zeros((h, w * bpp), np.int32)
    prev = np.zeros(w * bpp, np.int32)
    for y in range(h):
        f, line = raw[y, 0], raw[y, 1:].astype(np.int32)
        if f == 0:
            cur = line
        elif f == 2:
            cur = (line + prev) & 0xFF
        else:
            cur = np.zeros_like(line)
            for i in range(line.size):
                a = cur[i - bpp] if i >= bpp else 0
                c = prev[i - bpp] if i >= bpp else 0
                b = prev[i]
                pred = a if f == 1 else ((a + b) >> 1 if f == 3 else _paeth(a, b, c))
                cur[i] = (line[i] + pred) & 0xFF
        img[y] = cur
        prev = cur
    img = img.astype(np.uint8).reshape(h, w, bpp)
    if bpp == 4:
        return img
    out = np.full((h, w, 4), 255, np.uint8)
    out[..., :3] = img
    return out


def read_pfm(data: bytes) -> np.ndarray:
    (magic, w, h, scale), off = _tokens(data, 4)
    if magic != "PF":
        raise ValueError("only colour PFM (PF) is read")
    w, h, scale = int(w), int(h), float(scale)
    dt = "<f4" if scale < 0 else ">f4"
    return np.frombuffer(data, dt, w * h * 3, off).reshape(h, w, 3)[::-1].astype(np.float32)


def read_image(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] == b"\x89PNG\r\n\x1a\n":
        return read_png(data)
    if data[:2] == b"P6":
        return read_ppm(data)
    if data[:2] == b"PF":
        return read_pfm(data)
    raise ValueError(f"{path}: unknown image format")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("a")
    ap.add_argument("b")
    ap.add_argument("--records", action="store_true", help="inputs are .npy ray-record arrays")
    ap.add_argument("--min-within1", type=float, default=0.999)
    ap.add_argument("--max-angle", type=float, default=1e-3)
    ap.add_argument("--hdr-rtol", type=float, default=1e-3)
    args = ap.parse_args(argv)
    if args.records:
        ra, rb = np.load(args.a), np.load(args.b)
        ang, st = bt.exit_angle_diff(ra, rb)
        within = float((ang[ang >= 0] <= args.max_angle).mean()) if st.bothEscaped else 1.0
        res = {"kind": "records", "both_escaped": st.bothEscaped, "outcome_mismatch": st.outcomeMismatch,
               "median_rad": st.median, "p999_rad": st.p999, "max_rad": st.max, "within": within}
        ok = st.outcomeMismatch == 0 and within >= args.min_within1
    else:
        A, B = read_image(args.a), read_image(args.b)
        if A.shape != B.shape:
            print(json.dumps({"error": f"shapes differ: {A.shape} vs {B.shape}"}))
            return 1
        if A.dtype == np.float32:
            d = np.abs(A.astype(np.float64) - B)
            rel = d / np.maximum(np.maximum(np.abs(A), np.abs(B)), 1e-30)
            res = {"kind": "hdr", "pixels": int(A.shape[0] * A.shape[1]), "max_abs": float(d.max()),
                   "max_rel": float(rel.max()), "nonfinite": int((~np.isfinite(A)).sum() + (~np.isfinite(B)).sum())}
            ok = res["max_rel"] <= args.hdr_rtol and res["nonfinite"] == 0
        else:
            st = bt.image_diff(A, B)
            res = {"kind": "ldr", "pixels": st.pixels, "max_diff": list(st.maxDiff), "hist": list(st.hist),
                   "within1": st.fracWithin1}
            ok = st.fracWithin1 >= args.min_within1
    res["pass"] = bool(ok)
    print(json.dumps(res))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())

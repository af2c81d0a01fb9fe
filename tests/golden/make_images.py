"""Write tests/golden/images.npz: small PNG and Netpbm files made here with
Python's zlib (every PNG colour type and bit depth the reference reads, all
five scanline filters, stored / fixed-Huffman / dynamic-Huffman DEFLATE
blocks) and the reference's read_image outputs on them (stereonet.py:
236-324, :3190-3613).

Run in the development container only (the reference checkout is not present
anywhere else):

    python tests/golden/make_images.py /root/reference

The reference is loaded the way its own harnesses load it (exec of the
source text); only read_image is called.  Outputs are compressed .npz (no
pickles).
"""

import os
import struct
import sys
import tempfile
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_reference(ref_root):
    src = open(os.path.join(ref_root, "stereonet.py")).read().replace(
        "if __name__ == '__main__':\n    run_demo()\n", "")
    SN = {"__name__": "stereonet_reference"}
    exec(src, SN)
    return SN


def _chunk(t, body):
    return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _filter(rows, bpp, types):
    """Apply PNG filter types[y] to each raw row (forward direction)."""
    out = bytearray()
    prev = bytes(len(rows[0]))
    for y, row in enumerate(rows):
        t = types[y]
        out.append(t)
        for i in range(len(row)):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i] if y > 0 else 0
            c = prev[i - bpp] if (y > 0 and i >= bpp) else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][t]
            out.append((row[i] - pred) & 0xFF)
        prev = row
    return bytes(out)


def make_png(rng, W, H, ctype, depth, comp, palette_n=None):
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    maxv = (1 << depth) - 1
    if ctype == 3:
        maxv = min(maxv, palette_n - 1)
    samples = rng.integers(0, maxv + 1, size=(H, W, ch))
    rows = []
    for y in range(H):
        if depth == 8:
            rows.append(bytes(samples[y].astype(np.uint8).ravel()))
        elif depth == 16:
            rows.append(samples[y].astype(">u2").tobytes())
        else:
            bits = "".join(format(int(v), "0%db" % depth) for v in samples[y, :, 0])
            bits += "0" * (-len(bits) % 8)
            rows.append(bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8)))
    bpp = max(1, ch * depth // 8)
    types = [int(t) for t in rng.integers(0, 5, size=H)]
    raw = _filter(rows, bpp, types)
    if comp == "stored":
        z = zlib.compress(raw, 0)
    elif comp == "fixed":
        co = zlib.compressobj(9, zlib.DEFLATED, 15, 9, zlib.Z_FIXED)
        z = co.compress(raw) + co.flush()
    elif comp == "rle":
        co = zlib.compressobj(9, zlib.DEFLATED, 15, 9, zlib.Z_RLE)
        z = co.compress(raw) + co.flush()
    else:
        z = zlib.compress(raw, 9)
    body = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, depth, ctype, 0, 0, 0))
    if ctype == 3:
        body += _chunk(b"PLTE", bytes(rng.integers(0, 256, size=3 * palette_n).astype(np.uint8)))
    body += _chunk(b"tEXt", b"Comment\x00made by make_images.py")
    # IDAT split in two chunks
    h = len(z) // 2
    body += _chunk(b"IDAT", z[:h]) + _chunk(b"IDAT", z[h:]) + _chunk(b"IEND", b"")
    return body


def make_pnm(rng, W, H, magic, maxval):
    ch = 3 if magic in ("P3", "P6") else 1
    s = rng.integers(0, maxval + 1, size=W * H * ch)
    head = ("%s\n# a comment line\n%d %d\n# another\n%d\n" % (magic, W, H, maxval)).encode()
    if magic in ("P2", "P3"):
        return head + (" ".join(str(int(v)) for v in s) + "\n").encode()
    if maxval < 256:
        return head + bytes(s.astype(np.uint8))
    return head + s.astype(">u2").tobytes()


def cases(rng):
    out = []
    for ctype, depths in ((0, (1, 2, 4, 8, 16)), (2, (8, 16)), (3, (1, 2, 4, 8)), (4, (8, 16)),
                          (6, (8, 16))):
        for depth in depths:
            for comp in ("stored", "fixed", "dynamic", "rle"):
                W, H = int(rng.integers(1, 40)), int(rng.integers(1, 24))
                pal = int(rng.integers(1, (1 << min(depth, 8)) + 1)) if ctype == 3 else None
                out.append(("png_c%d_d%d_%s" % (ctype, depth, comp),
                            make_png(rng, W, H, ctype, depth, comp, pal)))
    out.append(("png_big_rgb", make_png(rng, 96, 48, 2, 8, "dynamic")))
    for magic in ("P2", "P3", "P5", "P6"):
        for maxval in (255, 1000, 65535):
            W, H = int(rng.integers(1, 30)), int(rng.integers(1, 20))
            out.append(("pnm_%s_%d" % (magic, maxval), make_pnm(rng, W, H, magic, maxval)))
    return out


def main(ref_root):
    SN = load_reference(ref_root)
    rng = np.random.default_rng(2024)
    data = {}
    names = []
    with tempfile.TemporaryDirectory() as td:
        for name, blob in cases(rng):
            path = os.path.join(td, name)
            with open(path, "wb") as f:
                f.write(blob)
            gray, rgb, W, H = SN["read_image"](path)
            data["file_" + name] = np.frombuffer(blob, dtype=np.uint8)
            data["gray_" + name] = np.array(gray, dtype=np.float64)
            data["rgb_" + name] = np.array(rgb, dtype=np.float64)
            data["wh_" + name] = np.array([W, H], dtype=np.int32)
            names.append(name)
    data["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "images.npz"), **data)
    print("wrote images.npz (%d files)" % len(names))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

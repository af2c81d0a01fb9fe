"""Image input: read_pnm, read_png and read_image (stereonet.py:236-324,
:3499-3613) with the reference's names and return contract -- (gray, rgb,
W, H), gray W*H floats and rgb 3*W*H interleaved floats in [0, 1] -- plus
read_image_gpu, which leaves both planes on the GPU for the dense path.

The file is decoded on the host by libstereo_mi.so's own PNG (DEFLATE,
scanline filters) and Netpbm decoder into integer samples; one HIP kernel
turns them into the float64 planes with the reference's arithmetic.
"""

import ctypes

import numpy as np
import torch

from . import _lib as L

_I, _P, _SZ, _LL = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_longlong
L.extra_signatures({
    "sv_image_size": (_I, [_P, _SZ, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "sv_image_decode": (_I, [_P, _SZ, _P, _SZ, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "sv_image_expand": (_I, [_P, _LL, _I, _I, _P, _P, _P]),
})

_PNG_SIG = b"\x89PNG\r\n\x1a\n"


def _size(buf, path):
    W, H = ctypes.c_int(0), ctypes.c_int(0)
    lib = L.lib()
    if lib.sv_image_size(buf.ctypes.data_as(_P), buf.size, ctypes.byref(W), ctypes.byref(H)):
        raise ValueError("%s: %s" % (path, lib.sv_last_error().decode(errors="replace")))
    return W.value, H.value


def _decode_into(buf, s, path):
    """Decode into the uint16 array s (3 samples per pixel); returns the
    scale rule.  ctypes drops the GIL for the call, so threads overlap."""
    mode, maxval = ctypes.c_int(0), ctypes.c_int(0)
    lib = L.lib()
    if lib.sv_image_decode(buf.ctypes.data_as(_P), buf.size, s.ctypes.data_as(_P), s.size,
                           ctypes.byref(mode), ctypes.byref(maxval)):
        raise ValueError("%s: %s" % (path, lib.sv_last_error().decode(errors="replace")))
    return mode.value, maxval.value


def _decode_bytes(data, path):
    buf = np.frombuffer(data, dtype=np.uint8)
    W, H = _size(buf, path)
    s = np.empty(3 * W * H, dtype=np.uint16)
    mode, maxval = _decode_into(buf, s, path)
    return s, mode, maxval, W, H


def _expand(s, mode, maxval, W, H):
    dev = L.device()
    t = torch.from_numpy(s.view(np.int16)).to(dev)  # 16-bit samples, bit pattern kept
    gray = torch.empty(W * H, dtype=torch.float64, device=dev)
    rgb = torch.empty(3 * W * H, dtype=torch.float64, device=dev)
    L.check(L.lib().sv_image_expand(L.ptr(t), W * H, mode, maxval, L.ptr(gray), L.ptr(rgb),
                                    L.stream()))
    return gray, rgb


def _read(path):
    with open(path, "rb") as f:
        return f.read()


def read_image_gpu(path):
    """Decode a PNG or Netpbm file (chosen by its magic bytes) and return
    (gray [H][W], rgb [H][W][3], W, H) as float64 CUDA tensors."""
    data = _read(path) if not isinstance(path, (bytes, bytearray)) else bytes(path)
    s, mode, maxval, W, H = _decode_bytes(data, path if isinstance(path, str) else "<bytes>")
    gray, rgb = _expand(s, mode, maxval, W, H)
    return gray.reshape(H, W), rgb.reshape(H, W, 3), W, H


def read_images_gpu(files, threads=8):
    """Decode N files of one size (paths or bytes; PNG or Netpbm each) on up
    to `threads` host threads into one pinned staging buffer, upload it with
    one copy and expand every image on the GPU.  Returns (gray [N][H][W],
    rgb [N][H][W][3], W, H), float64 CUDA tensors equal image by image to
    read_image_gpu's."""
    from concurrent.futures import ThreadPoolExecutor
    datas = [bytes(f) if isinstance(f, (bytes, bytearray)) else _read(f) for f in files]
    names = [f if isinstance(f, str) else "<bytes %d>" % i for i, f in enumerate(files)]
    bufs = [np.frombuffer(d, dtype=np.uint8) for d in datas]
    if not bufs:
        raise ValueError("no images")
    dims = [_size(b, n) for b, n in zip(bufs, names)]
    W, H = dims[0]
    for (w, h), n in zip(dims, names):
        if (w, h) != (W, H):
            raise ValueError("images are not all the same size: %dx%d and %dx%d (%s)"
                             % (W, H, w, h, n))
    Nn, npix = len(bufs), W * H
    stage = torch.empty((Nn, 3 * npix), dtype=torch.int16, pin_memory=True)
    arr = stage.numpy().view(np.uint16)
    with ThreadPoolExecutor(max_workers=max(1, min(threads, Nn))) as ex:
        rules = list(ex.map(lambda i: _decode_into(bufs[i], arr[i], names[i]), range(Nn)))
    dev = L.device()
    t = stage.to(dev, non_blocking=True)
    gray = torch.empty((Nn, npix), dtype=torch.float64, device=dev)
    rgb = torch.empty((Nn, 3 * npix), dtype=torch.float64, device=dev)
    i = 0
    while i < Nn:  # one launch per run of images sharing a scale rule
        j = i + 1
        while j < Nn and rules[j] == rules[i]:
            j += 1
        L.check(L.lib().sv_image_expand(L.ptr(t[i]), (j - i) * npix, rules[i][0], rules[i][1],
                                        L.ptr(gray[i]), L.ptr(rgb[i]), L.stream()))
        i = j
    return gray.reshape(Nn, H, W), rgb.reshape(Nn, H, W, 3), W, H


def _lists(path, data):
    s, mode, maxval, W, H = _decode_bytes(data, path)
    gray, rgb = _expand(s, mode, maxval, W, H)
    return gray.cpu().numpy().tolist(), rgb.cpu().numpy().tolist(), W, H


def read_pnm(path):
    """read_pnm, stereonet.py:262-324: a PGM or PPM file (P2/P3/P5/P6)."""
    data = _read(path)
    if data[0:2] not in (b"P2", b"P3", b"P5", b"P6"):
        raise ValueError("not a PGM/PPM file: " + str(path))
    return _lists(path, data)


def read_png(path):
    """read_png, stereonet.py:3499-3601: a non-interlaced PNG."""
    data = _read(path)
    if data[0:8] != _PNG_SIG:
        raise ValueError("not a PNG file: " + str(path))
    return _lists(path, data)


def read_image(path):
    """read_image, stereonet.py:3604-3613: PNG or Netpbm by the magic bytes."""
    data = _read(path)
    if data[0:8] == _PNG_SIG:
        return _lists(path, data)
    if data[0:2] not in (b"P2", b"P3", b"P5", b"P6"):
        raise ValueError("not a PGM/PPM file: " + str(path))
    return _lists(path, data)


def _quantise(v, lo, hi):
    """write_pgm's mapping: int(255 (v - lo) / span + 0.5) clamped to 0..255."""
    t = np.trunc(255.0 * (v - lo) / (hi - lo if hi - lo > 1.0e-12 else 1.0) + 0.5)
    return np.clip(t, 0, 255).astype(np.uint8)


def write_pgm(path, gray, width, height, lo=None, hi=None):
    """write_pgm, stereonet.py:326-355: binary P5, scaled from [lo, hi]
    (the values' own range when not given) to 0..255."""
    v = gray.detach().double().cpu().numpy() if isinstance(gray, torch.Tensor) else \
        np.asarray(gray, dtype=np.float64)
    v = v.ravel()
    lo = float(v.min()) if lo is None else lo
    hi = float(v.max()) if hi is None else hi
    with open(path, "wb") as f:
        f.write(("P5\n%d %d\n255\n" % (width, height)).encode("ascii"))
        f.write(_quantise(v, lo, hi).tobytes())


def write_ppm(path, rgb, width, height):
    """write_ppm, stereonet.py:358-372: binary P6 from interleaved [0, 1]
    floats, int(255 v + 0.5) clamped."""
    v = rgb.detach().double().cpu().numpy() if isinstance(rgb, torch.Tensor) else \
        np.asarray(rgb, dtype=np.float64)
    t = np.clip(np.trunc(255.0 * v.ravel() + 0.5), 0, 255).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(("P6\n%d %d\n255\n" % (width, height)).encode("ascii"))
        f.write(t.tobytes())

"""ctypes binding of libstereo_mi.so (declarations: include/stereo_mi.h).

The product path has no CPU fallback: if the library is not built, or no
ROCm GPU is visible, every compute call raises.
"""

import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstereo_mi.so")

SV_OK, SV_EINVAL, SV_EUNSUPPORTED, SV_EHIP, SV_EWORKSPACE = 0, -1, -2, -3, -4
SV_U8, SV_F32, SV_F64 = 0, 1, 2

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_F = ctypes.c_float
_SZ = ctypes.c_size_t
_PSZ = ctypes.POINTER(ctypes.c_size_t)
_L = ctypes.c_longlong

# name -> (restype, argtypes); mirrors include/*.h
SIGNATURES = {
    "sv_last_error": (ctypes.c_char_p, []),
    "sv_version": (_I, []),
    "sv_census": (_I, [_P, _I, _I, _I, _I, _I, _I, _P, _P]),
    "sv_band_costs": (_I, [_P, _P, _P, _P, _I, _P, _I, _I, _I, _I, _D, _D, _D, _P, _D, _D, _P, _P]),
    "sv_aggregate_path_workspace_bytes": (_I, [_I, _I, _I, _I, _PSZ]),
    "sv_aggregate_path": (_I, [_P, _P, _I, _I, _I, _I, _D, _D, _I, _I, _P, _P, _SZ, _P]),
    "sv_refine_workspace_bytes": (_I, [_I, _I, _I, _I, _I, _D, _D, _D, _D, _I, _I, _I, _D, _D, _PSZ]),
    "sv_refine_fullres": (_I, [_P, _P, _I, _P, _I, _I, _I, _I, _D, _D, _D, _D, _I, _D, _I, _I,
                               _D, _D, _P, _P, _P, _SZ, _P]),
    "sv_full_range_workspace_bytes": (_I, [_I, _I, _I, _I, _I, _D, _D, _I, _I, _I, _PSZ]),
    "sv_full_range_fast": (_I, [_I, _I, _I, _I, _D, _D, _I, _I, _I, ctypes.POINTER(_I)]),
    "sv_full_range": (_I, [_I, _P, _P, _I, _I, _I, _I, _I, _P, _D, _D, _I, _I, _I, _D, _P, _P, _P,
                           _P, _SZ, _P]),
    "sv_left_to_right": (_I, [_P, _I, _I, _I, _P, _P]),
    "sv_lr_refine": (_I, [_P, _P, _I, _P, _I, _I, _I, _D, _I, _D, _D, _D, _D, _I, _D, _I, _I,
                          _D, _D, _P, _P, _P, _SZ, _P]),
    "sv_median_filter": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _P]),
    "sv_profile_enable": (_I, [_I]),
    "sv_tune_set": (_I, [_I, _I]),
    "sv_tune_get": (_I, [_I, ctypes.POINTER(_I)]),
    "sv_profile_read": (_I, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I)]),
    "sv_refine_disparity": (_I, [_P, _P, _I, _P, _I, _I, _I, _I, _D, _D, _I, _I, _D, _P, _P, _P,
                                 _SZ, _P]),
    "sv_refine_disparity_noalloc": (_I, [_P, _P, _I, _P, _I, _I, _I, _I, _D, _D, _I, _I, _D, _P,
                                         _P, _P]),
}

_lib = None


def extra_signatures(table):
    """Register further entry points (used by the network module)."""
    SIGNATURES.update(table)
    if _lib is not None:
        _declare(_lib, table)


def _declare(l, table):
    for name, (res, args) in table.items():
        f = getattr(l, name)
        f.restype = res
        f.argtypes = args


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libstereo_mi.so is not built; run "
                              "python from-scratch-stereo-vision_amd/build.py")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib, SIGNATURES)
    return _lib


def check(rc):
    if rc == SV_OK:
        return
    msg = lib().sv_last_error().decode(errors="replace")
    if rc == SV_EINVAL:
        raise ValueError(msg)
    if rc == SV_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError("libstereo_mi error %d: %s" % (rc, msg))


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("libstereo_mi needs a ROCm GPU; none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class _Workspace:
    """One growing scratch buffer per device and stream; kernels queued on a
    stream use its buffer in order, so reuse across calls on one stream is
    safe, and calls on different streams may run concurrently."""

    def __init__(self):
        self.bufs = {}

    def get(self, nbytes):
        dev = device()
        key = (dev.index, torch.cuda.current_stream(dev).cuda_stream)
        buf = self.bufs.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=dev)
            self.bufs[key] = buf
        return buf


workspace = _Workspace()

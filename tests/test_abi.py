"""CPU checks of the C ABI: the library loads without a GPU and exports
every entry point include/*.h declares; the ctypes table matches them.
No compute call is made."""

import ctypes
import glob
import importlib
import os
import re

from conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(sv_\w+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert {"sv_census", "sv_refine_fullres", "sv_lr_refine", "sv_refine_disparity",
            "sv_left_to_right", "sv_median_filter", "sv_last_error"} <= names


def test_library_exports_every_declared_symbol():
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_covers_header():
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    importlib.import_module("from-scratch-stereo-vision_amd")  # registers every module's table
    assert _declared() <= set(L.SIGNATURES), sorted(_declared() - set(L.SIGNATURES))


def test_version_and_error_text_without_gpu():
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    lib = L.lib()
    assert lib.sv_version() >= 1
    assert isinstance(lib.sv_last_error(), bytes)

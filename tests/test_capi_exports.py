"""The C-ABI library loads and exports every function include/gpu_masstree_c.h declares
(CPU-only: no compute call is made)."""
import ctypes
import os
import re

from indexingpu_amd import build as B

HEADER = os.path.join(B.REPO_DIR, "include", "gpu_masstree_c.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gmt_[a-z_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_functions()
    for want in ("gmt_create", "gmt_destroy", "gmt_find", "gmt_insert", "gmt_update",
                 "gmt_erase", "gmt_scan", "gmt_mixed_batch"):
        assert want in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(B.LIB_PATH):
        B.build()
    lib = ctypes.CDLL(B.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_null_handle_is_rejected_without_gpu():
    if not os.path.exists(B.LIB_PATH):
        B.build()
    lib = ctypes.CDLL(B.LIB_PATH)
    lib.gmt_destroy.argtypes = [ctypes.c_void_p]
    assert lib.gmt_destroy(None) != 0
    lib.gmt_last_error_message.restype = ctypes.c_char_p
    assert b"null" in lib.gmt_last_error_message()

"""CPU checks of the C-ABI library: it loads, exports every function declared
in include/mdt_gpu.h, and fails loudly (no CPU fallback) without a GPU."""
import ctypes as C
import os
import re

import pytest

from md_tian2_amd import mdt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "mdt_gpu.h")).read()
    return sorted(set(re.findall(r"^(?:int|double)\s+(mdt_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_the_binding_symbols():
    assert _declared() == sorted(mdt.SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(mdt.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_errors_without_device():
    L = mdt.lib()
    assert L.mdt_version() >= 100
    if mdt.device_count() == 0:
        with pytest.raises(mdt.MdtError, match="no GPU"):
            mdt.Context(0)


def test_null_arguments_rejected():
    L = mdt.lib()
    assert L.mdt_create(0, None) == -1
    buf = C.create_string_buffer(256)
    n = L.mdt_last_error(buf, 256)
    assert n > 0 and b"NULL" in buf.value

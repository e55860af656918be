"""The C-ABI library loads, exports every entry point of include/lipjoin.h,
and its struct layouts match the ctypes mirror.  No compute calls (no GPU)."""
import ctypes as C
import os
import re

import pytest

from lipjoin import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "lipjoin.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lj_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert sorted(_abi.SIGNATURES) == names          # the binding covers the whole header


def test_struct_layouts_match():
    assert _abi.check_layout() == []
    assert _abi.load().lj_abi_version() == 5


def test_status_strings():
    lib = _abi.load()
    assert lib.lj_status_str(0) == b"ok"
    assert lib.lj_status_str(6).startswith(b"key outside")


def test_create_without_gpu_reports_an_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _abi.load()
    h = C.c_void_p()
    st = lib.lj_create(0, 0, None, C.byref(h))
    assert st == _abi.EHIP and not h.value
    assert b"device" in lib.lj_last_error(None)


def test_host_generator_argument_checks():
    lib = _abi.load()
    assert lib.lj_gen_host(0, 1, 1, 5, 4, 0, 10, None) == _abi.EINVAL     # hi < lo
    assert lib.lj_gen_host(9, 1, 1, 0, 4, 0, 10, None) == _abi.EINVAL     # unknown kind


def test_tune_model_argument_checks():
    lib = _abi.load()
    m = _abi.TuneModel(scan_l2=2.0, scan_big=3.0, probe_extra=0.2, line=1.0, gather=6.0, scan_mod=0.5,
                       scan_blocked=0.7, gather_queue=8.0, gather_lds=9.0,
                       gather_queue_lds=14.0)
    assert lib.lj_tune_set_model(None, C.byref(m)) == _abi.EINVAL
    assert lib.lj_tune_get_model(None, C.byref(m)) == _abi.EINVAL

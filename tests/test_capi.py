"""The C ABI library loads without a GPU and exports every declared entry point."""

import ctypes
import re

import pytest

from conftest import ROOT

HEADER = ROOT / 'include' / 'fresnel_hip.h'
LIB = ROOT / 'fresnel_amd' / 'libfresnel_hip.so'


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(fh_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert 'fh_tracer_render' in names and 'fh_batch_render' in names
    assert len(names) > 50


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_without_a_gpu_are_statuses():
    """Calls that need a GPU return a status and a message instead of crashing."""
    lib = ctypes.CDLL(str(LIB))
    lib.fh_last_error.restype = ctypes.c_char_p
    n = ctypes.c_int(-1)
    assert lib.fh_get_num_available_devices(ctypes.byref(n)) == 0
    if n.value > 0:
        pytest.skip("a GPU is visible")
    dev = ctypes.c_void_p()
    status = lib.fh_device_create(-1, ctypes.byref(dev))
    assert status != 0
    assert b"GPU" in lib.fh_last_error()
    # null handles are rejected
    assert lib.fh_scene_create(None, ctypes.byref(ctypes.c_void_p())) != 0
    assert b"null" in lib.fh_last_error()

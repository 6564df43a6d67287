"""The C-ABI libraries load and export every symbol their headers declare;
without a GPU the tracking entry points fail loudly (no CPU fallback)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phc_[a-z0-9_]+)\s*\(", text)))


def test_hip_library_exports_header():
    from polyhomocont_amd import device
    lib = device.lib()
    names = declared("phc.h")
    assert "phc_track_paths" in names and len(names) > 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_library_exports_header():
    from polyhomocont_amd import host
    lib = host.library().lib
    names = declared("polyhomocont.h")
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_wire_library_exports_header():
    from polyhomocont_amd import wire
    lib = wire.lib()
    names = declared("phc_wire.h")
    assert "phc_wire_decode_request" in names and len(names) == 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_default_options_match_reference_defaults():
    from polyhomocont_amd import abi, device
    o = abi.SolveOptions()
    device.lib().phc_default_options(ctypes.byref(o))
    want = abi.default_options()
    assert bytes(o) == bytes(want)


def test_no_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from polyhomocont_amd import device
    with pytest.raises(device.DeviceError):
        device.Context(0)

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


# ctypes mirrors of the headers' structs: (C type, ctypes class, header)
_STRUCTS = [
    ("phc_tracker_options", "TrackerOptions", "phc.h"),
    ("phc_endgame_options", "EndgameOptions", "phc.h"),
    ("phc_solve_options", "SolveOptions", "phc.h"),
    ("phc_system", "System", "phc.h"),
    ("phc_homotopy", "Homotopy", "phc.h"),
    ("phc_endpoints", "Endpoints", "phc.h"),
    ("phc_server_config", "ServerConfig", "phc.h"),
    ("phc_user_system", "UserSystem", "polyhomocont.h"),
    ("phc_solve_summary", "SolveSummary", "polyhomocont.h"),
    ("phc_path_results", "PathResults", "polyhomocont.h"),
]


def test_ctypes_structs_match_header_layout(tmp_path):
    """Every ctypes struct has the size and field offsets the C compiler
    gives the header's struct (fields in the same order)."""
    import shutil
    import subprocess
    from polyhomocont_amd import abi
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    lines = ["#include <stddef.h>", "#include <stdio.h>",
             '#include "phc.h"', '#include "polyhomocont.h"', "int main(void) {"]
    for cname, pyname, _ in _STRUCTS:
        lines.append(f'  printf("{pyname} size %zu\\n", sizeof({cname}));')
        for field, _t in getattr(abi, pyname)._fields_:
            lines.append(f'  printf("{pyname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for _, pyname, _ in _STRUCTS:
        cls = getattr(abi, pyname)
        assert got[(pyname, "size")] == ctypes.sizeof(cls), pyname
        for field, _t in cls._fields_:
            assert got[(pyname, field)] == getattr(cls, field).offset, (pyname, field)

"""Register and scratch budget of the built kernels (CPU: reads the gfx950
code object inside polyhomocont_amd/lib/libphc_hip.so).

The grid kernels hold a path's state in LDS and keep two wavefronts per
SIMD: their VGPRs plus AGPRs fit 256 (512 registers per lane per SIMD).
No kernel spills to scratch memory except the 16-lane grid kernel reading
its program from global memory, which is held to two wavefronts with a few
spilled registers (at most 64 bytes per lane). A change that silently
pushes a grid kernel to one wavefront per SIMD halves its throughput, so it
fails here.
"""
SPILLING = "_ZN3phc12track_kernelILi16ELb0ELi2EEEvNS_6LaunchE"
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "polyhomocont_amd", "lib", "libphc_hip.so")
LLVM = "/opt/rocm/llvm/bin"


def kernel_notes(tmp_path):
    objcopy = shutil.which("objcopy")
    bundler = os.path.join(LLVM, "clang-offload-bundler")
    readelf = os.path.join(LLVM, "llvm-readelf")
    if not (os.path.exists(LIB) and objcopy and os.path.exists(bundler)
            and os.path.exists(readelf)):
        pytest.skip("library or LLVM tools not available")
    fat = tmp_path / "fat.bin"
    dev = tmp_path / "dev.co"
    subprocess.run([objcopy, f"--dump-section=.hip_fatbin={fat}", LIB,
                    str(tmp_path / "copy.so")], check=True)
    subprocess.run([bundler, "--unbundle", "--type=o",
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                    f"--input={fat}", f"--output={dev}"], check=True)
    text = subprocess.run([readelf, "--notes", str(dev)], check=True,
                          capture_output=True, text=True).stdout
    kernels, cur = [], {}
    for line in text.splitlines():
        m = re.match(r"\s*-?\s*\.(\w+):\s+(\S+)", line)
        if not m:
            continue
        key, val = m.groups()
        if key == "agpr_count" and cur:
            kernels.append(cur)
            cur = {}
        cur[key] = val
    if cur:
        kernels.append(cur)
    return [k for k in kernels if "track_kernel" in k.get("name", "")]


def test_grid_kernels_two_waves_per_simd_and_no_scratch(tmp_path):
    ks = kernel_notes(tmp_path)
    assert len(ks) >= 12, [k.get("name") for k in ks]
    grid = 0
    for k in ks:
        name = k["name"]
        if name == SPILLING:
            assert int(k["private_segment_fixed_size"]) <= 64, name
            assert int(k["vgpr_spill_count"]) <= 16, name
        else:
            assert int(k["private_segment_fixed_size"]) == 0, name
            assert int(k["vgpr_spill_count"]) == 0, name
        # track_kernel<G, STAGED, MODE>: MODE 2 is the grid mapping
        if name.endswith("ELi2EEEvNS_6LaunchE"):
            grid += 1
            regs = int(k["vgpr_count"]) + int(k["agpr_count"])
            assert regs <= 256, (name, regs)
    assert grid == 6

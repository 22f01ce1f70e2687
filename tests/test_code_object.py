"""Checks on the gfx950 code object of the built libholo.so (CPU only: it disassembles the
library's own device code with the ROCm LLVM tools).

A global load under a per-thread branch is followed by ``s_waitcnt vmcnt(0)`` in the generated
code, which serialises a thread's loads (the background-mode kernel once ran 3x slower for it).
The hot kernels must keep their loads in flight together: no full wait for memory may separate
two of their loads that are issued close together.
"""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "particleholography.jl_amd" / "lib" / "libholo.so"
LLVM = Path("/opt/rocm/llvm/bin")

# mangled-name prefixes of the kernels on the measured paths
HOT = {
    "pass A, 1024": "_ZN4holo12k_recon_colsILi16ELi1024ELi0E",
    "pass B, 1024": "_ZN4holo17k_recon_rows_projILi16ELi1024E",
    "pass B, 1280": "_ZN4holo17k_recon_rows_projILi20ELi1280E",
    "N0f8 store, 1024": "_ZN4holo18k_recon_rows_storeILi16ELi1024ELi0ELb0E",
    "GS columns, 1024": "_ZN4holo11k_prop_colsILi8ELi1024ELi2E",
    "GS rows, 1024": "_ZN4holo9k_gs_rowsILi4ELi1024E",
    "GS rows, 2048": "_ZN4holo9k_gs_rowsILi8ELi2048E",
    "background mode": "_ZN4holo17k_background_modeILb1E",
    "background mode, 8-bit counters": "_ZN4holo18k_background_mode8ILi1E",
    "background mode, 8-bit counters, buffer loads": "_ZN4holo18k_background_mode8ILi2E",
    "ingest with background": "_ZN4holo11k_ingest_u8ILb1ELb1E",
}


@pytest.fixture(scope="module")
def disassembly(tmp_path_factory):
    tools = [LLVM / t for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-objdump")]
    if not LIB.exists() or not all(t.exists() for t in tools):
        pytest.skip("libholo.so or the ROCm LLVM tools are missing")
    d = tmp_path_factory.mktemp("co")
    fat, co = d / "fat.bin", d / "gfx950.co"
    subprocess.run([str(tools[0]), f"--dump-section=.hip_fatbin={fat}", str(LIB), str(d / "stripped.so")],
                   check=True, capture_output=True)
    subprocess.run([str(tools[1]), "--unbundle", "--type=o", f"--input={fat}",
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True, capture_output=True)
    out = subprocess.run([str(tools[2]), "-d", "--no-show-raw-insn", str(co)], check=True,
                         capture_output=True, text=True).stdout
    shutil.rmtree(d, ignore_errors=True)
    return out


def _kernels(text):
    """{symbol: [instruction lines]} from llvm-objdump output."""
    blocks, cur = {}, None
    for line in text.splitlines():
        m = re.match(r"^[0-9a-f]+ <([^>]+)>:$", line)
        if m:
            cur = m.group(1)
            blocks[cur] = []
        elif cur is not None and line.startswith("\t"):
            blocks[cur].append(line.strip().split("//")[0].strip())
    return blocks


def _is_load(ins):
    return re.match(r"(global|buffer)_load", ins) is not None


def _loads_then_full_wait(insns, window=24):
    """Loads followed (within two instructions) by a full wait for memory when another load
    comes within `window` instructions after that wait: two loads that could have been in
    flight together. A full wait after the last load of a group, before the computation that
    uses the group, is the normal case and not counted."""
    n_loads = waits = 0
    for i, ins in enumerate(insns):
        if not _is_load(ins):
            continue
        n_loads += 1
        for j in range(i + 1, min(i + 3, len(insns))):
            if insns[j].startswith("s_waitcnt") and "vmcnt(0)" in insns[j]:
                if any(_is_load(x) for x in insns[j + 1:j + 1 + window]):
                    waits += 1
                break
    return n_loads, waits


def test_code_object_is_gfx950_and_has_the_kernels(disassembly):
    blocks = _kernels(disassembly)
    for name, prefix in HOT.items():
        assert any(s.startswith(prefix) for s in blocks), f"{name}: kernel {prefix}* not in the code object"


@pytest.mark.parametrize("name", sorted(HOT))
def test_hot_kernel_loads_stay_in_flight(disassembly, name):
    blocks = _kernels(disassembly)
    sym = next(s for s in blocks if s.startswith(HOT[name]))
    n_loads, waits = _loads_then_full_wait(blocks[sym])
    assert n_loads > 0
    assert waits == 0, f"{name}: {waits} of {n_loads} global loads are followed by s_waitcnt vmcnt(0)"

"""CPU-side checks of the boundary: libholo.so loads, exports every symbol declared in
include/holo.h, the ctypes table matches the header, host-only entry points work, and the
Python mirror validates arguments like the reference before touching a GPU."""
import ctypes
import re
from pathlib import Path

import pytest
import torch

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "holo.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(holo_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from particleholography import _lib

    lib = _lib.load()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.SIGNATURES) == names


def test_header_is_plain_c():
    text = HEADER.read_text()
    assert 'extern "C"' in text
    for banned in ("torch", "std::", "hipStream_t", "template"):
        assert banned not in re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def test_supported_lengths_host_only():
    from particleholography import _lib

    lib = _lib.load()
    for n in (1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 40, 48, 80, 96, 512, 768, 1024, 1280, 2048, 4096):
        assert lib.holo_supported_length(n) == 1, n
    for n in (0, 7, 9, 14, 15, 25, 8192, 4097):
        assert lib.holo_supported_length(n) == 0, n
    assert lib.holo_version() >= 1


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """ctypes mirrors of holo_tf / holo_job / holo_gs_job against gcc's layout of the header."""
    import subprocess

    from particleholography import _lib

    fields = {"holo_tf": ["data", "z", "wavlen", "dx", "generated"],
              "holo_job": ["nx", "ny", "src", "src_real", "front", "dz", "slices", "proj"],
              "holo_gs_job": ["nx", "ny", "holo1", "holo2", "fwd", "bwd", "iters", "out"]}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "holo.h"', "int main(void) {"]
    for st, fs in fields.items():
        lines.append(f'  printf("{st} %zu", sizeof({st}));')
        for f in fs:
            lines.append(f'  printf(" %zu", offsetof({st}, {f}));')
        lines.append('  printf("\\n");')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(HEADER.parent), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        name, size, *offs = line.split()
        cls = getattr(_lib, name)
        assert ctypes.sizeof(cls) == int(size), name
        for f, o in zip(fields[name], offs):
            assert getattr(cls, f).offset == int(o), (name, f)


def test_mirror_validates_sizes_before_device_work():
    import particleholography as ph

    wf = ph.CuWavefront(torch.zeros((8, 8), dtype=torch.complex64))
    h8 = ph.CuTransfer(torch.zeros((8, 8), dtype=torch.complex64))
    h4 = ph.CuTransfer(torch.zeros((4, 4), dtype=torch.complex64))
    with pytest.raises(AssertionError, match="same size"):
        ph.cu_get_reconst_xyprojection(wf, h8, h4, 3)
    with pytest.raises(ph.ArgumentError):
        ph.cu_get_reconst_vol(wf, h8, h8, 3, return_type=torch.int32)
    with pytest.raises(AssertionError, match="size"):
        ph.cu_phase_retrieval_holo(torch.zeros(8, 8), torch.zeros(4, 4), h8, h8, 3, 8)
    with pytest.raises(AssertionError, match="2\\*size"):
        ph.cu_get_reconst_vol_and_xyprojection_padded(wf, h8, h8, 2)
    # host tensors are rejected (no CPU fallback)
    with pytest.raises(AssertionError, match="device"):
        ph.cu_get_reconst_xyprojection(wf, h8, h8, 3)


def test_types_forward_julia_size():
    import particleholography as ph

    t = ph.CuTransferSqrtPart(torch.zeros((3, 5)))
    assert t.size() == (5, 3) and t.ndims() == 2 and t.eltype() == torch.float32


def test_transfer_arrays_are_checked_before_any_launch():
    """An array-backed transfer is read by the kernels: a wrong element type, shape or a host
    tensor raises on the host (the reference's CuTransfer holds CuArray{ComplexF32,2} only)."""
    import particleholography as ph
    from particleholography.holofunc import _tf_struct

    good = torch.zeros((8, 8), dtype=torch.complex64)
    with pytest.raises(AssertionError, match="ComplexF32"):
        _tf_struct(ph.CuTransfer(torch.zeros((8, 8), dtype=torch.complex128)), False, (8, 8), 0)
    with pytest.raises(AssertionError, match="ComplexF32"):
        _tf_struct(ph.CuTransfer(torch.zeros((8, 8), dtype=torch.float32)), False, (8, 8), 0)
    with pytest.raises(AssertionError, match=r"\(8, 8\)"):
        _tf_struct(ph.CuTransfer(torch.zeros((8, 4), dtype=torch.complex64)), False, (8, 8), 0)
    with pytest.raises(AssertionError, match=r"got \(4, 8\)"):  # a transposed view of the wrong shape
        _tf_struct(ph.CuTransfer(good[:, :4].t()), False, (8, 4), 0)
    with pytest.raises(AssertionError, match="device"):
        _tf_struct(ph.CuTransfer(good), False, (8, 8), 0)
    # generated transfers carry no array
    s = _tf_struct(ph.CuTransfer(good, z=-1e4, wavlen=0.6328, dx=10.0), True, (8, 8), 0)
    assert s.generated == 1 and not s.data


def test_low_pass_filter_is_checked_before_any_launch():
    import particleholography as ph

    wf = ph.CuWavefront(torch.zeros((8, 8), dtype=torch.complex64))
    with pytest.raises(AssertionError, match="device"):
        ph.cu_apply_low_pass_filter(wf, ph.CuLowPassFilter(torch.zeros((8, 8))))

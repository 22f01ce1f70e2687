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


def test_holo_tf_struct_layout():
    from particleholography import _lib

    assert ctypes.sizeof(_lib.holo_tf) == 40
    assert _lib.holo_tf.z.offset == 8 and _lib.holo_tf.generated.offset == 32
    assert ctypes.sizeof(_lib.holo_job) == 120
    assert _lib.holo_job.front.offset == 24 and _lib.holo_job.slices.offset == 104
    assert _lib.holo_job.proj.offset == 112


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

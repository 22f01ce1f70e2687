"""The allocation-free reconstruction entry point (holo_reconst_xyprojection_noalloc) against
holo_reconst_xyprojection, through the C ABI, on two calls in a row (the second smaller, so it
runs on scratch contents the first left behind), plus its size limit.

Bars: <= 1e-5 relative L2 against the allocating entry point (both run the same kernels; the
rounds differ, so the per-slice rotation restarts at different slices) and <= 1e-4 against
the FP64 oracle (the projection bar)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import asm

pytestmark = pytest.mark.gpu

LAM, DX = 0.6328, 10.0


def _tf(z):
    from particleholography import _lib

    return _lib.holo_tf(None, float(z), LAM, DX, 1)


def _wavefront(n, seed, dev):
    rng = np.random.default_rng(seed)
    holo = (1.0 + 0.2 * rng.standard_normal((n, n))).clip(0.0).astype(np.float32)
    psi = torch.from_numpy(np.sqrt(holo).astype(np.complex64)).to(dev)
    return holo, psi


def _call(lib, name, h, n, psi, z0, dz, slices, out):
    from particleholography import _lib

    st = torch.cuda.current_stream().cuda_stream
    f, d = _tf(z0), _tf(dz)
    return getattr(lib, name)(h, st, n, n, psi.data_ptr(), ctypes.byref(f), ctypes.byref(d), slices,
                              out.data_ptr())


def test_noalloc_matches_allocating_entry_point(gpu):
    from particleholography import _lib

    lib = _lib.load()
    hd = _lib.Handle(0)  # fresh handle: no cached plans, empty workspace
    try:
        cases = [(1024, 40, -1.2e4, -250.0, 31), (512, 20, -8e3, -400.0, 32), (1024, 24, -1.5e4, -300.0, 33)]
        for n, slices, z0, dz, seed in cases:
            print(f"noalloc case: {n}^2, {slices} slices, z0={z0}, dz={dz}, seed={seed}", flush=True)
            holo, psi = _wavefront(n, seed, gpu)
            a = torch.empty((n, n), dtype=torch.float32, device=gpu)
            b = torch.empty_like(a)
            ws_before = hd.workspace_size()
            _lib.check(_call(lib, "holo_reconst_xyprojection_noalloc", hd.ptr, n, psi, z0, dz, slices, b))
            torch.cuda.synchronize()
            assert hd.workspace_size() == ws_before  # the allocation-free call left the handle's workspace alone
            _lib.check(_call(lib, "holo_reconst_xyprojection", hd.ptr, n, psi, z0, dz, slices, a))
            torch.cuda.synchronize()
            an, bn = a.cpu().numpy(), b.cpu().numpy()
            err = asm.rel_l2(bn, an)
            print(f"  noalloc vs allocating: rel L2 {err:.3e}", flush=True)
            assert np.isfinite(bn).all()
            assert err <= 1e-5
            if n == 512:
                ref = asm.reconst_xyprojection(asm.gabor_wavefront(holo), slices, "exact", z_front=z0,
                                               z_dz=dz, wavlen=LAM, dx=DX)
                e2 = asm.rel_l2(bn, ref)
                print(f"  noalloc vs FP64 oracle: rel L2 {e2:.3e}", flush=True)
                assert e2 <= 1e-4
    finally:
        hd.close()


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    from particleholography import _lib

    lib = _lib.load()
    h = _lib.handle(0)
    n = 4096  # spectrum + one slice = 256 MiB > 128 MiB
    psi = torch.empty((8,), dtype=torch.complex64, device=gpu)  # never read: the size check fails first
    out = torch.empty((8,), dtype=torch.float32, device=gpu)
    rc = _call(lib, "holo_reconst_xyprojection_noalloc", h.ptr, n, psi, -1e4, -100.0, 4, out)
    assert rc == _lib.HOLO_ERR_BUDGET
    assert "scratch" in lib.holo_last_error().decode()

"""GPU parity of the low-pass filters (frequency_filters.jl) against the oracle, through the
C ABI. The rectangle filter is bit-exact; the super-Gaussian is within one Float32 ulp (FP64
exp of the device library vs numpy, then rounded); the filtered wavefront is within 2e-6
relative L2 of the FP64 oracle (two FP32 FFT round trips)."""
import numpy as np
import pytest
import torch

from oracle import asm

pytestmark = pytest.mark.gpu

LAM, DX = 0.6328, 10.0
CASES = [(0.01, 5.3e-7, 128, 6.5e-6),            # runtests.jl:516-535
         (80000.0 + 100.0 * 1000, LAM, 1024, DX),  # lowpassfilter.md:37
         (2000.0, LAM, 1024, DX),                  # runtests.jl:424
         (2000.0, LAM, 768, DX),
         (3000.0, LAM, 2048, DX),
         (12.0, 1.0, 8, 1.0),
         (40.0, 1.0, 9, 1.0)]


def _ph():
    import particleholography as ph

    return ph


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("z,lam,n,dx", CASES)
def test_rectangle_filter_bit_exact(gpu, z, lam, n, dx):
    ph = _ph()
    f = ph.cu_rectangle_filter(z, lam, n, dx)
    assert isinstance(f, ph.CuLowPassFilter) and tuple(f.data.shape) == (n, n)
    assert f.data.dtype == torch.float32
    np.testing.assert_array_equal(f.data.cpu().numpy(), asm.rectangle_filter(z, lam, n, dx))


@pytest.mark.parametrize("z,lam,n,dx", CASES)
def test_super_gaussian_filter_one_ulp(gpu, z, lam, n, dx):
    ph = _ph()
    f = ph.cu_super_gaussian_filter(z, lam, n, dx).data.cpu().numpy()
    want = asm.super_gaussian_filter(z, lam, n, dx)
    assert f.shape == (n, n)
    assert np.all(np.abs(f - want) <= np.spacing(np.abs(want)))


@pytest.mark.parametrize("nx,ny", [(128, 128), (512, 512), (768, 768), (1024, 1024), (1280, 1024),
                                   (2048, 2048), (6, 10), (5, 3)])
def test_apply_low_pass_filter_matches_oracle(gpu, nx, ny):
    ph = _ph()
    rng = np.random.default_rng(nx + 3 * ny)
    psi = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex64)
    lpf = rng.random((ny, nx)).astype(np.float32)
    want = asm.apply_low_pass_filter(psi, lpf)
    wf = ph.CuWavefront(_t(psi, gpu))
    out = ph.cu_apply_low_pass_filter(wf, ph.CuLowPassFilter(_t(lpf, gpu)))
    assert out is not wf and out.data.dtype == torch.complex64
    assert asm.rel_l2(out.data.cpu().numpy(), want) < 2e-6
    # in place (cu_apply_low_pass_filter!) gives the same values
    ph.cu_apply_low_pass_filter_(wf, ph.CuLowPassFilter(_t(lpf, gpu)))
    np.testing.assert_array_equal(wf.data.cpu().numpy(), out.data.cpu().numpy())


def test_apply_with_built_filters_matches_oracle(gpu):
    """The tutorial's sequence (lowpassfilter.md:33-38) on a synthetic 1024^2 hologram."""
    ph = _ph()
    from particleholography import synth

    holo = synth.hologram(7, 200, 1024)
    wf = ph.cu_gabor_wavefront(_t(holo, gpu))
    for make, ref in ((ph.cu_super_gaussian_filter, asm.super_gaussian_filter),
                      (ph.cu_rectangle_filter, asm.rectangle_filter)):
        lpf = make(2000.0, LAM, 1024, DX)
        got = ph.cu_apply_low_pass_filter(wf, lpf).data.cpu().numpy()
        want = asm.apply_low_pass_filter(np.sqrt(holo.astype(np.float64)), ref(2000.0, LAM, 1024, DX))
        assert asm.rel_l2(got, want) < 2e-6


def test_reference_low_pass_tests(gpu):
    """runtests.jl:538-575: an all-ones filter keeps the shape and type, and the FP32 round
    trip changes the data (sum of absolute differences above 1e-7)."""
    ph = _ph()
    i = np.arange(1, 129)
    data = ((i[None, :] + i[:, None]) + 1j * (i[None, :] - i[:, None])).astype(np.complex64)
    holo = ph.CuWavefront(_t(data, gpu))
    ones = ph.CuLowPassFilter(torch.ones((128, 128), dtype=torch.float32, device=gpu))
    original = holo.data.clone()
    filtered = ph.cu_apply_low_pass_filter(holo, ones)
    assert filtered is not holo and tuple(filtered.data.shape) == (128, 128)
    assert filtered.data.dtype == torch.complex64
    assert float((filtered.data - holo.data).abs().sum()) > 1e-7
    ph.cu_apply_low_pass_filter_(holo, ones)
    assert tuple(holo.data.shape) == (128, 128) and holo.data.dtype == torch.complex64
    assert float((holo.data - original).abs().sum()) > 1e-7
    assert asm.rel_l2(holo.data.cpu().numpy(), data) < 2e-6


def test_apply_rejects_mismatched_filter(gpu):
    ph = _ph()
    wf = ph.CuWavefront(torch.zeros((64, 64), dtype=torch.complex64, device=gpu))
    with pytest.raises(AssertionError, match="same size"):
        ph.cu_apply_low_pass_filter(wf, ph.CuLowPassFilter(torch.ones((32, 32), device=gpu)))

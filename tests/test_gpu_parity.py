"""GPU parity of libholo against the FP64 oracle (through the C ABI).

Tolerances: FFTs <= 2e-6 relative L2 (complex FP32); projections <= 1e-4 relative L2 vs the
FP64 oracle (the stated bar); array-provider reconstructions <= 2e-5 vs the oracle's
reference-emulation mode; GS phase <= 1e-3 rad amplitude-weighted RMS; the Float32 d array
bit-exact."""
import numpy as np
import pytest
import torch

from oracle import asm

pytestmark = pytest.mark.gpu

LAM, DX = 0.6328, 10.0


def _ph():
    import particleholography as ph

    return ph


def _holo(n, seed, ny=None, amp=0.2):
    ny = n if ny is None else ny
    rng = np.random.default_rng(seed)
    return (1.0 + amp * rng.standard_normal((ny, n))).clip(0.0).astype(np.float32)


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


# ---------------------------------------------------------------------------------------
# A12: FFT
# ---------------------------------------------------------------------------------------
FFT_SIZES = [(1, 1), (2, 2), (4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (3, 3), (5, 5), (6, 6),
             (12, 12), (20, 20), (24, 24), (40, 40), (48, 48), (96, 80), (128, 256), (384, 384),
             (640, 640), (512, 512), (768, 768), (1024, 1024), (1280, 1024), (2048, 2048), (4096, 64),
             (64, 4096), (4096, 4096)]


@pytest.mark.parametrize("nx,ny", FFT_SIZES)
def test_fft2_matches_numpy(gpu, nx, ny):
    ph = _ph()
    rng = np.random.default_rng(nx * 7 + ny)
    x = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex64)
    xt = _t(x, gpu)
    f = ph.cu_fft2(xt).cpu().numpy()
    want = np.fft.fft2(x.astype(np.complex128))
    assert asm.rel_l2(f, want) < 2e-6
    b = ph.cu_ifft2(_t(f, gpu)).cpu().numpy()
    assert asm.rel_l2(b, x) < 4e-6


# ---------------------------------------------------------------------------------------
# A2 / A3 / A4
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [4, 8, 33, 512, 1024])
def test_transfer_sqrt_bit_exact(gpu, n):
    ph = _ph()
    d = ph.cu_transfer_sqrt_arr(n, LAM, DX).data.cpu().numpy()
    np.testing.assert_array_equal(d, asm.transfer_sqrt_arr(n, LAM, DX))
    d64 = ph.cu_transfer_sqrt_arr(n, LAM, DX, dtype=torch.float64).data.cpu().numpy()
    np.testing.assert_array_equal(d64, asm.transfer_sqrt_arr(n, LAM, DX, dtype=np.float64))


@pytest.mark.parametrize("z", [-100.0, 80000.0, -2e5])
def test_transfer_modes(gpu, z):
    ph = _ph()
    n = 256
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    href = ph.cu_transfer(z, n, LAM, dsq, precision="reference").data.cpu().numpy()
    want = asm.transfer_reference(z, LAM, asm.transfer_sqrt_arr(n, LAM, DX))
    # same F32 sqrt, FP64 sincos: equal to within one float32 ulp of 1
    assert np.max(np.abs(href - want)) <= 1.2e-7
    hex_ = ph.cu_transfer(z, n, LAM, dsq).data.cpu().numpy()
    assert np.max(np.abs(hex_ - asm.transfer_exact(z, LAM, DX, n))) < 1e-6


def test_gabor_wavefront_exact(gpu):
    ph = _ph()
    h = _holo(64, 1)
    wf = ph.cu_gabor_wavefront(_t(h, gpu)).data.cpu().numpy()
    np.testing.assert_array_equal(wf, asm.gabor_wavefront(h))


# ---------------------------------------------------------------------------------------
# A6: projection, generated transfers vs the FP64 oracle
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("n,slices,z0,dz", [(64, 5, -1e4, -1e3), (128, 37, -2e4, -250.0),
                                            (96, 9, -5e3, -700.0), (256, 20, -8e4, -100.0)])
def test_projection_generated_vs_fp64_oracle(gpu, n, slices, z0, dz):
    ph = _ph()
    h = _holo(n, n + slices)
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    h0, hdz = ph.cu_transfer(z0, n, LAM, dsq), ph.cu_transfer(dz, n, LAM, dsq)
    wf = ph.cu_gabor_wavefront(_t(h, gpu))
    got = ph.cu_get_reconst_xyprojection(wf, h0, hdz, slices).cpu().numpy()
    want = asm.reconst_xyprojection(asm.gabor_wavefront(h), slices, "exact", z_front=z0, z_dz=dz, wavlen=LAM, dx=DX)
    assert asm.rel_l2(got, want) < 1e-4
    # sqrt(I) fused into the first FFT pass: same arithmetic up to FMA contraction of the
    # zero imaginary part
    fused = ph.reconst_xyprojection_holo(_t(h, gpu), h0, hdz, slices).cpu().numpy()
    assert asm.rel_l2(fused, got) < 1e-6


@pytest.mark.parametrize("nx,ny,slices", [(512, 512, 40), (768, 768, 60), (1024, 1024, 70), (1280, 1024, 30),
                                          (2048, 2048, 66), (4096, 4096, 3)])
def test_projection_configured_sizes_vs_fp64_oracle(gpu, nx, ny, slices):
    """The production kernels and tiles at the configured sides (compile-time plans, shipped
    per-size rounds: 32 slices at 1024^2, 25 at 1280x1024, 64 at 2048^2, so the slice counts
    cross round boundaries) and at the largest supported side, against the FP64 oracle."""
    from particleholography.batch import ReconJob, reconst_batch

    ph = _ph()
    z0, dz = -1e4, -300.0
    h = _holo(nx, nx + slices, ny=ny)
    want = asm.reconst_xyprojection(asm.gabor_wavefront(h), slices, "exact", z_front=z0, z_dz=dz, wavlen=LAM, dx=DX)
    got = reconst_batch([ReconJob(_t(h, gpu), z0, dz, slices, LAM, DX)])[0].cpu().numpy()
    print(f"projection {nx}x{ny} x {slices}: rel L2 vs FP64 oracle {asm.rel_l2(got, want):.2e}")
    assert asm.rel_l2(got, want) < 1e-4
    if nx == ny:
        dsq = ph.cu_transfer_sqrt_arr(nx, LAM, DX)
        h0, hdz = ph.cu_transfer(z0, nx, LAM, dsq), ph.cu_transfer(dz, nx, LAM, dsq)
        single = ph.cu_get_reconst_xyprojection(ph.cu_gabor_wavefront(_t(h, gpu)), h0, hdz, slices).cpu().numpy()
        assert asm.rel_l2(single, want) < 1e-4


@pytest.mark.parametrize("batch,spgc,spgr", [(1, 1, 1), (3, 2, 1), (4, 3, 5), (64, 7, 2)])
def test_projection_batching_is_invisible(gpu, batch, spgc, spgr):
    ph = _ph()
    from particleholography import _lib

    n, slices = 128, 23
    h = _holo(n, 3)
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    h0, hdz = ph.cu_transfer(-3e4, n, LAM, dsq), ph.cu_transfer(-321.0, n, LAM, dsq)
    wf = ph.cu_gabor_wavefront(_t(h, gpu))
    base = ph.cu_get_reconst_xyprojection(wf, h0, hdz, slices).cpu().numpy()
    hd = _lib.handle(0)
    try:
        hd.set_tuning(batch, spgc, spgr, 0, 0)
        got = ph.cu_get_reconst_xyprojection(wf, h0, hdz, slices).cpu().numpy()
    finally:
        hd.set_tuning()
    # slices see the same FP64 phase up to the order of the z sum: equal to ~1 ulp
    assert asm.rel_l2(got, base) < 1e-6


@pytest.mark.parametrize("nx,ny", [(1024, 1024), (1280, 1024), (512, 512), (2048, 2048)])
def test_pass_a_exchange_variants_bitwise_equal(gpu, nx, ny):
    """Pass A's exchange variants move the same values through different routes: one LDS buffer;
    two alternating by exchange, twiddles from registers; for 1024-tall columns the last exchange by lane swaps (v_permlane16/32_swap, swizzled thread layout) with
    the other in two buffers alternating by slice, twiddles from LDS or from registers (other
    sides run their two-buffer form for those settings). The projection is bitwise the same for
    each, at every tile width up to 512 threads, over a slice count that leaves a short last
    group (odd slice counts per group use both buffers first), and within the bar of the FP64
    oracle. Tiles of 2 and 4 columns store the intermediate slices in 4 x 4 blocks and 8-column
    tiles row-major, so the comparison across widths also covers the two layouts of pass A's
    stores and pass B's loads."""
    import ctypes

    from particleholography import _lib, batch

    slices = 37
    h = _holo(nx, 21, ny)
    ht = _t(h, gpu)
    lib = _lib.load()
    proj = torch.empty((ny, nx), dtype=torch.float32, device=gpu)
    f = _lib.holo_tf(None, -2e4, LAM, DX, 1)
    d = _lib.holo_tf(None, -333.0, LAM, DX, 1)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    hd = _lib.handle(0)
    v0 = batch._variants(nx, ny, 0)
    outs = {}
    try:
        for cd in (-1, 1, 2, 3):
            batch._variants(nx, ny, 0, {"col_db": cd, "row_pf": v0["row_pf"]})
            for w in ((2, 4, 8) if ny <= 1024 else (2, 4)):
                hd.set_tuning(0, 0, 0, w, 0)
                _lib.check(lib.holo_reconst_xyprojection_holo(hd.ptr, st, nx, ny, ctypes.c_void_p(ht.data_ptr()),
                                                               ctypes.byref(f), ctypes.byref(d), slices,
                                                               ctypes.c_void_p(proj.data_ptr())))
                outs[(cd, w)] = proj.cpu().numpy()
    finally:
        hd.set_tuning()
        batch._variants(nx, ny, 0, v0)
    base = outs[(-1, 4)]
    for k, v in outs.items():
        assert np.array_equal(v, base), f"col_db {k[0]}, {k[1]} columns per workgroup"
    want = asm.reconst_xyprojection(asm.gabor_wavefront(h), slices, "exact", z_front=-2e4, z_dz=-333.0,
                                    wavlen=LAM, dx=DX)
    err = asm.rel_l2(base, want)
    print(f"pass-A exchange variants {nx}x{ny} x {slices}: bitwise equal, {err:.2e} rel L2 vs FP64 oracle")
    assert err < 1e-4


@pytest.mark.parametrize("n", [768, 2048])
def test_blocked_and_row_major_intermediate_slices_bitwise_equal(gpu, n, monkeypatch):
    """768^2 and 2048^2 run pass A on 4-column tiles, whose intermediate slices are stored in 4 x 4
    blocks; a handle created with HOLO_GBLK_W=0 stores them row-major with the same tiles. The
    projection and a Float32 volume with its projection (the store pass reads the slices too)
    are the same bit for bit."""
    import ctypes

    from particleholography import _lib

    slices = 19
    h = _holo(n, 23)
    wf = _t(np.sqrt(h).astype(np.complex64), gpu)
    lib = _lib.load()
    f = _lib.holo_tf(None, -1.5e4, LAM, DX, 1)
    d = _lib.holo_tf(None, -275.0, LAM, DX, 1)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    monkeypatch.setenv("HOLO_GBLK_W", "0")
    row_major = _lib.Handle(0)  # the layout knob is read when a handle is created
    monkeypatch.delenv("HOLO_GBLK_W")
    outs = []
    for hd in (_lib.handle(0), row_major):
        proj = torch.empty((n, n), dtype=torch.float32, device=gpu)
        vproj = torch.empty((n, n), dtype=torch.float32, device=gpu)
        vol = torch.empty((slices, n, n), dtype=torch.float32, device=gpu)
        _lib.check(lib.holo_reconst_xyprojection(hd.ptr, st, n, n, ctypes.c_void_p(wf.data_ptr()), ctypes.byref(f),
                                                 ctypes.byref(d), slices, ctypes.c_void_p(proj.data_ptr())))
        _lib.check(lib.holo_reconst_vol(hd.ptr, st, n, n, ctypes.c_void_p(wf.data_ptr()), ctypes.byref(f),
                                        ctypes.byref(d), slices, 2, ctypes.c_void_p(vol.data_ptr()),
                                        ctypes.c_void_p(vproj.data_ptr())))
        torch.cuda.synchronize()
        outs.append((proj.cpu().numpy(), vol.cpu().numpy(), vproj.cpu().numpy()))
    for a, b, what in zip(outs[0], outs[1], ("projection", "volume", "volume projection")):
        assert np.array_equal(a, b), f"{n}^2 {what}: blocked and row-major intermediate slices differ"
    print(f"{n}^2 x {slices}: blocked == row-major intermediate slices (projection, volume)")


@pytest.mark.parametrize("nx,ny", [(768, 512), (512, 768)])
def test_blocked_layout_non_square(gpu, nx, ny):
    """Non-square planes whose sides are multiples of 4: 4-column pass-A tiles store the
    intermediate slices in 4 x 4 blocks (nx / 4 blocks per block row), 8-column tiles row-major;
    the projections are bitwise equal and within the bar of the FP64 oracle."""
    import ctypes

    from particleholography import _lib

    slices = 21
    h = _holo(nx, 29, ny)
    ht = _t(h, gpu)
    lib = _lib.load()
    f = _lib.holo_tf(None, -1.2e4, LAM, DX, 1)
    d = _lib.holo_tf(None, -410.0, LAM, DX, 1)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    hd = _lib.handle(0)
    outs = {}
    try:
        for w in (4, 8):
            hd.set_tuning(0, 0, 0, w, 0)
            proj = torch.empty((ny, nx), dtype=torch.float32, device=gpu)
            _lib.check(lib.holo_reconst_xyprojection_holo(hd.ptr, st, nx, ny, ctypes.c_void_p(ht.data_ptr()),
                                                           ctypes.byref(f), ctypes.byref(d), slices,
                                                           ctypes.c_void_p(proj.data_ptr())))
            outs[w] = proj.cpu().numpy()
    finally:
        hd.set_tuning()
    assert np.array_equal(outs[4], outs[8])
    want = asm.reconst_xyprojection(asm.gabor_wavefront(h), slices, "exact", z_front=-1.2e4, z_dz=-410.0,
                                    wavlen=LAM, dx=DX)
    err = asm.rel_l2(outs[4], want)
    print(f"{nx}x{ny} x {slices}: blocked == row-major, {err:.2e} rel L2 vs FP64 oracle")
    assert err < 1e-4


@pytest.mark.parametrize("nx,ny", [(80, 64), (64, 96), (160, 48)])
def test_projection_non_square(gpu, nx, ny):
    from particleholography import _lib
    import ctypes

    h = _holo(nx, 4, ny=ny)
    lib = _lib.load()
    ht = _t(h, gpu)
    proj = torch.empty((ny, nx), dtype=torch.float32, device=gpu)
    f = _lib.holo_tf(None, -1.5e4, LAM, DX, 1)
    d = _lib.holo_tf(None, -400.0, LAM, DX, 1)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.holo_reconst_xyprojection_holo(_lib.handle(0).ptr, st, nx, ny, ctypes.c_void_p(ht.data_ptr()),
                                                   ctypes.byref(f), ctypes.byref(d), 6, ctypes.c_void_p(proj.data_ptr())))
    got = proj.cpu().numpy()
    want = asm.reconst_xyprojection(asm.gabor_wavefront(h), 6, "exact", z_front=-1.5e4, z_dz=-400.0, wavlen=LAM, dx=DX)
    assert asm.rel_l2(got, want) < 1e-4


def test_projection_512_seeded_100_slices_vs_fp64_oracle(gpu):
    """512^2, 100 slices, z0 = 1e4, dz = 1e3 um on a seeded disk hologram (the CPU config)."""
    ph = _ph()
    from particleholography import synth

    holo = synth.hologram(101, 200, 512, zmin=1e4, zmax=1e5)
    dsq = ph.cu_transfer_sqrt_arr(512, LAM, DX)
    h0, hdz = ph.cu_transfer(-1e4, 512, LAM, dsq), ph.cu_transfer(-1e3, 512, LAM, dsq)
    got = ph.cu_get_reconst_xyprojection(ph.cu_gabor_wavefront(_t(holo, gpu)), h0, hdz, 100).cpu().numpy()
    want = asm.reconst_xyprojection(asm.gabor_wavefront(holo), 100, "exact", z_front=-1e4, z_dz=-1e3,
                                    wavlen=LAM, dx=DX)
    err = asm.rel_l2(got, want)
    print(f"512^2 x 100 seeded projection rel L2 vs FP64 oracle: {err:.3e}")
    assert err < 1e-4


def test_projection_1024_seeded_1000_slices_vs_fp64_oracle(gpu):
    """The headline workload end to end: the seed-102 1000-disk 1024^2 hologram, 1000 slices at
    dz = 200 um from z0 = 1e4 um (to 2.098e5 um, holofunc.jl:259-263), through the fused path
    (sqrt fused into the forward FFT, the shipped tiles), against the FP64 oracle."""
    ph = _ph()
    from particleholography import synth

    n, slices = 1024, 1000
    holo = synth.hologram(102, 1000, n, zmin=1e4, zmax=2e5, wavlen=LAM, dx=DX)
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    h0, hdz = ph.cu_transfer(-1e4, n, LAM, dsq), ph.cu_transfer(-200.0, n, LAM, dsq)
    got = ph.reconst_xyprojection_holo(_t(holo, gpu), h0, hdz, slices).cpu().numpy()
    want = asm.reconst_xyprojection_running(asm.gabor_wavefront(holo), slices, z_front=-1e4, z_dz=-200.0,
                                            wavlen=LAM, dx=DX)
    err = asm.rel_l2(got, want)
    print(f"1024^2 x 1000 seeded projection rel L2 vs FP64 oracle: {err:.3e} (max abs {np.abs(got - want).max():.3e})")
    assert err < 1e-4


def test_gs_1024_seeded_pair_100_iterations_vs_fp64_oracle(gpu):
    """The GS configuration: the seeded 500-disk scene recorded at planes 2000 um apart
    (1024^2), 100 iterations on the production path (graphs, line shapes), against the FP64
    oracle's 100 iterations: amplitude-weighted phase RMS."""
    ph = _ph()
    from particleholography import synth

    n, z, iters = 1024, 2000.0, 100
    i1, i2 = synth.gs_pair(103, 500, n, z)
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    out = ph.cu_phase_retrieval_holo(_t(i1, gpu), _t(i2, gpu), ph.cu_transfer(z, n, LAM, dsq),
                                     ph.cu_transfer(-z, n, LAM, dsq), iters, n).data.cpu().numpy()
    want = asm.phase_retrieval(i1, i2, iters, z=z, wavlen=LAM, dx=DX)
    np.testing.assert_allclose(np.abs(out), np.sqrt(i1), rtol=2e-6)
    err = asm.weighted_phase_rms(out, want)
    print(f"GS 1024^2 seeded pair x 100 iterations: weighted phase RMS vs FP64 oracle {err:.2e} rad")
    assert err < 1e-3


# ---------------------------------------------------------------------------------------
# A6 array provider: the reference's ComplexF32 running product
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("n,slices,batch", [(64, 12, 0), (128, 30, 7)])
def test_projection_arrays_vs_reference_emulation(gpu, n, slices, batch):
    ph = _ph()
    from particleholography import _lib

    h = _holo(n, 21)
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    h0 = ph.cu_transfer(-8e4, n, LAM, dsq, precision="reference")
    hdz = ph.cu_transfer(-100.0, n, LAM, dsq, precision="reference")
    assert not h0.generated
    wf = ph.cu_gabor_wavefront(_t(h, gpu))
    hd = _lib.handle(0)
    try:
        hd.set_tuning(batch, 0, 0, 0, 0)
        got = ph.cu_get_reconst_xyprojection(wf, h0, hdz, slices).cpu().numpy()
    finally:
        hd.set_tuning()
    want = asm.reconst_xyprojection(asm.gabor_wavefront(h), slices, "reference",
                                    h_front=h0.data.cpu().numpy(), h_dz=hdz.data.cpu().numpy())
    assert asm.rel_l2(got, want) < 2e-5


# ---------------------------------------------------------------------------------------
# A7 / A8 / A9 volumes
# ---------------------------------------------------------------------------------------
def _vol_case(gpu, n=64, slices=6):
    ph = _ph()
    h = (0.5 * _holo(n, 31, amp=0.6)).astype(np.float32)  # straddles the clamp at 1
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    h0, hdz = ph.cu_transfer(-6e3, n, LAM, dsq), ph.cu_transfer(-900.0, n, LAM, dsq)
    return ph, h, ph.cu_gabor_wavefront(_t(h, gpu)), h0, hdz, dict(z_front=-6e3, z_dz=-900.0, wavlen=LAM, dx=DX)


def test_volume_n0f8(gpu):
    ph, h, wf, h0, hdz, kw = _vol_case(gpu)
    vol = ph.cu_get_reconst_vol(wf, h0, hdz, 6).cpu().numpy()
    want = asm.reconst_vol(asm.gabor_wavefront(h), 6, np.uint8, "exact", **kw)
    diff = np.abs(vol.astype(int) - want.astype(int))
    assert diff.max() <= 1 and np.mean(diff) < 0.01


@pytest.mark.parametrize("rt,np_t,tol", [("f32", np.float32, 2e-5), ("f64", np.float64, 2e-5), ("f16", np.float16, 1e-3)])
def test_volume_float_types(gpu, rt, np_t, tol):
    ph, h, wf, h0, hdz, kw = _vol_case(gpu)
    t = {"f32": ph.Float32, "f64": ph.Float64, "f16": ph.Float16}[rt]
    vol = ph.cu_get_reconst_vol(wf, h0, hdz, 6, return_type=t).cpu().numpy()
    want = asm.reconst_vol(asm.gabor_wavefront(h), 6, np_t, "exact", **kw)
    assert vol.dtype == np_t
    assert np.max(np.abs(vol.astype(np.float64) - want.astype(np.float64))) < tol


def test_complex_volume(gpu):
    ph, h, wf, h0, hdz, kw = _vol_case(gpu)
    cv = ph.cu_get_reconst_complex_vol(wf, h0, hdz, 6).cpu().numpy()
    want = asm.reconst_complex_vol(asm.gabor_wavefront(h), 6, "exact", **kw)
    for k in range(6):
        assert asm.rel_l2(cv[k], want[k]) < 1e-5


@pytest.mark.parametrize("n,slices,rt", [(1024, 40, "n0f8"), (2048, 8, "n0f8"), (768, 60, "f32"), (1024, 4, "c64")])
def test_volumes_configured_sizes_vs_fp64_oracle(gpu, n, slices, rt):
    """The store epilogues on the production kernels and tiles (1024^2 crosses a 32-slice
    round): N0f8 within 1 LSB, Float32 within 2e-5, ComplexF32 within 1e-5 rel L2 per slice."""
    ph = _ph()
    h = (0.5 * _holo(n, n + slices, amp=0.6)).astype(np.float32)  # straddles the clamp at 1
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    z0, dz = -8e3, -700.0
    h0, hdz = ph.cu_transfer(z0, n, LAM, dsq), ph.cu_transfer(dz, n, LAM, dsq)
    wf = ph.cu_gabor_wavefront(_t(h, gpu))
    kw = dict(z_front=z0, z_dz=dz, wavlen=LAM, dx=DX)
    if rt == "c64":
        cv = ph.cu_get_reconst_complex_vol(wf, h0, hdz, slices).cpu().numpy()
        want = asm.reconst_complex_vol(asm.gabor_wavefront(h), slices, "exact", **kw)
        assert max(asm.rel_l2(cv[k], want[k]) for k in range(slices)) < 1e-5
    elif rt == "n0f8":
        vol = ph.cu_get_reconst_vol(wf, h0, hdz, slices).cpu().numpy()
        want = asm.reconst_vol(asm.gabor_wavefront(h), slices, np.uint8, "exact", **kw)
        diff = np.abs(vol.astype(int) - want.astype(int))
        assert diff.max() <= 1 and np.mean(diff) < 0.01
    else:
        vol = ph.cu_get_reconst_vol(wf, h0, hdz, slices, return_type=ph.Float32).cpu().numpy()
        want = asm.reconst_vol(asm.gabor_wavefront(h), slices, np.float32, "exact", **kw)
        assert np.max(np.abs(vol.astype(np.float64) - want.astype(np.float64))) < 2e-5


@pytest.mark.parametrize("rt", ["n0f8", "f32"])
def test_volume_and_projection(gpu, rt):
    ph, h, wf, h0, hdz, kw = _vol_case(gpu)
    t = ph.N0f8 if rt == "n0f8" else ph.Float32
    vol, proj = ph.cu_get_reconst_vol_and_xyprojection(wf, h0, hdz, 6, return_type=t)
    vol, proj = vol.cpu().numpy(), proj.cpu().numpy()
    # the projection is exactly the min of the stored values (holofunc.jl:268-282)
    np.testing.assert_array_equal(proj, asm.xy_projection_from_vol(vol))


# ---------------------------------------------------------------------------------------
# f2: pad + padded reconstruction
# ---------------------------------------------------------------------------------------
def test_2d_pad_known_answer(gpu):
    ph = _ph()
    arr = torch.tensor([[1 + 1j, 2 + 1j], [3 + 1j, 4 + 1j]], dtype=torch.complex64, device=gpu)
    p = ph.cu_2d_pad(arr).cpu().numpy()
    assert p.shape == (4, 4)
    np.testing.assert_array_equal(p[1:3, 1:3], arr.cpu().numpy())
    np.testing.assert_array_equal(p, asm.pad2(arr.cpu().numpy()))


@pytest.mark.parametrize("n,slices", [(4, 2), (32, 5), (512, 4), (1024, 3)])
def test_padded_reconstruction(gpu, n, slices):
    """Padded volume + projection against the FP64 oracle; 512 and 1024 run the 1024- and
    2048-point compile-time plans with the cropped store window."""
    ph = _ph()
    h = _holo(n, 41, amp=0.5) * 0.6
    dsq = ph.cu_transfer_sqrt_arr(2 * n, LAM, DX)
    h0, hdz = ph.cu_transfer(-100.0, 2 * n, LAM, dsq), ph.cu_transfer(-10.0, 2 * n, LAM, dsq)
    wf = ph.cu_gabor_wavefront(_t(h.astype(np.float32), gpu))
    vol, proj = ph.cu_get_reconst_vol_and_xyprojection_padded(wf, h0, hdz, slices)
    assert tuple(vol.shape) == (slices, n, n) and vol.dtype == torch.uint8 and proj.dtype == torch.float32
    want_vol, _ = asm.reconst_vol_and_xyprojection_padded(asm.gabor_wavefront(h.astype(np.float32)), slices,
                                                          "exact", z_front=-100.0, z_dz=-10.0, wavlen=LAM, dx=DX)
    assert np.abs(vol.cpu().numpy().astype(int) - want_vol.astype(int)).max() <= 1
    np.testing.assert_array_equal(proj.cpu().numpy(), asm.xy_projection_from_vol(vol.cpu().numpy()))


@pytest.mark.parametrize("n", [512, 1024])
def test_padded_crop_at_zero_distance(gpu, n):
    """Closed form of the padded reconstruction's window: at z = 0 (H = 1 for every slice) each
    stored slice is n0f8(clamp(|psi|^2, 0, 1)) of the unpadded wavefront itself, the centre window
    of the 2n x 2n round trip pad -> FFT -> IFFT (holofunc.jl:339-364); within one N0f8 step
    (the round trip rounds in FP32), and exact for nearly every pixel."""
    ph = _ph()
    rng = np.random.default_rng(n + 3)
    h = (0.15 + 0.7 * rng.random((n, n))).astype(np.float32)  # |psi|^2 = h in (0, 1): no clamping
    dsq = ph.cu_transfer_sqrt_arr(2 * n, LAM, DX)
    h0 = ph.cu_transfer(0.0, 2 * n, LAM, dsq)
    vol, proj = ph.cu_get_reconst_vol_and_xyprojection_padded(ph.cu_gabor_wavefront(_t(h, gpu)), h0, h0, 2)
    want = asm.n0f8(np.abs(asm.gabor_wavefront(h).astype(np.complex128)) ** 2)
    got = vol.cpu().numpy().astype(int)
    for k in range(2):
        diff = np.abs(got[k] - want.astype(int))
        assert diff.max() <= 1
        assert (diff == 0).mean() > 0.99
    np.testing.assert_array_equal(proj.cpu().numpy(), asm.xy_projection_from_vol(vol.cpu().numpy()))


# ---------------------------------------------------------------------------------------
# A10: Gerchberg-Saxton
# ---------------------------------------------------------------------------------------
def _gs_case(n, seed, z=2000.0):
    rng = np.random.default_rng(seed)
    psi = np.exp(1j * 0.7 * rng.standard_normal((n, n))) * (1 + 0.3 * rng.random((n, n)))
    i1 = (np.abs(psi) ** 2).astype(np.float32)
    i2 = (np.abs(asm.asm_prop(psi, z, LAM, DX)) ** 2).astype(np.float32)
    return i1, i2


@pytest.mark.parametrize("n,iters", [(64, 1), (64, 10), (128, 25), (512, 20), (1024, 20), (2048, 3), (5, 6)])
def test_gs_generated_vs_fp64_oracle(gpu, n, iters):
    """Small sides on runtime plans; 512^2, 1024^2 and 2048^2 on the single-plane line shapes
    and their compile-time plans (20 iterations: one 16-iteration graph plus a remainder). Even
    sides read their generated transfers as quadrant tables; the odd side 5 (odd supported sides
    are 3 and 5) reads the whole table."""
    ph = _ph()
    z = 2000.0
    i1, i2 = _gs_case(n, n + iters, z)
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    hf, hb = ph.cu_transfer(z, n, LAM, dsq), ph.cu_transfer(-z, n, LAM, dsq)
    out = ph.cu_phase_retrieval_holo(_t(i1, gpu), _t(i2, gpu), hf, hb, iters, n).data.cpu().numpy()
    want = asm.phase_retrieval(i1, i2, iters, z=z, wavlen=LAM, dx=DX)
    np.testing.assert_allclose(np.abs(out), np.sqrt(i1), rtol=2e-6)
    err = asm.weighted_phase_rms(out, want)
    print(f"GS {n}^2 x {iters}: weighted phase RMS {err:.2e} rad")
    assert err < 1e-3


def test_gs_graph_equals_eager(gpu):
    ph = _ph()
    n, z = 128, 2000.0
    i1, i2 = _gs_case(n, 5, z)
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    hf, hb = ph.cu_transfer(z, n, LAM, dsq), ph.cu_transfer(-z, n, LAM, dsq)
    a = ph.cu_phase_retrieval_holo(_t(i1, gpu), _t(i2, gpu), hf, hb, 7, n, use_graph=False).data.cpu().numpy()
    for _ in range(3):  # graph replays are repeatable, with eager work between them
        b = ph.cu_phase_retrieval_holo(_t(i1, gpu), _t(i2, gpu), hf, hb, 7, n, use_graph=True).data.cpu().numpy()
        np.testing.assert_array_equal(a, b)
        torch.cuda.synchronize()


def test_gs_graphs_reused_across_inputs_and_iteration_counts(gpu):
    """The cached graphs hold only workspace addresses: new input tensors, other iteration
    counts (graphs of 16 iterations plus a remainder) and another transfer pair replay them and
    still equal the eager path bit for bit."""
    ph = _ph()
    n, z = 128, 2000.0
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    mirrored = (ph.cu_transfer(z, n, LAM, dsq), ph.cu_transfer(-z, n, LAM, dsq))
    h_other = ph.cu_transfer(-1500.0, n, LAM, dsq)
    arrays = (ph.CuTransfer(mirrored[0].data.clone()), ph.CuTransfer(h_other.data.clone()))  # not mirrored
    for seed, iters, (hf, hb) in [(5, 7, mirrored), (6, 40, mirrored), (7, 18, arrays), (8, 40, arrays), (9, 7, mirrored)]:
        i1, i2 = _gs_case(n, seed, z)
        t1, t2 = _t(i1, gpu), _t(i2, gpu)
        a = ph.cu_phase_retrieval_holo(t1, t2, hf, hb, iters, n, use_graph=False).data.cpu().numpy()
        b = ph.cu_phase_retrieval_holo(_t(i1, gpu), _t(i2, gpu), hf, hb, iters, n, use_graph=True).data.cpu().numpy()
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n", [128, 1024])
def test_gs_quadrant_tables_equal_full_tables(gpu, n):
    """Generated transfers of even sides go to the column pass as the quadrant
    (n/2 + 1) x (n/2 + 1) of their natural-order table, read at min(i, n - i), min(j, n - j);
    array-backed copies of the same transfers go as the whole table. Both hold the same values
    (one FP64 generator), so the retrieved fields are equal bit for bit (a wrong mirror index
    would shift H by one frequency step, about 0.04 rad per pass at 2000 um)."""
    ph = _ph()
    i1, i2 = _gs_case(n, 11, 2000.0)
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    gen = (ph.cu_transfer(2000.0, n, LAM, dsq), ph.cu_transfer(-1500.0, n, LAM, dsq))
    arr = tuple(ph.CuTransfer(t.data.clone()) for t in gen)
    assert gen[0].generated and not arr[0].generated
    a = ph.cu_phase_retrieval_holo(_t(i1, gpu), _t(i2, gpu), *gen, 5, n).data.cpu().numpy()
    b = ph.cu_phase_retrieval_holo(_t(i1, gpu), _t(i2, gpu), *arr, 5, n).data.cpu().numpy()
    print(f"GS {n}^2: quadrant vs full tables, weighted phase RMS {asm.weighted_phase_rms(a, b):.2e} rad")
    np.testing.assert_array_equal(a, b)


def test_gs_arrays_vs_oracle(gpu):
    ph = _ph()
    n, z = 64, 2000.0
    i1, i2 = _gs_case(n, 9, z)
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    hf = ph.cu_transfer(z, n, LAM, dsq, precision="reference")
    hb = ph.cu_transfer(-z, n, LAM, dsq, precision="reference")
    out = ph.cu_phase_retrieval_holo(_t(i1, gpu), _t(i2, gpu), hf, hb, 6, n).data.cpu().numpy()
    want = asm.phase_retrieval(i1, i2, 6, hf.data.cpu().numpy(), hb.data.cpu().numpy())
    assert asm.weighted_phase_rms(out, want) < 1e-3


def test_gs_zero_iterations_is_sqrt(gpu):
    ph = _ph()
    i1, i2 = _gs_case(32, 2)
    dsq = ph.cu_transfer_sqrt_arr(32, LAM, DX)
    hf, hb = ph.cu_transfer(2000.0, 32, LAM, dsq), ph.cu_transfer(-2000.0, 32, LAM, dsq)
    out = ph.cu_phase_retrieval_holo(_t(i1, gpu), _t(i2, gpu), hf, hb, 0, 32).data.cpu().numpy()
    np.testing.assert_array_equal(out, asm.gabor_wavefront(i1))


# ---------------------------------------------------------------------------------------
# A11: one propagation
# ---------------------------------------------------------------------------------------
def test_asm_prop_reference_case(gpu):
    """runtests.jl:385-394: ones(4,4), z = 100 um; the exact answer has |out| = 1."""
    ph = _ph()
    inh = ph.CuWavefront(torch.ones((4, 4), dtype=torch.complex64, device=gpu))
    outh = ph.CuWavefront(torch.zeros((4, 4), dtype=torch.complex64, device=gpu))
    dsq = ph.cu_transfer_sqrt_arr(4, LAM, DX)
    ph.cu_asm_prop_(outh, inh, dsq, 100.0, 4, LAM)
    o = outh.data.cpu().numpy()
    assert o.shape == (4, 4) and np.sum(np.abs(o)) > 0
    np.testing.assert_allclose(np.abs(o), 1.0, atol=1e-6)
    np.testing.assert_allclose(o, asm.asm_prop(np.ones((4, 4), complex), 100.0, LAM, DX), atol=1e-6)


@pytest.mark.parametrize("n,z", [(64, 5000.0), (256, -3e4), (768, 1e4), (512, 2000.0), (1024, -8e4)])
def test_asm_prop_vs_oracle(gpu, n, z):
    ph = _ph()
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    inh = ph.CuWavefront(_t(x, gpu))
    outh = ph.CuWavefront(torch.empty_like(inh.data))
    ph.cu_asm_prop_(outh, inh, ph.cu_transfer_sqrt_arr(n, LAM, DX), z, n, LAM)
    assert asm.rel_l2(outh.data.cpu().numpy(), asm.asm_prop(x, z, LAM, DX)) < 4e-6


def test_determinism(gpu):
    ph = _ph()
    n = 256
    h = _holo(n, 77)
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    h0, hdz = ph.cu_transfer(-2e4, n, LAM, dsq), ph.cu_transfer(-200.0, n, LAM, dsq)
    wf = ph.cu_gabor_wavefront(_t(h, gpu))
    a = ph.cu_get_reconst_xyprojection(wf, h0, hdz, 50).cpu().numpy()
    b = ph.cu_get_reconst_xyprojection(wf, h0, hdz, 50).cpu().numpy()
    np.testing.assert_array_equal(a, b)

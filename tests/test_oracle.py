"""Pins the CPU oracle (oracle/asm.py) with checks that do not restate its formulas:
scalar brute force, closed forms of the propagation, invariants and the reference's own
known-answer test. CPU only."""
import math

import numpy as np
import pytest

from oracle import asm

LAM, DX = 0.6328, 10.0


# ---------------------------------------------------------------------------------------
# A2: d array, bit-exact against a scalar re-evaluation with Python floats
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [4, 5, 8, 16, 33])
def test_transfer_sqrt_scalar_bruteforce(n):
    d = asm.transfer_sqrt_arr(n, LAM, DX)
    for y in range(1, n + 1):          # Julia 1-based indices, holofunc.jl:12
        for x in range(1, n + 1):
            v = 1.0 - ((x - n / 2 - 1.0) * LAM / n / DX) ** 2 - ((y - n / 2 - 1.0) * LAM / n / DX) ** 2
            assert d[y - 1, x - 1] == np.float32(v)


def test_transfer_sqrt_centre_and_symmetry():
    n = 64
    d = asm.transfer_sqrt_arr(n, LAM, DX, dtype=np.float64)
    assert d[n // 2, n // 2] == 1.0                       # x = n/2 + 1 is the DC term
    assert np.all(d[n // 2, 1:] == d[n // 2, 1:][::-1])    # symmetric about the centre
    fx = (np.arange(n) - n / 2) / (n * DX)
    np.testing.assert_allclose(d[n // 2], 1.0 - (LAM * fx) ** 2, rtol=0, atol=1e-15)


# ---------------------------------------------------------------------------------------
# A3: transfer functions
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("nx,ny", [(8, 8), (64, 64), (1024, 1024), (768, 512), (512, 768)])
def test_natural_transfer_is_determined_by_its_quadrant(nx, ny):
    """For even sides the natural-order transfer depends on each axis through the square of
    (ic - N/2), and natural indices n and N - n carry opposite values of it, exactly in FP64:
    H(i, j) = H(min(i, nx - i), min(j, ny - j)). The GS column pass reads generated transfers as
    that quadrant of (nx/2 + 1) x (ny/2 + 1) entries; rebuilt by its index rule, it is the whole
    table bit for bit."""
    h = asm.natural(asm.transfer_exact(2000.0, 0.6328, 10.0, nx, ny))
    q = h[: ny // 2 + 1, : nx // 2 + 1]
    j = np.minimum(np.arange(ny), ny - np.arange(ny))
    i = np.minimum(np.arange(nx), nx - np.arange(nx))
    assert np.array_equal(q[j[:, None], i[None, :]], h)
    # an asymmetric mistake (one index off) is caught
    assert not np.array_equal(q[np.minimum(j + 1, ny // 2)[:, None], i[None, :]], h)


def test_transfer_reference_properties():
    n = 64
    d32 = asm.transfer_sqrt_arr(n, LAM, DX)
    h0 = asm.transfer_reference(0.0, LAM, d32)
    assert np.all(h0 == np.complex64(1.0))
    h = asm.transfer_reference(-8e4, LAM, d32)
    np.testing.assert_allclose(np.abs(h), 1.0, atol=2e-7)
    # DC entry: sqrt(d) = 1 exactly, so H = exp(2 pi i z / lambda)
    want = np.exp(2j * np.pi * (-8e4) / LAM)
    assert abs(h[n // 2, n // 2] - want) < 2e-7


def test_transfer_reference_phase_error_matches_float32_sqrt_bound():
    # The reference rounds d and sqrt(d) to Float32 (holofunc.jl:12,42); the phase error vs
    # the exact transfer grows linearly in z, bounded by |2 pi z / lambda| * (ulp terms).
    n = 256
    d32 = asm.transfer_sqrt_arr(n, LAM, DX)
    for z in (2e3, 8e4):
        href = asm.transfer_reference(z, LAM, d32)
        hex_ = asm.transfer_exact(z, LAM, DX, n)
        dphi = np.angle(href * np.conj(hex_))
        bound = 2 * np.pi * abs(z) / LAM * 2 * np.finfo(np.float32).eps + 1e-6
        assert np.max(np.abs(dphi)) <= bound
    # and it is not zero at large z: the emulation really carries the F32 defect
    assert np.max(np.abs(dphi)) > 1e-3


def test_transfer_exact_composition():
    n = 32
    a = asm.transfer_exact(-1234.5, LAM, DX, n)
    b = asm.transfer_exact(-777.25, LAM, DX, n)
    c = asm.transfer_exact(-1234.5 - 777.25, LAM, DX, n)
    np.testing.assert_allclose(a * b, c, atol=1e-12)


# ---------------------------------------------------------------------------------------
# brute-force O(N^4) DFT propagation
# ---------------------------------------------------------------------------------------
def _dft_propagate(psi, z):
    ny, nx = psi.shape
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    out = np.zeros_like(psi, dtype=np.complex128)
    spec = np.zeros_like(out)
    for ky in range(ny):
        for kx in range(nx):
            spec[ky, kx] = np.sum(psi * np.exp(-2j * np.pi * (kx * ii / nx + ky * jj / ny)))
    for ky in range(ny):
        for kx in range(nx):
            fx = (kx if kx < nx - nx // 2 else kx - nx) / (nx * DX)
            fy = (ky if ky < ny - ny // 2 else ky - ny) / (ny * DX)
            d = 1.0 - (LAM * fx) ** 2 - (LAM * fy) ** 2
            spec[ky, kx] *= np.exp(2j * np.pi * z / LAM * math.sqrt(d))
    for j in range(ny):
        for i in range(nx):
            out[j, i] = np.sum(spec * np.exp(2j * np.pi * (ii * i / nx + jj * j / ny))) / (nx * ny)
    return out


@pytest.mark.parametrize("n", [8, 16])
def test_propagate_matches_bruteforce_dft(n):
    rng = np.random.default_rng(n)
    psi = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    z = -3000.0
    want = _dft_propagate(psi, z)
    got = asm.asm_prop(psi, z, LAM, DX)
    np.testing.assert_allclose(got, want, atol=1e-10 * np.abs(want).max())


# ---------------------------------------------------------------------------------------
# closed forms of the propagation
# ---------------------------------------------------------------------------------------
def test_constant_field_is_invariant():
    n = 32
    psi = np.ones((n, n), complex)
    out = asm.asm_prop(psi, -5e4, LAM, DX)
    np.testing.assert_allclose(np.abs(out) ** 2, 1.0, atol=1e-12)
    proj = asm.reconst_xyprojection(psi, 7, "exact", z_front=-1e4, z_dz=-1e3, wavlen=LAM, dx=DX)
    np.testing.assert_allclose(proj, 1.0, atol=1e-6)


def _grating_intensity(n, m, a, z):
    f = m / (n * DX)
    delta = 2 * np.pi * z / LAM * (math.sqrt(1 - (LAM * f) ** 2) - 1.0)
    c = np.cos(2 * np.pi * m * np.arange(n) / n)
    field = 1.0 + a * c * np.exp(1j * delta)
    return np.broadcast_to(np.abs(field) ** 2, (n, n))


def test_grating_closed_form_projection():
    n, m, a = 64, 5, 0.3
    psi = np.broadcast_to(1.0 + a * np.cos(2 * np.pi * m * np.arange(n) / n), (n, n)).astype(complex)
    z0, dz, s = -1e4, -750.0, 12
    want = np.min([_grating_intensity(n, m, a, z0 + k * dz) for k in range(s)], axis=0)
    got = asm.reconst_xyprojection(psi, s, "exact", z_front=z0, z_dz=dz, wavlen=LAM, dx=DX)
    np.testing.assert_allclose(got, want, atol=2e-7)


def test_delta_constant_pair():
    n = 16
    delta = np.zeros((n, n), complex)
    delta[0, 0] = 1.0
    np.testing.assert_allclose(asm.fft2(delta), np.ones((n, n)), atol=0)
    np.testing.assert_allclose(asm.ifft2(np.ones((n, n))), delta, atol=1e-15)


def test_reversibility_and_parseval():
    n = 64
    rng = np.random.default_rng(3)
    psi = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    fwd = asm.asm_prop(psi, 2500.0, LAM, DX)
    back = asm.asm_prop(fwd, -2500.0, LAM, DX)
    np.testing.assert_allclose(back, psi, atol=1e-11)
    assert abs(np.sum(np.abs(fwd) ** 2) - np.sum(np.abs(psi) ** 2)) < 1e-9 * np.sum(np.abs(psi) ** 2)


def test_reference_mode_running_product_equals_exact_at_small_depth():
    # H0 .* Hdz^(k-1) == H(z0 + (k-1) dz): the reference's running product converges to the
    # exact slice fields when the F32 phase defect is small (short distances, few slices).
    n = 64
    rng = np.random.default_rng(5)
    holo = (1 + 0.2 * rng.standard_normal((n, n))).clip(0).astype(np.float32)
    wf = asm.gabor_wavefront(holo)
    d32 = asm.transfer_sqrt_arr(n, LAM, DX)
    h0, hdz = asm.transfer_reference(-200.0, LAM, d32), asm.transfer_reference(-50.0, LAM, d32)
    ref = asm.reconst_xyprojection(wf, 6, "reference", h_front=h0, h_dz=hdz)
    ex = asm.reconst_xyprojection(wf, 6, "exact", z_front=-200.0, z_dz=-50.0, wavlen=LAM, dx=DX)
    assert asm.rel_l2(ref, ex) < 1e-5


def test_volume_slices_equal_single_propagations():
    n = 32
    rng = np.random.default_rng(9)
    wf = asm.gabor_wavefront((1 + 0.3 * rng.random((n, n))).astype(np.float32))
    cv = asm.reconst_complex_vol(wf, 3, "exact", z_front=-1e4, z_dz=-300.0, wavlen=LAM, dx=DX)
    for k in range(3):
        np.testing.assert_allclose(cv[k], asm.asm_prop(wf, -1e4 - 300.0 * k, LAM, DX), atol=1e-10)


# ---------------------------------------------------------------------------------------
# epilogues
# ---------------------------------------------------------------------------------------
def test_n0f8_rounding_ties_to_even():
    # v*255 = k + 1/2 exactly: round(UInt8, .) rounds half to even
    v = np.array([0.5 / 255, 1.5 / 255, 2.5 / 255, 254.5 / 255, -1.0, 2.0, 1.0, 0.0], np.float64)
    got = asm.n0f8(v)
    assert list(got) == [0, 2, 2, 254, 0, 255, 255, 0]


def test_projection_from_volume_nan_rule():
    vol = np.array([[[np.nan]], [[0.5]], [[0.25]]], np.float32)     # NaN first: sticks
    assert np.isnan(asm.xy_projection_from_vol(vol)[0, 0])
    vol = np.array([[[0.5]], [[np.nan]], [[0.25]]], np.float32)     # NaN later: ignored
    assert asm.xy_projection_from_vol(vol)[0, 0] == np.float32(0.25)
    u8 = np.array([[[200]], [[13]], [[90]]], np.uint8)
    assert asm.xy_projection_from_vol(u8)[0, 0] == np.float32(13) / np.float32(255)


@pytest.mark.parametrize("where", ["front", "dz", "none"])
def test_projection_nan_rules_agree_on_reconstructed_volumes(where):
    """The reference's projection from a volume keeps a NaN of slice 1 and ignores later ones
    (holofunc.jl:268-282); the build takes fminf over all slices (a NaN counts only where every
    slice is NaN). On volumes the reconstruction produces the two rules agree: a NaN anywhere in
    the spectrum times the transfer makes the whole inverse FFT NaN, and the running product
    (holofunc.jl:189-197) carries it into every later slice. A NaN planted in one element of the
    front transfer makes every slice NaN; one in the slice-step transfer every slice after the
    first."""
    rng = np.random.default_rng(11)
    n, slices = 16, 5
    wf = asm.gabor_wavefront((1.0 + 0.2 * rng.standard_normal((n, n))).clip(0).astype(np.float32))
    d = asm.transfer_sqrt_arr(n, 0.6328, 10.0)
    h0 = asm.transfer_reference(-3e3, 0.6328, d).astype(np.complex64)
    hdz = asm.transfer_reference(-150.0, 0.6328, d).astype(np.complex64)
    if where == "front":
        h0[3, 5] = np.nan
    elif where == "dz":
        hdz[7, 2] = np.nan
    vol = asm.reconst_vol(wf, slices, np.float32, "reference", h_front=h0, h_dz=hdz)
    nan_slices = [bool(np.isnan(v).all()) for v in vol]
    assert all(np.isnan(v).all() or not np.isnan(v).any() for v in vol)  # NaN-ness is slice-wide
    assert nan_slices == {"front": [True] * slices, "dz": [False] + [True] * (slices - 1),
                          "none": [False] * slices}[where]
    ref_rule = asm.xy_projection_from_vol(vol)
    fmin_rule = np.fmin.reduce(vol, axis=0).astype(np.float32)
    np.testing.assert_array_equal(ref_rule, fmin_rule)


# ---------------------------------------------------------------------------------------
# f2: cu_2d_pad known-answer test (runtests.jl:362-368)
# ---------------------------------------------------------------------------------------
def test_pad2_known_answer():
    arr = np.array([[1 + 1j, 2 + 1j], [3 + 1j, 4 + 1j]], np.complex64)
    padded = asm.pad2(arr)
    assert padded.shape == (4, 4)
    np.testing.assert_array_equal(padded[1:3, 1:3], arr)
    assert padded[0, 0] == np.complex64(2.5 + 1j)


# ---------------------------------------------------------------------------------------
# A10: Gerchberg-Saxton
# ---------------------------------------------------------------------------------------
def test_gs_plane_wave_fixed_point():
    n = 16
    i1 = np.full((n, n), 0.81, np.float32)
    out = asm.phase_retrieval(i1, i1, 5, z=2000.0, wavlen=LAM, dx=DX)
    np.testing.assert_allclose(out, np.sqrt(np.float32(0.81)) * np.exp(1j * np.angle(out[0, 0])), atol=1e-12)
    # the phase of a plane wave after +z then -z is 0 again
    np.testing.assert_allclose(np.angle(out), 0.0, atol=1e-12)


def test_gs_amplitude_and_error_reduction():
    n = 32
    rng = np.random.default_rng(11)
    psi_true = np.exp(1j * 0.5 * rng.standard_normal((n, n))) * (1 + 0.2 * rng.random((n, n)))
    z = 3000.0
    i1 = (np.abs(psi_true) ** 2).astype(np.float32)
    i2 = (np.abs(asm.asm_prop(psi_true, z, LAM, DX)) ** 2).astype(np.float32)
    hf = asm.natural(asm.transfer_exact(z, LAM, DX, n))
    errs = []
    for it in range(1, 8):
        out = asm.phase_retrieval(i1, i2, it, z=z, wavlen=LAM, dx=DX)
        np.testing.assert_allclose(np.abs(out), np.sqrt(i1.astype(np.float64)), rtol=1e-6)
        errs.append(asm.gs_error(out, i2, hf))
    assert all(b <= a * (1 + 1e-9) for a, b in zip(errs, errs[1:])), errs


def test_gs_one_iteration_bruteforce():
    n = 8
    rng = np.random.default_rng(13)
    i1 = (1 + 0.5 * rng.random((n, n))).astype(np.float32)
    i2 = (1 + 0.5 * rng.random((n, n))).astype(np.float32)
    z = 1500.0
    a1 = np.sqrt(i1.astype(np.float64))
    a2 = np.sqrt(i2.astype(np.float64))
    p2 = _dft_propagate(a1.astype(complex), z)
    p2 = a2 * p2 / np.abs(p2)
    p1 = _dft_propagate(p2, -z)
    p1 = a1 * p1 / np.abs(p1)
    got = asm.phase_retrieval(i1, i2, 1, z=z, wavlen=LAM, dx=DX)
    np.testing.assert_allclose(got, p1, atol=1e-10)


def test_weighted_phase_rms_metric():
    psi = np.exp(1j * np.array([[0.1, -0.2]]))
    ref = np.ones((1, 2), complex)
    assert abs(asm.weighted_phase_rms(psi, ref) - math.sqrt((0.01 + 0.04) / 2)) < 1e-12


# ---------------------------------------------------------------------------------------
# Low-pass filters (frequency_filters.jl)
# ---------------------------------------------------------------------------------------
LPF_CASES = [(0.01, 5.3e-7, 128, 6.5e-6),            # runtests.jl:516-535
             (80000.0 + 100.0 * 1000, LAM, 1024, DX),  # lowpassfilter.md:37
             (2000.0, LAM, 1024, DX),                  # runtests.jl:424 (pr_dist)
             (12.0, 1.0, 8, 1.0)]


@pytest.mark.parametrize("z,lam,n,dx", LPF_CASES)
def test_lpf_cutoff_is_the_aperture_angle_frequency(z, lam, n, dx):
    """Fugal 2009: the highest frequency a point at distance z spreads onto a sensor of
    width n dx is sin(theta)/wavlen with tan(theta) = (n dx / 2) / z; in samples of the
    n-point spectrum (spacing 1/(n dx)) that is n dx sin(theta) / wavlen."""
    theta = math.atan((n * dx / 2) / z)
    want = n * dx * math.sin(theta) / lam
    assert abs(asm.lpf_cutoff(z, lam, n, dx) / want - 1) < 1e-12


def test_rectangle_filter_hand_mask():
    """8x8, cutoff radius 2.5 samples: the pass band is the 21 lattice points with
    dr^2 + dc^2 <= 6 around numpy [2, 3] (Julia x = n/2 - 1 on dim 2, y = n/2 on dim 1)."""
    n, lam, dx = 8, 1.0, 1.0
    z = math.sqrt((64.0 / 2.5) ** 2 - 64.0) / 2.0  # cutoff 64 / sqrt(4 z^2 + 64) = 2.5
    assert abs(asm.lpf_cutoff(z, lam, n, dx) - 2.5) < 1e-12
    want = np.zeros((n, n), np.float32)
    for dr in range(-3, 4):
        for dc in range(-3, 4):
            if dr * dr + dc * dc <= 6:
                want[2 + dr, 3 + dc] = 1.0
    assert want.sum() == 21
    got = asm.rectangle_filter(z, lam, n, dx)
    np.testing.assert_array_equal(got, want)


def test_rectangle_filter_area():
    z, lam, n, dx = LPF_CASES[1]
    maxi = asm.lpf_cutoff(z, lam, n, dx)
    area = asm.rectangle_filter(z, lam, n, dx).sum()
    assert abs(area - math.pi * maxi * maxi) < 2 * math.pi * maxi  # lattice count vs disk


@pytest.mark.parametrize("z,lam,n,dx", LPF_CASES)
def test_super_gaussian_half_value_at_cutoff(z, lam, n, dx):
    maxi = asm.lpf_cutoff(z, lam, n, dx)
    sig = asm.super_gaussian_sigma(z, lam, n, dx)
    c = n / 2 - 1  # 1-based centre on both axes
    assert asm.super_gaussian_at(c, c, n, dx, sig) == 1.0
    for x, y in ((c + maxi, c), (c, c - maxi), (c + maxi / math.sqrt(2), c + maxi / math.sqrt(2))):
        assert abs(asm.super_gaussian_at(x, y, n, dx, sig) - 0.5) < 1e-12


def test_super_gaussian_grid_symmetry_and_monotonicity():
    z, lam, n, dx = LPF_CASES[2]
    f = asm.super_gaussian_filter(z, lam, n, dx)
    np.testing.assert_array_equal(f, f.T)
    c = n // 2 - 2  # numpy index of the centre
    assert f[c, c] == 1.0
    row = f[c, c:]
    assert np.all(np.diff(row.astype(np.float64)) <= 0)


def _dft2(x, sign):
    ny, nx = x.shape
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    out = np.zeros((ny, nx), np.complex128)
    for ky in range(ny):
        for kx in range(nx):
            out[ky, kx] = np.sum(x * np.exp(sign * 2j * np.pi * (kx * ii / nx + ky * jj / ny)))
    return out


@pytest.mark.parametrize("shape", [(8, 8), (6, 10), (5, 3)])
def test_apply_low_pass_filter_bruteforce(shape):
    """Natural-order spectrum sample (ky, kx) meets filter element
    ((ky + ny//2) % ny, (kx + nx//2) % nx): the fftshift / ifftshift pair, odd sizes included."""
    ny, nx = shape
    rng = np.random.default_rng(ny * 100 + nx)
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    lpf = rng.random(shape)
    spec = _dft2(psi, -1)
    for ky in range(ny):
        for kx in range(nx):
            spec[ky, kx] *= lpf[(ky + ny // 2) % ny, (kx + nx // 2) % nx]
    want = _dft2(spec, +1) / (nx * ny)
    np.testing.assert_allclose(asm.apply_low_pass_filter(psi, lpf), want, atol=1e-12)


def test_apply_low_pass_filter_closed_forms():
    ny, nx = 16, 12
    rng = np.random.default_rng(5)
    psi = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    np.testing.assert_allclose(asm.apply_low_pass_filter(psi, np.ones((ny, nx))), psi, atol=1e-13)
    dc = np.zeros((ny, nx))
    dc[ny // 2, nx // 2] = 1.0
    np.testing.assert_allclose(asm.apply_low_pass_filter(psi, dc), np.full((ny, nx), psi.mean()), atol=1e-13)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    kx, ky = 3, -5
    wave = np.exp(2j * np.pi * (kx * ii / nx + ky * jj / ny))
    lpf = rng.random((ny, nx))
    gain = lpf[(ky % ny + ny // 2) % ny, (kx + nx // 2) % nx]
    np.testing.assert_allclose(asm.apply_low_pass_filter(wave, lpf), gain * wave, atol=1e-13)


# ---------------------------------------------------------------------------------------
# Ingest (utils.jl:23-25, 217-317; preprocessings.md:21-23)
# ---------------------------------------------------------------------------------------
def test_n0f8_to_f32_exhaustive():
    from fractions import Fraction

    got = asm.n0f8_to_f32(np.arange(256))
    for v in range(256):
        exact = Fraction(v, 255)
        f = float(got[v])
        # correctly rounded: no other float32 is closer to v/255
        lo = float(np.nextafter(np.float32(f), np.float32(-1)))
        hi = float(np.nextafter(np.float32(f), np.float32(2)))
        d = abs(Fraction(f) - exact)
        assert d <= abs(Fraction(lo) - exact) and d <= abs(Fraction(hi) - exact), v


def test_ingest_scalar_reevaluation():
    rng = np.random.default_rng(3)
    u8 = np.arange(256, dtype=np.uint8).reshape(16, 16)
    bkg = rng.random((16, 16))
    off = 64 / 255
    got = asm.ingest_u8(u8, bkg, off)
    for j in range(16):
        for i in range(16):
            f = float(np.float32(int(u8[j, i]) / 255))  # v/255 has no rounding tie at 24 bits
            want = np.float32((f - float(bkg[j, i])) + off)
            assert got[j, i] == want
    np.testing.assert_array_equal(asm.ingest_u8(u8), asm.n0f8_to_f32(u8))


def test_background_mode_known_answer():
    """runtests.jl:125-138, Julia matrices [10 20; 30 40], [10 25; 30 45], [10 20; 30 40] in
    the reference layout (tensor [j, i] = Julia [i+1, j+1])."""
    imgs = [np.array([[10, 30], [20, 40]], np.uint8), np.array([[10, 30], [25, 45]], np.uint8),
            np.array([[10, 30], [20, 40]], np.uint8)]
    bg = asm.background_mode(np.stack(imgs))
    assert bg.dtype == np.float64 and bg.shape == (2, 2)
    for (j, i), v in {(0, 0): 10, (1, 0): 20, (0, 1): 30, (1, 1): 40}.items():
        assert abs(bg[j, i] - v / 255) < 1e-6


def test_background_mode_counter_bruteforce():
    from collections import Counter

    rng = np.random.default_rng(11)
    stack = rng.integers(0, 4, size=(9, 7, 5)).astype(np.uint8) * 60  # frequent ties
    bg = asm.background_mode(stack)
    for j in range(7):
        for i in range(5):
            c = Counter(int(v) for v in stack[:, j, i])
            top = max(c.values())
            assert bg[j, i] == min(v for v, n in c.items() if n == top) / 255.0


def test_background_mean_sequential_and_fsum():
    rng = np.random.default_rng(12)
    stack = rng.integers(0, 256, size=(13, 4, 6)).astype(np.uint8)
    bg = asm.background_mean(stack)
    for j in range(4):
        for i in range(6):
            acc = 0.0
            for m in range(13):
                acc = acc + float(np.float32(int(stack[m, j, i]) / 255))
            assert bg[j, i] == acc / 13
            exact = math.fsum(float(np.float32(int(stack[m, j, i]) / 255)) for m in range(13)) / 13
            assert abs(bg[j, i] - exact) <= 1e-15 * max(exact, 1e-300)


def test_n0f8_reciprocal_product_rounds_like_the_quotient():
    """The device conversion rounds the FP64 product v * fl(1/255) to Float32; for every byte
    this equals the correctly rounded quotient the oracle uses."""
    v = np.arange(256)
    prod = (v.astype(np.float64) * (1.0 / 255.0)).astype(np.float32)
    np.testing.assert_array_equal(prod, asm.n0f8_to_f32(v))


def test_running_projection_equals_per_slice_fp64():
    """The long-stack FP64 oracle (running complex128 product, threaded FFTs) against the
    per-slice FP64 evaluation, non-square, with slices on both sides of z = 0."""
    rng = np.random.default_rng(5)
    holo = (1.0 + 0.3 * rng.standard_normal((64, 96))).clip(0.0).astype(np.float32)
    wf = asm.gabor_wavefront(holo)
    kw = dict(z_front=-3000.0, z_dz=75.0, wavlen=0.6328, dx=10.0)
    fast = asm.reconst_xyprojection_running(wf, 60, workers=2, **kw)
    slow = asm.reconst_xyprojection(wf, 60, "exact", **kw)
    assert asm.rel_l2(fast, slow) < 1e-10

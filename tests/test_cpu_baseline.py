"""The CPU baseline (the reference's op sequence on torch CPU) agrees with the oracle's
reference-emulation mode; the seeded generator is deterministic. CPU only."""
import numpy as np

from oracle import asm, cpu_baseline

LAM, DX = 0.6328, 10.0


def test_cpu_restatement_matches_reference_emulation():
    rng = np.random.default_rng(17)
    holo = (1 + 0.3 * rng.standard_normal((96, 96))).clip(0).astype(np.float32)
    d32 = asm.transfer_sqrt_arr(96, LAM, DX)
    h0, hdz = asm.transfer_reference(-1e4, LAM, d32), asm.transfer_reference(-1e3, LAM, d32)
    got = cpu_baseline.reconst_xyprojection_cpu(holo, h0, hdz, 9)
    want = asm.reconst_xyprojection(asm.gabor_wavefront(holo), 9, "reference", h_front=h0, h_dz=hdz)
    assert asm.rel_l2(got, want) < 1e-5


def test_measure_reports_cores_and_sample():
    rng = np.random.default_rng(1)
    holo = (1 + 0.1 * rng.standard_normal((64, 64))).astype(np.float32)
    r = cpu_baseline.measure(holo, -1e4, -200.0, LAM, DX, 50, budget_s=0.3, threads=2)
    assert r["cores"] == 2 and r["kind"] == "port" and r["value"] > 0 and "of 50 slices" in r["sample"]


def test_generator_is_seeded():
    from particleholography import synth

    a = synth.hologram(5, 30, 64, zmin=1e4, zmax=5e4)
    b = synth.hologram(5, 30, 64, zmin=1e4, zmax=5e4)
    c = synth.hologram(6, 30, 64, zmin=1e4, zmax=5e4)
    assert a.dtype == np.float32 and a.shape == (64, 64) and a.min() >= 0
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)
    i1, i2 = synth.gs_pair(5, 30, 64, 2000.0, zmin=1e4, zmax=5e4)
    assert i1.shape == i2.shape == (64, 64) and not np.array_equal(i1, i2)

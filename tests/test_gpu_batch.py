"""Batched reconstruction of holograms of uneven sizes (one column-pass and one row-pass launch
per round over all holograms of a size) and the tile autotuner, through the C ABI."""
import numpy as np
import pytest
import torch

from oracle import asm

pytestmark = pytest.mark.gpu

LAM, DX = 0.6328, 10.0


def _holo(nx, ny, seed):
    rng = np.random.default_rng(seed)
    return (1.0 + 0.2 * rng.standard_normal((ny, nx))).clip(0.0).astype(np.float32)


def _single(ph, h, z0, dz, slices, nx, ny):
    import ctypes
    from particleholography import _lib

    proj = torch.empty((ny, nx), dtype=torch.float32, device=h.device)
    f = _lib.holo_tf(None, z0, LAM, DX, 1)
    d = _lib.holo_tf(None, dz, LAM, DX, 1)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.load().holo_reconst_xyprojection_holo(_lib.handle(0).ptr, st, nx, ny, ctypes.c_void_p(h.data_ptr()),
                                                          ctypes.byref(f), ctypes.byref(d), slices,
                                                          ctypes.c_void_p(proj.data_ptr())))
    return proj


def test_batch_uneven_sizes_equal_single_calls(gpu):
    import particleholography as ph
    from particleholography.batch import ReconJob, reconst_batch

    cases = [(512, 512, 17, -1e4, -700.0), (768, 768, 40, -2e4, -300.0), (1024, 1024, 33, -1e4, -200.0),
             (1280, 1024, 8, -5e4, -1000.0), (256, 256, 25, -8e4, -100.0), (1024, 1024, 60, -3e4, -150.0),
             (512, 512, 3, -1e4, -50.0)]
    jobs, holos = [], []
    for i, (nx, ny, s, z0, dz) in enumerate(cases):
        h = torch.from_numpy(_holo(nx, ny, 100 + i)).to(gpu)
        holos.append(h)
        jobs.append(ReconJob(h, z0, dz, s, LAM, DX))
    outs = reconst_batch(jobs)
    for (nx, ny, s, z0, dz), h, out in zip(cases, holos, outs):
        want = _single(ph, h, z0, dz, s, nx, ny)
        np.testing.assert_array_equal(out.cpu().numpy(), want.cpu().numpy())


def test_batch_configured_mix_at_its_depths(gpu):
    """The batched workload's shapes at its depths: every configured side, a 2048^2 hologram
    included, with 128-1024 slices over the 1e4-2e5 um range in one call. Each result equals its
    single call bit for bit (shipped tiles: 32-slice rounds at 1024^2, 25 at 1280x1024, 64 at
    2048^2, so every depth crosses several rounds); the 512^2 and 768^2 results also against the
    FP64 oracle."""
    import particleholography as ph
    from particleholography.batch import ReconJob, reconst_batch

    cases = [(512, 512, 128), (768, 768, 300), (1024, 1024, 1024), (1280, 1024, 257), (2048, 2048, 140),
             (512, 512, 777)]
    jobs, holos = [], []
    for i, (nx, ny, s) in enumerate(cases):
        h = _holo(nx, ny, 700 + i)
        holos.append(h)
        jobs.append(ReconJob(torch.from_numpy(h).to(gpu), -1e4, -(2e5 - 1e4) / s, s, LAM, DX))
    outs = [o.cpu().numpy() for o in reconst_batch(jobs)]
    for (nx, ny, s), j, out in zip(cases, jobs, outs):
        want = _single(ph, j.holo, j.z_front, j.z_dz, s, nx, ny).cpu().numpy()
        np.testing.assert_array_equal(out, want, err_msg=f"{nx}x{ny} x {s}")
    for (nx, ny, s), h, j, out in zip(cases, holos, jobs, outs):
        if nx > 768 or s > 300:
            continue
        want = asm.reconst_xyprojection(asm.gabor_wavefront(h), s, "exact", z_front=j.z_front, z_dz=j.z_dz,
                                        wavlen=LAM, dx=DX)
        err = asm.rel_l2(out, want)
        print(f"batched {nx}x{ny} x {s} to z = 2e5 um: rel L2 vs FP64 oracle {err:.2e}")
        assert err < 1e-4


def test_batch_vs_fp64_oracle_and_chunking(gpu):
    from particleholography.batch import ReconJob, reconst_batch

    rng = np.random.default_rng(5)
    jobs, hs = [], []
    for i in range(37):  # more than one 32-hologram launch
        n = 64 if i % 2 == 0 else 96
        s = int(rng.integers(1, 9))
        h = _holo(n, n, 300 + i)
        hs.append((h, s))
        jobs.append(ReconJob(torch.from_numpy(h).to(gpu), -1e4 - 100.0 * i, -250.0, s, LAM, DX))
    outs = reconst_batch(jobs)
    for (h, s), j, out in zip(hs, jobs, outs):
        want = asm.reconst_xyprojection(asm.gabor_wavefront(h), s, "exact", z_front=j.z_front, z_dz=j.z_dz,
                                        wavlen=LAM, dx=DX)
        assert asm.rel_l2(out.cpu().numpy(), want) < 1e-4


def test_batch_rejects_array_transfers_and_bad_jobs(gpu):
    import ctypes
    from particleholography import _lib

    arr = (_lib.holo_job * 1)()
    arr[0].nx = arr[0].ny = 64
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    with pytest.raises(AssertionError, match="NULL"):
        _lib.check(_lib.load().holo_reconst_batch(_lib.handle(0).ptr, st, arr, 1))
    assert _lib.load().holo_reconst_batch(_lib.handle(0).ptr, st, arr, 0) == 0


def test_autotune_keeps_results_correct(gpu):
    import particleholography as ph
    from particleholography import batch

    n = 256
    choice = batch.autotune(n, n, LAM, DX, save=False)
    assert choice["batch"] >= 1 and choice["cols_per_wg"] in (1, 2, 4, 8, 16)
    # rounds are whole numbers of generation groups: no workgroup runs a short remainder group
    assert choice["batch"] % choice["spg_gen"] == 0 or choice["batch"] < choice["spg_gen"], choice
    eff = batch.get_tuning_for_size(n, n, 1000)
    assert eff["batch"] == choice["batch"] and eff["spg_gen"] == choice["spg_gen"]
    # the reported choice is the tiles in force, defaults resolved (also when the incumbent,
    # the default round of a size without an entry, keeps its place)
    assert batch.get_tuning_for_size(n, n, 0) == choice
    try:
        h = _holo(n, n, 9)
        ht = torch.from_numpy(h).to(gpu)
        got = _single(ph, ht, -2e4, -300.0, 40, n, n).cpu().numpy()
        want = asm.reconst_xyprojection(asm.gabor_wavefront(h), 40, "exact", z_front=-2e4, z_dz=-300.0,
                                        wavlen=LAM, dx=DX)
        assert asm.rel_l2(got, want) < 1e-4
    finally:
        batch.set_tuning_for_size(n, n, {})


def _gs_pair(n, seed, z=2000.0):
    rng = np.random.default_rng(seed)
    psi = np.exp(1j * 0.7 * rng.standard_normal((n, n))) * (1 + 0.3 * rng.random((n, n)))
    i1 = (np.abs(psi) ** 2).astype(np.float32)
    i2 = (np.abs(asm.asm_prop(psi, z, LAM, DX)) ** 2).astype(np.float32)
    return i1, i2


def test_gs_batch_equals_single_calls(gpu):
    import particleholography as ph
    from particleholography.batch import GSJob, phase_retrieval_batch

    specs = [(64, 5, 2000.0), (128, 12, 1500.0), (64, 1, 3000.0), (128, 7, 2000.0), (96, 0, 2000.0), (64, 9, 2500.0)]
    jobs, ins = [], []
    for k, (n, it, z) in enumerate(specs):
        i1, i2 = _gs_pair(n, 40 + k, z)
        t1, t2 = torch.from_numpy(i1).to(gpu), torch.from_numpy(i2).to(gpu)
        ins.append((t1, t2, n, it, z))
        jobs.append(GSJob(t1, t2, z, it, LAM, DX))
    # one pair with array-backed transfers in the reference's arithmetic
    n = 64
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    hf, hb = ph.cu_transfer(1800.0, n, LAM, dsq, precision="reference"), ph.cu_transfer(-1800.0, n, LAM, dsq, precision="reference")
    a1, a2 = _gs_pair(n, 99, 1800.0)
    ta1, ta2 = torch.from_numpy(a1).to(gpu), torch.from_numpy(a2).to(gpu)
    jobs.append(GSJob(ta1, ta2, 1800.0, 6, fwd=hf.data, bwd=hb.data))
    outs = phase_retrieval_batch(jobs)
    for (t1, t2, m, it, z), out in zip(ins, outs):
        dsq_m = ph.cu_transfer_sqrt_arr(m, LAM, DX)
        want = ph.cu_phase_retrieval_holo(t1, t2, ph.cu_transfer(z, m, LAM, dsq_m), ph.cu_transfer(-z, m, LAM, dsq_m), it, m)
        np.testing.assert_array_equal(out.cpu().numpy(), want.data.cpu().numpy())
    want = ph.cu_phase_retrieval_holo(ta1, ta2, hf, hb, 6, n)
    np.testing.assert_array_equal(outs[-1].cpu().numpy(), want.data.cpu().numpy())
    ref = asm.phase_retrieval(a1, a2, 6, hf.data.cpu().numpy(), hb.data.cpu().numpy())
    assert asm.weighted_phase_rms(outs[-1].cpu().numpy(), ref) < 1e-3


def test_gs_batch_many_pairs_vs_oracle(gpu):
    from particleholography.batch import GSJob, phase_retrieval_batch

    jobs, pairs = [], []
    for k in range(34):  # two launches per pass (32 pairs max per launch)
        i1, i2 = _gs_pair(32, 200 + k)
        pairs.append((i1, i2))
        jobs.append(GSJob(torch.from_numpy(i1).to(gpu), torch.from_numpy(i2).to(gpu), 2000.0, 3))
    outs = phase_retrieval_batch(jobs)
    for (i1, i2), out in zip(pairs, outs):
        ref = asm.phase_retrieval(i1, i2, 3, z=2000.0, wavlen=LAM, dx=DX)
        assert asm.weighted_phase_rms(out.cpu().numpy(), ref) < 1e-3


@pytest.mark.parametrize("nx,ny,iters", [(96, 64, 4), (1280, 1024, 3)])
def test_gs_batch_non_square_vs_oracle(gpu, nx, ny, iters):
    """The batched entry point takes (nx, ny) pairs (the reference's GS is square-only)."""
    from particleholography.batch import GSJob, phase_retrieval_batch

    rng = np.random.default_rng(nx + ny)
    psi = np.exp(1j * 0.7 * rng.standard_normal((ny, nx))) * (1 + 0.3 * rng.random((ny, nx)))
    i1 = (np.abs(psi) ** 2).astype(np.float32)
    i2 = (np.abs(asm.asm_prop(psi, 2000.0, LAM, DX)) ** 2).astype(np.float32)
    out = phase_retrieval_batch([GSJob(torch.from_numpy(i1).to(gpu), torch.from_numpy(i2).to(gpu), 2000.0, iters)])[0]
    assert tuple(out.shape) == (ny, nx)
    ref = asm.phase_retrieval(i1, i2, iters, z=2000.0, wavlen=LAM, dx=DX)
    np.testing.assert_allclose(np.abs(out.cpu().numpy()), np.sqrt(i1), rtol=2e-6)
    assert asm.weighted_phase_rms(out.cpu().numpy(), ref) < 1e-3


def test_gs_batch_graphs_equal_direct_launches(gpu):
    """The chunk-uniform iterations of batched GS run as cached hipGraphs: every replay (three
    calls, with direct-launch calls, other device work and synchronizations between them, and
    new intensities each time) equals the directly launched result bit for bit."""
    from particleholography import _lib
    from particleholography.batch import GSJob, phase_retrieval_batch

    hd = _lib.handle(0)
    specs = [(128, 37, 2000.0), (128, 21, 1500.0), (128, 40, 2500.0)]  # pairs with different counts
    try:
        for rep in range(3):
            jobs = []
            for k, (n, it, z) in enumerate(specs):
                i1, i2 = _gs_pair(n, 300 + 10 * rep + k, z)
                jobs.append(GSJob(torch.from_numpy(i1).to(gpu), torch.from_numpy(i2).to(gpu), z, it, LAM, DX))
            hd.set_gs_graphs(False)
            direct = [o.clone() for o in phase_retrieval_batch(jobs)]
            torch.cuda.synchronize()
            _ = torch.randn(1 << 20, device=gpu).sum().item()  # unrelated device work between replays
            hd.set_gs_graphs(True)
            graphed = phase_retrieval_batch(jobs)
            torch.cuda.synchronize()
            for a, b in zip(direct, graphed):
                np.testing.assert_array_equal(b.cpu().numpy(), a.cpu().numpy())
    finally:
        hd.set_gs_graphs(True)


def test_workspace_size_is_queryable(gpu):
    """holo_workspace_size reports the handle's device workspace, which holds at least the
    spectrum and one round of column-transformed slices after a reconstruction."""
    import particleholography as ph
    from particleholography import _lib

    n, slices = 256, 40
    dsq = ph.cu_transfer_sqrt_arr(n, LAM, DX)
    wf = ph.cu_gabor_wavefront(torch.from_numpy(_holo(n, n, 3)).to(gpu))
    ph.cu_get_reconst_xyprojection(wf, ph.cu_transfer(-1e4, n, LAM, dsq), ph.cu_transfer(-200.0, n, LAM, dsq), slices)
    torch.cuda.synchronize()
    h = _lib.handle(0)
    from particleholography.batch import get_tuning_for_size

    rounds = get_tuning_for_size(n, n, slices)["batch"]
    assert h.workspace_size() >= (1 + rounds) * n * n * 8


def test_gs_batch_chunks_leave_each_pair_unchanged(gpu):
    """At 1024^2 a batched GS call runs 16 pairs per launch (their fields and amplitudes fill
    256 MiB): 18 pairs run as chunks of 16 and 2, and every result equals the pair computed
    alone through the same entry point, bit for bit."""
    from particleholography.batch import GSJob, phase_retrieval_batch

    n, z, iters = 1024, 2000.0, 3
    pairs = [_gs_pair(n, 300 + k % 3, z) for k in range(3)]
    ts = [(torch.from_numpy(a).to(gpu), torch.from_numpy(b).to(gpu)) for a, b in pairs]
    jobs = [GSJob(ts[k % 3][0], ts[k % 3][1], z + 10.0 * k, iters, LAM, DX) for k in range(18)]
    outs = phase_retrieval_batch(jobs)
    for k in (0, 7, 15, 16, 17):
        alone = phase_retrieval_batch([jobs[k]])[0]
        np.testing.assert_array_equal(outs[k].cpu().numpy(), alone.cpu().numpy(), err_msg=f"pair {k}")

#!/usr/bin/env python3
"""Benchmark: z-slices/s of the fused angular-spectrum reconstruction + xy min-projection.

Workload (the 1024^2 GPU configuration): one 1024 x 1024 Float32 inline hologram of a seeded
1000-disk scene (depths 1e4-2e5 um, lambda = 0.6328 um, pitch 10 um), reconstructed at 1000
slices from z0 = 1e4 um in steps of 200 um; complex-FP32 FFTs, FP64-wrapped transfer phase,
fused multiply / inverse FFT / |.|^2 / running min. One step = one full reconstruction
(sqrt + forward spectrum + 1000 slices) of the device-resident hologram.

Prints one JSON line. --gpus N > 1 is refused: the library runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "particleholography.jl_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "z-slices/s at 1024² CF32, % of HBM roofline; p50/p99 per-hologram latency"
HBM_PEAK_GBS = 8000.0
LAM, DX = 0.6328, 10.0
N, SLICES, Z0, DZ, NDISKS, SEED = 1024, 1000, 1e4, 200.0, 1000, 102
# Traffic models, bytes per pixel and slice:
#   fused (this design, the roofline's algorithmic bytes): pass A writes the column-transformed
#   slice (8) and reads the spectrum once per slice group of spg_gen slices (8 / spg_gen); pass B
#   reads the slice (8) and read-modify-writes the projection once per group of spg_out slices
#   (8 / spg_out). 17 B/px/slice at the default 16 / 16.
#   reference (the per-slice sequence of holofunc.jl:259-263 as one fused pass pair would see it
#   if the spectrum were re-read every slice): 16 + 16 = 32 B/px/slice. Reported as
#   "reference traffic", never as a roofline.
REFERENCE_BYTES_PER_PX = 32.0


def pass_a_bytes(spg_gen):
    return 8.0 + 8.0 / spg_gen


def fused_bytes(spg_gen, spg_out):
    return 16.0 + 8.0 / spg_gen + 8.0 / spg_out


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_steps(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(st)
    for _ in range(steps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, e0.elapsed_time(e1) / 1e3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the plain-path and GS figures")
    ap.add_argument("--batch-holograms", type=int, default=64, help="holograms in the batched workload")
    ap.add_argument("--serve-requests", type=int, default=5000,
                    help="requests in the served Poisson stream (the full stream has 5000)")
    ap.add_argument("--profile-gs", action="store_true",
                    help="for rocprofv3: Gerchberg-Saxton at 1024^2 (one warm-up call, then --steps "
                         "calls of 20 iterations: eager launches, then the hipGraph form); no extras")
    ap.add_argument("--profile-run", action="store_true",
                    help="for rocprofv3: one warm-up + --steps steps, then a 256 MiB device copy "
                         "(a kernel of known traffic to scale FETCH_SIZE / WRITE_SIZE); no extras")
    ap.add_argument("--profile-gs-batch", action="store_true",
                    help="for rocprofv3: batched GS at 1024^2, 32 pairs x 20 iterations per call (one warm-up "
                         "call, then --steps calls); no other work")
    ap.add_argument("--profile-detect", action="store_true",
                    help="for rocprofv3: the particle-detection extra (1024^2 x 1000 N0f8 volume, "
                         "labels and rectangles of every slice, --steps calls); no other work")
    ap.add_argument("--profile-ingest", action="store_true",
                    help="for rocprofv3: 8-bit ingest of 64 1024^2 holograms and the background mode / mean of a "
                         "256-image stack (--steps calls of each after warm-up); no other work")
    ap.add_argument("--profile-batch", action="store_true",
                    help="for rocprofv3: the 64-hologram batched reconstruction (shipped tiles, "
                         "--steps calls after one warm-up); no other work")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (a reconstruction "
              "would shard by z-range and merge the projections with one min-reduction); run with "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import particleholography as ph
    from particleholography import _lib, synth

    dev = torch.device("cuda:0")
    log(f"bench: generating the {N}^2 hologram (seed {SEED}, {NDISKS} disks)")
    holo_np = synth.hologram(SEED, NDISKS, N, zmin=1e4, zmax=2e5, wavlen=LAM, dx=DX)
    holo = torch.from_numpy(holo_np).to(dev)
    d_sqr = ph.cu_transfer_sqrt_arr(N, LAM, DX)
    h0 = ph.cu_transfer(-Z0, N, LAM, d_sqr)
    hdz = ph.cu_transfer(-DZ, N, LAM, d_sqr)
    proj = torch.empty((N, N), dtype=torch.float32, device=dev)
    hd = _lib.handle(0)

    def step():
        ph.reconst_xyprojection_holo(holo, h0, hdz, SLICES, out=proj)

    if args.profile_detect:
        r = bench_detect(ph, torch, dev, reps=args.steps)
        log("bench: detection profile run done")
        print(json.dumps({"profile_detect": True, "steps": args.steps, **r}), flush=True)
        return
    if args.profile_ingest:
        r = bench_ingest(ph, torch, dev, reps=args.steps)
        log("bench: ingest profile run done")
        print(json.dumps({"profile_ingest": True, "steps": args.steps, **r}), flush=True)
        return
    if args.profile_batch:
        r = bench_batched(torch, dev, args.batch_holograms, autotune=False, reps=args.steps)
        log("bench: batched profile run done")
        print(json.dumps({"profile_batch": True, "steps": args.steps, "slices_1024eq_per_s": r["slices_1024eq_per_s"],
                          "algorithmic_GBs": r["algorithmic_GBs"]}), flush=True)
        return

    if args.profile_gs_batch:
        from particleholography.batch import GSJob, phase_retrieval_batch

        pairs = [synth.gs_pair_device(103000 + k, 500, N, N, separation=2000.0) for k in range(4)]
        jobs = [GSJob(pairs[k % 4][0], pairs[k % 4][1], 2000.0, 20, LAM, DX) for k in range(32)]
        for _ in range(1 + args.steps):
            outs = phase_retrieval_batch(jobs)
        torch.cuda.synchronize()
        log("bench: batched GS profile run done")
        print(json.dumps({"profile_gs_batch": True, "steps": args.steps,
                          "finite": bool(all(torch.isfinite(o).all() for o in outs))}), flush=True)
        return

    if args.profile_gs:
        i2 = torch.from_numpy(synth.gs_pair(103, 500, N, 2000.0)[1]).to(dev)
        hf, hb = ph.cu_transfer(2000.0, N, LAM, d_sqr), ph.cu_transfer(-2000.0, N, LAM, d_sqr)
        out = None
        for use_graph in (False, True):
            for _ in range(1 + args.steps):
                out = ph.cu_phase_retrieval_holo(holo, i2, hf, hb, 20, N, use_graph=use_graph)
            torch.cuda.synchronize()
        log("bench: GS profile run done")
        print(json.dumps({"profile_gs": True, "steps": args.steps, "finite": bool(torch.isfinite(out.data).all())}), flush=True)
        return

    if args.profile_run:
        timed_steps(step, args.steps, 1, torch)
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        log("bench: profile run done (copy kernel: 268435456 B read, 268435456 B written)")
        print(json.dumps({"profile_run": True, "steps": args.steps, "copy_bytes": 268435456}), flush=True)
        return

    log(f"bench: {args.warmup} warm-up + {args.steps} timed steps")
    wall, gpu_s = timed_steps(step, args.steps, args.warmup, torch)
    slices_per_s = args.steps * SLICES / wall
    ms_per_step = wall / args.steps * 1e3
    log(f"bench: {slices_per_s:.0f} slices/s ({ms_per_step:.3f} ms/step, events {gpu_s / args.steps * 1e3:.3f} ms)")

    # per-pass HIP-event timing over one more step (same kernels, same stream)
    hd.set_profiling(True)
    for p in range(5):
        hd.profile_read(p, reset=True)
    step()
    torch.cuda.synchronize()
    passes = {}
    for pid, name in ((_lib.HOLO_PASS_RECON_COLS, "recon_cols"), (_lib.HOLO_PASS_RECON_ROWS, "recon_rows"),
                      (_lib.HOLO_PASS_FWD_FFT, "forward_fft")):
        ms, n, units = hd.profile_read(pid)
        passes[name] = {"ms_total": ms, "launches": n, "slices": units}
    hd.set_profiling(False)
    from particleholography import batch as _batch

    tune = _batch.get_tuning_for_size(N, N, SLICES)
    # pass A (the column pass that generates H_k, k_recon_cols) is the dominant kernel
    d = passes["recon_cols"]
    launches = max(d["launches"], 1)
    avg_s = d["ms_total"] * 1e-3 / launches
    alg_bytes = pass_a_bytes(tune["spg_gen"]) * N * N * d["slices"] / launches
    achieved = alg_bytes / avg_s / 1e9
    traffic, dram = None, None
    # the pass-A instantiation in use: <E, L, generated H, exchange variant>; the size's col_db
    # 1-3 name the variant, -1 and 0 (two buffers where occupancy allows) leave it to the library
    # (a fifth argument is the tile width when it is compiled in)
    xb = tune["col_db"] if tune["col_db"] in (1, 2, 3) else None
    wc = ", 8" if xb in (2, 3) and tune["cols_per_wg"] == 8 else ""
    kname = f"k_recon_cols<16, {N}, 0, {xb}{wc}>" if xb is not None else f"k_recon_cols<16, {N}, 0, *>"
    xb_pat = str(xb) if xb is not None else r"-?\d+"
    kpat = re.compile(rf"k_recon_cols<16, {N}, 0, {xb_pat}(, \d+)?>")
    pmc_file = ROOT / "profiles" / "pmc_traffic_kernels.json"
    if pmc_file.exists():  # HBM bytes per launch from the committed rocprofv3 PMC runs
        pmc = json.loads(pmc_file.read_text())
        for kn, v in pmc["kernels"].items():
            if kpat.search(kn):
                traffic = int(v["fetch_bytes_mean"] + v["write_bytes_mean"])
                dram = v.get("dram_read_requests_mean")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": pmc_file.name if traffic else None,
                "kernel": f"{kname} (pass A)", "avg_launch_ms": round(avg_s * 1e3, 4),
                "bytes_per_launch": alg_bytes,
                "bytes_model": f"8 B/px/slice written + the 8 B/px spectrum read once per {tune['spg_gen']}-slice group",
                "counter_frac": round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                "dram_read_requests_per_launch": dram,
                "reference_traffic_GBs": round(REFERENCE_BYTES_PER_PX / 2 * N * N * d["slices"] / launches / avg_s / 1e9, 1)}
    step_bytes = fused_bytes(tune["spg_gen"], tune["spg_out"])
    extras = {"passes": passes, "tuning": tune,
              "step_algorithmic_GBs": round(step_bytes * N * N * slices_per_s / 1e9, 1),
              "step_bytes_per_px_slice": step_bytes,
              "step_reference_traffic_GBs": round(REFERENCE_BYTES_PER_PX * N * N * slices_per_s / 1e9, 1)}

    if not args.no_extras:
        # the plain one-slice-at-a-time GPU path: one pass-A + pass-B launch pair per slice
        hd.set_tuning(1, 1, 1, 0, 0)
        try:
            pw, _ = timed_steps(step, 2, 1, torch)
        finally:
            hd.set_tuning()
        plain = 2 * SLICES / pw
        extras["plain_path_slices_per_s"] = round(plain, 1)
        extras["speedup_vs_plain"] = round(slices_per_s / plain, 2)
        extras["gs_1024"] = bench_gs(ph, synth, torch, dev, d_sqr)
        extras["gs_batched_1024"] = bench_gs_batched(torch, dev)
        extras["volumes_1024"] = bench_volumes(ph, torch, holo)
        extras["lpf_1024"] = bench_lpf(ph, torch, holo)
        extras["ingest_1024"] = bench_ingest(ph, torch, dev)
        extras["detect_reference_holograms"] = bench_detect(ph, torch, dev)
        extras["batched_64"] = bench_batched(torch, dev, args.batch_holograms)
        extras["hbm_practical"] = bench_hbm_practical(torch, dev)
        if args.serve_requests > 0:
            extras["serving"] = bench_serving(torch, args.serve_requests)
            # the same service on a stream of 1024^2 requests only (the size the speedup target
            # is stated at)
            extras["serving_1024"] = bench_serving(torch, max(1, args.serve_requests // 5), [(N, N)])

    cpu, parity = None, None
    if not args.no_cpu_baseline:
        from oracle import asm, cpu_baseline

        log("bench: CPU baseline (reference op sequence, torch CPU complex64)")
        cpu = cpu_baseline.measure(holo_np, -Z0, -DZ, LAM, DX, SLICES, budget_s=15.0)
        cpu["value"] = round(cpu["value"], 2)
        # the CPU configuration: 512^2, seeded 200-disk scene, 100 slices at 1e3 um from 1e4 um
        small_holo = synth.hologram(101, 200, 512, zmin=1e4, zmax=1e5, wavlen=LAM, dx=DX)
        small = cpu_baseline.measure(small_holo, -1e4, -1e3, LAM, DX, 100, budget_s=6.0)
        cpu["seeded_512x512x100"] = {"value": round(small["value"], 2), "unit": small["unit"], "sample": small["sample"]}
        # the measured projection of the timed workload against the FP64 oracle
        log("bench: FP64 oracle of the 1024^2 x 1000 workload (parity of the measured projection)")
        want = asm.reconst_xyprojection_running(asm.gabor_wavefront(holo_np), SLICES, z_front=-Z0, z_dz=-DZ,
                                                wavlen=LAM, dx=DX)
        got = proj.cpu().numpy()
        parity = {"projection_rel_l2_vs_fp64_oracle": float(f"{asm.rel_l2(got, want):.3e}"),
                  "max_abs_vs_fp64_oracle": float(f"{float(np.abs(got - want).max()):.3e}"),
                  "bar": 1e-4, "workload": "the timed 1024^2 x 1000-slice projection"}

    # per-hologram latency of the headline workload: submission to completion of one 1000-slice
    # reconstruction, 20 calls each synchronised (the stream is idle at every submission)
    lat = []
    for _ in range(20):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
    latency = {"hologram_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                               "p99": round(float(np.percentile(lat, 99)), 3), "n": len(lat),
                               "workload": "one 1024^2 hologram, 1000 slices, idle stream"}}
    if "serving" in extras:
        srv = extras["serving"]["poisson_at_75pct"]
        latency["served_ms"] = {"p50": srv["p50_ms"], "p99": srv["p99_ms"], "n": srv["requests"],
                                "workload": "Poisson stream at 75% of capacity, mixed 512^2-2048^2, 20% GS"}

    out = {
        "metric": METRIC, "value": round(slices_per_s, 1), "unit": "z-slices/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "cf32",
        "data": "synthetic",
        "config": {"workload": "1024x1024 Float32 hologram, 1000-disk seeded scene, 1000 slices "
                               "at dz=200 um from z0=1e4 um, xy min-projection", "size": [N, N],
                   "slices": SLICES, "z0_um": Z0, "dz_um": DZ, "wavelength_um": LAM, "pitch_um": DX,
                   "transfer": "generated FP64-wrapped phase", "parallelism": "single GPU"},
        "roofline": roofline, "latency": latency, "cpu_baseline": cpu, "parity": parity, "extras": extras,
    }
    print(json.dumps(out), flush=True)


def bench_gs(ph, synth, torch, dev, d_sqr):
    """Gerchberg-Saxton at 1024^2: a seeded 500-disk scene recorded at the sensor and 2000 um
    further, 100 iterations (hipGraph-captured iteration), then a 512-slice reconstruction."""
    i1_np, i2_np = synth.gs_pair(103, 500, N, 2000.0)
    i1, i2 = torch.from_numpy(i1_np).to(dev), torch.from_numpy(i2_np).to(dev)
    hf, hb = ph.cu_transfer(2000.0, N, LAM, d_sqr), ph.cu_transfer(-2000.0, N, LAM, d_sqr)
    h0, hdz = ph.cu_transfer(-Z0, N, LAM, d_sqr), ph.cu_transfer(-DZ, N, LAM, d_sqr)

    def run():
        wf = ph.cu_phase_retrieval_holo(i1, i2, hf, hb, 100, N)
        return ph.cu_get_reconst_xyprojection(wf, h0, hdz, 512)

    gw, _ = timed_steps(lambda: ph.cu_phase_retrieval_holo(i1, i2, hf, hb, 100, N), 3, 1, torch)
    tw, _ = timed_steps(run, 3, 1, torch)
    return {"iterations_per_s": round(300 / gw, 1), "gs_algorithmic_GBs": round(72.0 * N * N * 300 / gw / 1e9, 1),
            "pipeline_ms": round(tw / 3 * 1e3, 3), "config": "1024^2, 100 iterations at 2000 um, then 512 slices"}


def bench_volumes(ph, torch, holo, reps=3):
    """The volume variants on the headline hologram (1024^2, dz = 200 um from z0 = 1e4 um):
    cu_get_reconst_vol N0f8 and Float32 and cu_get_reconst_vol_and_xyprojection N0f8 over 1000
    slices, cu_get_reconst_complex_vol over 250 slices (2 GiB). Algorithmic bytes per pixel and
    slice (fused model): pass A 8 + 8/spg_gen, the column-transformed slice read 8 and the
    stored value (1 / 4 / 8), plus 8/spg_out for the projection's read-modify-write."""
    from particleholography import batch as _batch

    t = _batch.get_tuning_for_size(N, N, SLICES)
    a = pass_a_bytes(t["spg_gen"]) + 8.0
    d_sqr = ph.cu_transfer_sqrt_arr(N, LAM, DX)
    h0, hdz = ph.cu_transfer(-Z0, N, LAM, d_sqr), ph.cu_transfer(-DZ, N, LAM, d_sqr)
    wf = ph.cu_gabor_wavefront(holo)
    cases = {"n0f8": (SLICES, a + 1, lambda s: ph.cu_get_reconst_vol(wf, h0, hdz, s)),
             "f32": (SLICES, a + 4, lambda s: ph.cu_get_reconst_vol(wf, h0, hdz, s, return_type=ph.Float32)),
             "n0f8_and_projection": (SLICES, a + 1 + 8.0 / t["spg_out"],
                                     lambda s: ph.cu_get_reconst_vol_and_xyprojection(wf, h0, hdz, s)),
             "complex": (250, a + 8, lambda s: ph.cu_get_reconst_complex_vol(wf, h0, hdz, s))}
    out = {}
    for name, (s, bpx, fn) in cases.items():
        _, sec = timed_steps(lambda: fn(s), reps, 1, torch)
        rate = s * reps / sec
        out[name] = {"slices": s, "slices_per_s": round(rate, 1),
                     "algorithmic_GBs": round(rate * N * N * bpx / 1e9, 1), "bytes_per_px_slice": bpx}
        torch.cuda.empty_cache()
    return out


def bench_lpf(ph, torch, holo):
    """Super-Gaussian low-pass filter on the 1024^2 wavefront, applied in place (the tutorial's
    step before reconstruction). Algorithmic bytes per application: 52 per pixel (row pass
    8 in + 8 out, column pass 8 + 4 filter in + 8 out, inverse row pass 8 in + 8 out)."""
    wf = ph.cu_gabor_wavefront(holo)
    lpf = ph.cu_super_gaussian_filter(Z0 + DZ * SLICES, LAM, N, DX)
    reps = 200
    _, sec = timed_steps(lambda: ph.cu_apply_low_pass_filter_(wf, lpf), reps, 10, torch)
    return {"apply_us": round(sec / reps * 1e6, 2), "algorithmic_GBs": round(52.0 * N * N * reps / sec / 1e9, 1),
            "config": "1024^2 ComplexF32 wavefront, order-3 super-Gaussian cutoff for 2.1e5 um"}


def bench_ingest(ph, torch, dev, reps=20):
    """8-bit ingest at 1024^2: 64 device-resident uint8 holograms converted to Float32 with the
    background removal fused (1 B in + 8 B background + 4 B out per pixel), and the per-pixel
    background mode / mean of a 256-image stack (1 B per pixel per image, read once)."""
    g = torch.Generator(device="cpu").manual_seed(17)
    stack = torch.randint(0, 256, (256, N, N), dtype=torch.uint8, generator=g).to(dev)
    bkg = torch.rand((N, N), dtype=torch.float64, generator=g).to(dev)
    batch = stack[:64]
    _, t_ing = timed_steps(lambda: ph.cu_ingest_u8(batch, bkg=bkg), reps, 2, torch)
    _, t_mode = timed_steps(lambda: ph.cu_make_background(stack, "mode"), reps, 2, torch)
    _, t_mean = timed_steps(lambda: ph.cu_make_background(stack, "mean"), reps, 2, torch)
    px = float(N * N)
    return {"ingest_holograms_per_s": round(64 * reps / t_ing, 1),
            "ingest_GBs": round(13.0 * px * 64 * reps / t_ing / 1e9, 1),
            "background_mode_ms": round(t_mode / reps * 1e3, 3),
            "background_mode_GBs": round(px * 256 * reps / t_mode / 1e9, 1),
            "background_mean_ms": round(t_mean / reps * 1e3, 3),
            "background_mean_GBs": round(px * 256 * reps / t_mean / 1e9, 1),
            "config": "1024^2 uint8 holograms: 64 ingested with a Float64 background, 256 reduced"}


def bench_detect(ph, torch, dev, reps=5):
    """Particle detection on the reference's own test holograms (tests/golden/holo1.png and
    holo2.png, runtests.jl:153-193): GS 9 iterations at +-8e4 um, 1000 N0f8 slices of 1024^2
    from 8e4 um at 100 um, threshold 10/255 fused into the labelling. Timed: components and
    rectangles of every slice (one libholo call, 1 B read per voxel plus the union-find state),
    particle_bounding_boxes (that call, the copy of the used rectangles and the native z-linking),
    particle_coordinates, and the whole chain from the two 8-bit holograms."""
    from particleholography import _lib

    golden = ROOT / "tests" / "golden"
    lam, dx, z0, dz, n, slices, pr_dist, pr_iter, thr = 0.6328, 10.0, 8e4, 100.0, 1024, 1000, 8e4, 9, 10 / 255
    i1 = torch.from_numpy(ph.load_gray2float(str(golden / "holo1.png"))).to(dev)
    i2 = torch.from_numpy(ph.load_gray2float(str(golden / "holo2.png"))).to(dev)
    d_sqr = ph.cu_transfer_sqrt_arr(n, lam, dx)
    d_tf, d_slice = ph.cu_transfer(-z0, n, lam, d_sqr), ph.cu_transfer(-dz, n, lam, d_sqr)
    d_pr, d_pr_inv = ph.cu_transfer(pr_dist, n, lam, d_sqr), ph.cu_transfer(-pr_dist, n, lam, d_sqr)

    def chain():
        d_holo = ph.cu_phase_retrieval_holo(i1, i2, d_pr, d_pr_inv, pr_iter, n)
        vol = ph.cu_get_reconst_vol(d_holo, d_tf, d_slice, slices)
        pbbs = ph.particle_bounding_boxes(vol, threshold=thr)
        return vol, pbbs, ph.particle_coordinates(pbbs, vol)

    chain()  # warm-up (plans, graphs, workspace)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vol, pbbs, coords = chain()
    chain_s = time.perf_counter() - t0
    max_boxes = 4096
    boxes = torch.empty((slices, max_boxes, 5), dtype=torch.int32, device=vol.device)
    nbox = torch.empty(slices, dtype=torch.int32, device=vol.device)
    lib, h = _lib.load(), _lib.handle(0)

    def run():
        st = torch.cuda.current_stream().cuda_stream
        _lib.check(lib.holo_ccl_boxes(h.ptr, st, n, n, slices, vol.data_ptr(), _lib.HOLO_CCL_N0F8_LE,
                                      thr, max_boxes, boxes.data_ptr(), nbox.data_ptr()))

    _, sec = timed_steps(run, reps, 1, torch)
    t0 = time.perf_counter()
    pbbs = ph.particle_bounding_boxes(vol, threshold=thr)
    link_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    ph.particle_coordinates(pbbs, vol)
    coord_s = time.perf_counter() - t0
    return {"label_boxes_ms": round(sec / reps * 1e3, 3),
            "slices_per_s": round(slices * reps / sec, 1),
            "voxel_GBs": round(float(n * n) * slices * reps / sec / 1e9, 1),
            "components": int(nbox.sum().item()), "neighbourhoods": len(pbbs), "particles": len(coords),
            "particle_bounding_boxes_ms": round(link_s * 1e3, 2),
            "particle_coordinates_ms": round(coord_s * 1e3, 2),
            "holograms_to_particles_ms": round(chain_s * 1e3, 2),
            "config": "reference test holograms holo1/holo2 (1024^2): GS 9 iterations at +-8e4 um, 1000 N0f8 "
                      "slices from 8e4 um at 100 um, threshold 10/255, 8-connected, max 4096 boxes per slice"}


def bench_batched(torch, dev, nholo, autotune=True, reps=1):
    """64 seeded holograms, sides from {512, 768, 1024, 1280x1024, 2048}, 100-2000 disks,
    128-1024 slices each over z 1e4-2e5 um; one batched call (tiles autotuned per size unless
    ``autotune`` is False: then the shipped tiles), timed over ``reps`` calls."""
    import numpy as np

    from particleholography import batch, synth
    from particleholography.serve import SIZES

    rng = np.random.Generator(np.random.PCG64(104))
    jobs = []
    for i in range(nholo):
        nx, ny = SIZES[int(rng.integers(len(SIZES)))]
        nd = int(rng.integers(100, 2001))
        s = int(rng.integers(128, 1025))
        h = synth.hologram_device(104 * 1000 + i, nd, nx, ny, zmin=1e4, zmax=2e5)
        jobs.append(batch.ReconJob(h, -1e4, -(2e5 - 1e4) / s, s, LAM, DX))
    before = {(nx, ny): batch.get_tuning_for_size(nx, ny) for nx, ny in SIZES}
    tuned = {f"{nx}x{ny}": (batch.autotune(nx, ny, LAM, DX, save=False) if autotune else before[(nx, ny)])
             for nx, ny in SIZES}
    batch.reconst_batch(jobs)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.reconst_batch(jobs)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    slices = sum(j.slices for j in jobs)
    px_slices = sum(j.slices * j.holo.numel() for j in jobs)
    bpx = {}
    for nx, ny in SIZES:
        t = batch.get_tuning_for_size(nx, ny, 1024)
        bpx[(nx, ny)] = fused_bytes(t["spg_gen"], t["spg_out"])
    alg = sum(j.slices * j.holo.numel() * bpx[(j.holo.shape[1], j.holo.shape[0])] for j in jobs)
    out = {"holograms": nholo, "slices": slices, "slices_per_s": round(slices / dt, 1),
           "slices_1024eq_per_s": round(px_slices / (N * N) / dt, 1),
           "algorithmic_GBs": round(alg / dt / 1e9, 1), "frac_of_8TBs": round(alg / dt / 1e9 / HBM_PEAK_GBS, 3),
           "reference_traffic_GBs": round(REFERENCE_BYTES_PER_PX * px_slices / dt / 1e9, 1),
           "wall_ms": round(dt * 1e3, 2), "tuning": tuned, "per_size": {}}
    for nx, ny in (SIZES if reps == 1 else []):
        sub = [j for j in jobs if tuple(j.holo.shape) == (ny, nx)]
        if not sub:
            continue
        batch.reconst_batch(sub)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.reconst_batch(sub)
        torch.cuda.synchronize()
        d = time.perf_counter() - t0
        ss = sum(j.slices for j in sub)
        gbs = bpx[(nx, ny)] * nx * ny * ss / d / 1e9
        out["per_size"][f"{nx}x{ny}"] = {"holograms": len(sub), "slices": ss, "slices_per_s": round(ss / d, 1),
                                         "bytes_per_px_slice": round(bpx[(nx, ny)], 3),
                                         "algorithmic_GBs": round(gbs, 1), "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 3)}
    for nx, ny in SIZES:
        batch.set_tuning_for_size(nx, ny, before[(nx, ny)])
    return out


def bench_serving(torch, nreq, sizes=None):
    """Seeded Poisson stream of mixed 512^2-2048^2 requests (or of the given sizes only;
    128-1024 slices, 20% two-plane GS with 50 iterations on one scene recorded 2000 um apart):
    capacity with every request queued at once, then the stream at 75% of that capacity
    (p50/p99 latency, overall and per size). The comparator is the plain path, as for the
    headline: every request alone, in arrival order, its reconstruction with one pass-A + pass-B
    launch pair per slice (the host loop of holofunc.jl:259-263) and its GS with the single-pair
    call; measured on the stream's first 150 requests (the same mix), in 1024^2-equivalent
    slices/s."""
    from particleholography import _lib, serve

    kw = {} if sizes is None else {"sizes": sizes}
    reqs = serve.make_stream(nreq, rate=0.0, seed=105, **kw)
    srv = serve.Server(n_slots=2, max_batch=16, hbm_budget=64 << 30, workspace_per_slot=4 << 30,
                       keep_results=False)
    srv.run(reqs[: min(32, nreq)])  # warm-up: plans, graphs
    ws0 = srv.workspace_sizes()
    cap = srv.run(reqs)
    rate = 0.75 * cap["requests"] / cap["makespan_s"]
    reqs2 = serve.make_stream(nreq, rate=rate, seed=105, **kw)
    live = srv.run(reqs2)
    grew = srv.workspace_sizes() != ws0
    sample = serve.make_stream(min(150, nreq), rate=0.0, seed=105, **kw)
    hd = _lib.handle(0)
    hd.set_tuning(1, 1, 1, 0, 0)
    try:
        plain = serve.plain_one_at_a_time(sample)
    finally:
        hd.set_tuning()
    eq = sum(r.slices * r.holo.shape[0] * r.holo.shape[1] for r in sample) / (N * N)
    plain_eq = eq / plain["makespan_s"]
    r = lambda d: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items()}  # noqa: E731
    return {"capacity": r(cap), "poisson_at_75pct": r(live), "rate_req_per_s": round(rate, 2),
            "plain_per_slice": {"requests": plain["requests"], "slices_1024eq_per_s": round(plain_eq, 1),
                                "workload": "the stream's first 150 requests, each alone, one launch pair per slice"},
            "capacity_speedup_vs_plain": round(cap["slices_1024eq_per_s"] / plain_eq, 2),
            "load75_speedup_vs_plain": round(live["slices_1024eq_per_s"] / plain_eq, 2),
            "workspace_grew_after_warmup": grew}


def bench_hbm_practical(torch, dev, nbytes=1 << 30, reps=10):
    """What plain streaming kernels reach on this device (torch's fill and device copy over 1 GiB
    of float32): the practical write and read+write ceilings next to the 8 TB/s published peak
    that every frac above is quoted against."""
    n = nbytes // 4
    a = torch.empty(n, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    a.fill_(1.0)
    out = {}
    for name, fn, moved in (("fill", lambda: a.fill_(2.0), nbytes), ("copy", lambda: b.copy_(a), 2 * nbytes)):
        _, sec = timed_steps(fn, reps, 2, torch)
        out[f"{name}_GBs"] = round(moved * reps / sec / 1e9, 1)
    out["workload"] = "1 GiB float32: fill (write), device copy (read + write)"
    del a, b
    torch.cuda.empty_cache()
    return out


def bench_gs_batched(torch, dev, iters=100):
    """Batched Gerchberg-Saxton at 1024^2: 16 and 32 pairs (four seeded two-plane scenes, 2000 um
    apart, repeated) x 100 iterations in one call, the chunk-uniform iterations in hipGraphs.
    Algorithmic bytes per pixel and iteration: 72 (the fused two-pass propagation with generated
    transfers) and 76 as built (each column pass also reads the 2 B/px quadrant of its
    natural-order transfer table)."""
    from particleholography import synth
    from particleholography.batch import GSJob, phase_retrieval_batch

    pairs = [synth.gs_pair_device(103000 + k, 500, N, N, separation=2000.0) for k in range(4)]
    out = {}
    for npairs in (16, 32):
        jobs = [GSJob(pairs[k % 4][0], pairs[k % 4][1], 2000.0, iters, LAM, DX) for k in range(npairs)]
        sec_total, _ = timed_steps(lambda: phase_retrieval_batch(jobs), 3, 1, torch)
        it_s = npairs * iters * 3 / sec_total
        out[f"{npairs}_pairs"] = {"pair_iterations_per_s": round(it_s, 1),
                                  "algorithmic_GBs_72": round(72.0 * N * N * it_s / 1e9, 1),
                                  "frac_of_8TBs_72": round(72.0 * N * N * it_s / 1e9 / HBM_PEAK_GBS, 3),
                                  "as_built_GBs_76": round(76.0 * N * N * it_s / 1e9, 1),
                                  "call_ms": round(sec_total / 3 * 1e3, 3)}
    out["config"] = f"1024^2, {iters} iterations at 2000 um per pair, one batched call"
    return out


if __name__ == "__main__":
    main()

"""The compile-time lane-swap exchanges of the FFT engine (fft_core.hpp, make_lane_swap), checked
on the CPU: a host-only probe prints the engine's own plans (radices, spans, and for each
exchange whether it has the swap form, with its register bits and renaming). A numpy model of
one wavefront then runs the whole line FFT with that plan: Stockham stages on the per-thread
registers (element t + T*e in register e of thread t), every exchange either through an LDS
array (write at the autosort position, read back in the t + T*e pattern) or, where the plan has
the swap form, through the documented lane semantics of v_permlane32_swap_b32 (lanes 32-63 of
the first register with lanes 0-31 of the second) and v_permlane16_swap_b32 (odd 16-lane rows of
the first with even rows of the second) followed by the register renaming. The result must be
numpy's FFT of the line: a wrong register bit, lane bit or renaming scrambles it.
"""
import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "particleholography.jl_amd" / "csrc"
HIPCC = Path("/opt/rocm/bin/hipcc")

CASES = [(1024, 16), (512, 16), (2048, 16), (768, 12), (1280, 20), (1024, 8), (256, 16)]

PROBE = r"""
#include "fft_core.hpp"
#include <cstdio>
template <int L, int E, int S>
void swap_of() {
  constexpr holo::LaneSwap s = holo::LaneSwapPlan<L, E, S>::s;
  std::printf("%s{\"ok\": %d, \"rb5\": %d, \"rb4\": %d, \"sigma\": [", S > 1 ? ", " : "", s.ok ? 1 : 0, s.rb5, s.rb4);
  for (int e = 0; e < E; ++e) std::printf("%s%d", e ? ", " : "", s.ok ? s.sigma[e] : -1);
  std::printf("]}");
}
template <int L, int E, int... S>
void show(std::integer_sequence<int, S...>) {
  constexpr holo::CPlanData d = holo::CPlan<L, E>::d;
  std::printf("{\"L\": %d, \"E\": %d, \"R\": [", L, E);
  for (int s = 0; s < d.ns; ++s) std::printf("%s%d", s ? ", " : "", d.R[s]);
  std::printf("], \"P\": [");
  for (int s = 0; s < d.ns; ++s) std::printf("%s%d", s ? ", " : "", d.P[s]);
  std::printf("], \"swap\": [");
  (swap_of<L, E, S + 1>(), ...);
  std::printf("]}\n");
}
int main() {
@CALLS@
  return 0;
}
"""


@pytest.fixture(scope="module")
def plans(tmp_path_factory):
    if not HIPCC.exists() or shutil.which("g++") is None:
        pytest.skip("hipcc is not available")
    calls = []
    for L, E in CASES:
        ns = _stages(L, E)
        calls.append(f"  show<{L}, {E}>(std::make_integer_sequence<int, {ns - 1}>{{}});")
    d = tmp_path_factory.mktemp("lane_swap")
    src, exe = d / "probe.hip", d / "probe"
    src.write_text(PROBE.replace("@CALLS@", "\n".join(calls)))
    subprocess.run([str(HIPCC), "-std=c++17", "--cuda-host-only", f"-I{CSRC}", str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {(p["L"], p["E"]): p for p in map(json.loads, out.splitlines())}


def _stages(L, E):
    # stage count only, to size the probe's integer sequence (the plan itself comes from C++)
    n, rem = 0, L
    while rem > 1:
        r = next(x for x in (20, 16, 12, 10, 8, 6, 5, 4, 3, 2) if E % x == 0 and rem % x == 0)
        rem //= r
        n += 1
    return n


def _line_fft(x, plan, use_swaps, direction=-1):
    """One wavefront of 64 lanes holding 64 // T lines; returns the lines after the plan."""
    L, E = plan["L"], plan["E"]
    T = L // E
    lines = 64 // T
    # v[lane, e]: element t + T*e of the lane's line
    v = np.zeros((64, E), np.complex128)
    for lane in range(64):
        line, t = divmod(lane, T)
        v[lane] = x[line][t + T * np.arange(E)]
    R, P = plan["R"], plan["P"]
    for s, (r, p) in enumerate(zip(R, P)):
        if s > 0:
            sw = plan["swap"][s - 1]
            if use_swaps and sw["ok"]:
                v = _swap_exchange(v, sw)
            else:
                v = _lds_exchange(v, y, T, E)
        q_count = E // r
        y = [np.zeros(L, np.complex128) for _ in range(lines)]
        out = v.copy()
        for lane in range(64):
            line, t = divmod(lane, T)
            for q in range(q_count):
                b = t + T * q
                k = b % p
                j = b // p
                xs = np.array([v[lane, q + q_count * m] for m in range(r)])
                xs = xs * np.exp(direction * 2j * np.pi * np.arange(r) * k / (p * r))
                X = np.array([np.sum(xs * np.exp(direction * 2j * np.pi * np.arange(r) * mm / r)) for mm in range(r)])
                for m in range(r):
                    out[lane, q + q_count * m] = X[m]
                    y[line][j * p * r + k + m * p] = X[m]
        v = out
    res = [np.zeros(L, np.complex128) for _ in range(lines)]
    for lane in range(64):
        line, t = divmod(lane, T)
        res[line][t + T * np.arange(E)] = v[lane]
    return res


def _lds_exchange(v, y, T, E):
    out = np.empty_like(v)
    for lane in range(64):
        line, t = divmod(lane, T)
        out[lane] = y[line][t + T * np.arange(E)]
    return out


def _swap_exchange(v, sw):
    v = v.copy()
    E = v.shape[1]
    a, c = sw["rb5"], sw["rb4"]
    if a >= 0:  # v_permlane32_swap(A = v[r], B = v[r | 2^a])
        for r in range(E):
            if (r >> a) & 1:
                continue
            A, B = r, r | (1 << a)
            hiA, loB = v[32:, A].copy(), v[:32, B].copy()
            v[32:, A], v[:32, B] = loB, hiA
    if c >= 0:  # v_permlane16_swap: odd rows of A with even rows of B
        for r in range(E):
            if (r >> c) & 1:
                continue
            A, B = r, r | (1 << c)
            for row in (0, 2):
                odd = slice((row + 1) * 16, (row + 2) * 16)
                even = slice(row * 16, (row + 1) * 16)
                oA, eB = v[odd, A].copy(), v[even, B].copy()
                v[odd, A], v[even, B] = eB, oA
    return v[:, sw["sigma"]]


def test_swap_form_exactly_where_expected(plans):
    ok = {k: [s["ok"] for s in p["swap"]] for k, p in plans.items()}
    # the exchange in front of the last stage of 1024 = 16*16*4 and 512 = 16*16*2 moves only lane
    # bits 4-5; the first exchange (lane bits 0-3), 2048-point lines (128 threads), the mixed
    # 768 / 1280 plans and 8-element 1024 lines have no swap form
    assert ok[(1024, 16)] == [0, 1] and ok[(512, 16)] == [0, 1]
    for k in ((2048, 16), (768, 12), (1280, 20), (1024, 8), (256, 16)):
        assert not any(ok[k]), k
    assert plans[(512, 16)]["swap"][1]["rb5"] == -1  # two lines per wave: lane bit 5 is the line


@pytest.mark.parametrize("L,E", [(1024, 16), (512, 16)])
@pytest.mark.parametrize("direction", [-1, 1])
def test_line_fft_with_lane_swaps_is_the_fft(plans, L, E, direction):
    plan = plans[(L, E)]
    rng = np.random.default_rng(L + E + direction)
    lines = 64 // (L // E)
    x = [rng.standard_normal(L) + 1j * rng.standard_normal(L) for _ in range(lines)]
    want = [np.fft.fft(xi) if direction < 0 else np.fft.ifft(xi) * L for xi in x]
    for use_swaps in (False, True):
        got = _line_fft(x, plan, use_swaps, direction)
        for g, w in zip(got, want):
            assert np.max(np.abs(g - w)) < 1e-9 * L, (use_swaps, np.max(np.abs(g - w)))


def test_a_wrong_renaming_or_register_bit_is_caught(plans):
    """The model is sensitive to the plan's details: a swapped renaming or the other register
    bit no longer gives the FFT."""
    plan = plans[(1024, 16)]
    rng = np.random.default_rng(3)
    x = [rng.standard_normal(1024) + 1j * rng.standard_normal(1024)]
    want = np.fft.fft(x[0])
    bad = json.loads(json.dumps(plan))
    sg = bad["swap"][1]["sigma"]
    sg[1], sg[2] = sg[2], sg[1]
    assert np.max(np.abs(_line_fft(x, bad, True)[0] - want)) > 1e-3
    bad = json.loads(json.dumps(plan))
    bad["swap"][1]["rb5"], bad["swap"][1]["rb4"] = bad["swap"][1]["rb4"], bad["swap"][1]["rb5"]
    assert np.max(np.abs(_line_fft(x, bad, True)[0] - want)) > 1e-3

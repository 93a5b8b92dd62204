"""Mixed-length batches (plfft_mixed_plan_* of
include/plfft_gpu.h) against the FP64 oracle, member by member.

The reference has no mixed-length object; its caller runs one plan per
length, so the oracle for member i is the single-length transform of
oracle/oracle.c (orc_fft, which restates the reference's Cooley-Tukey,
Rader and Bluestein choices, src/planner/planner.cpp:38-47).
"""
import numpy as np
import pytest

import plfft_amd as P
from plfft_amd import rng
from gpu_util import bf16_bits, bf16_to_f32, from_dev, relerr, to_dev

pytestmark = pytest.mark.gpu

# one of each schedule: direct (<= 4096, smooth), four-step (smooth > 4096),
# Bluestein in one block (other n <= 2049), four-step with Bluestein rows
# (n = n1 * n2, n1 smooth, n2 <= 2049: 6054 = 3 * 2018, 16398 = 9 * 1822,
# 65344 = 32 * 2042), chirp-z (the rest)
CURATED = [1, 2, 3, 5, 7, 8, 16, 17, 23, 29, 31, 58, 64, 97, 100, 127, 243, 256, 1000,
           1009, 2047, 2048, 2053, 4096, 4097, 4099, 5000, 6054, 8192, 15015, 16384, 16398,
           40009, 65344, 65521, 65536, 65537, 131072]
# the single-plan bounds: f16 1e-3 (2^-11 output rounding), bf16 8e-3 (2^-8)
TOL_MIX = {"f32": 2e-6, "f64": 1e-12, "f16": 1e-3, "bf16": 8e-3}


def _inputs(lengths, seed, dtype=np.complex128):
    total = sum(lengths)
    return rng.complex_uniform(seed, total, np.complex128).astype(dtype)


def _oracle_members(orc, x, lengths, in_offs, direction):
    out = []
    for n, o in zip(lengths, in_offs):
        out.append(orc.fft(np.asarray(x[o:o + n], np.complex128), direction))
    return out


def _offsets(lengths):
    return list(np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(int))


def _check(y, ref, lengths, out_offs, tol, what=""):
    worst = 0.0
    for n, o, r in zip(lengths, out_offs, ref):
        e = relerr(y[o:o + n], r)
        worst = max(worst, e)
        assert e < tol, f"{what} n={n}: rel err {e:.3g} >= {tol}"
    return worst


def _run(torch, lengths, direction=P.PLFFT_FORWARD, dtype="f32", seed=1):
    plan = P.MixedPlan(lengths, direction, dtype)
    cdt = np.complex64 if dtype == "f32" else np.complex128
    x = _inputs(lengths, seed, cdt)
    dx = to_dev(torch, x)
    dy = torch.zeros_like(dx)
    plan.execute(dx, dy)
    return plan, x, from_dev(torch, dy)


@pytest.mark.parametrize("direction", [P.PLFFT_FORWARD, P.PLFFT_BACKWARD])
def test_mixed_curated_f32(orc, torch_cuda, direction):
    plan, x, y = _run(torch_cuda, CURATED, direction, "f32", seed=11)
    d = plan.describe()
    for tag in (" direct=", " fourstep=", " bluestein=", " ct_bluestein=", " chirpz="):
        assert tag in d and f"{tag}0 " not in d, d
    ref = _oracle_members(orc, x, CURATED, _offsets(CURATED), direction)
    _check(y, ref, CURATED, _offsets(CURATED), TOL_MIX["f32"])


def test_mixed_curated_f64(orc, torch_cuda):
    plan, x, y = _run(torch_cuda, CURATED, P.PLFFT_FORWARD, "f64", seed=12)
    ref = _oracle_members(orc, x, CURATED, _offsets(CURATED), P.PLFFT_FORWARD)
    _check(y, ref, CURATED, _offsets(CURATED), TOL_MIX["f64"])


def test_mixed_log_uniform_f32(orc, torch_cuda):
    lengths = rng.mixed_lengths(5, 300)
    assert sum(rng.is_prime(n) for n in lengths) >= 30
    plan, x, y = _run(torch_cuda, lengths, P.PLFFT_FORWARD, "f32", seed=13)
    ref = _oracle_members(orc, x, lengths, _offsets(lengths), P.PLFFT_FORWARD)
    _check(y, ref, lengths, _offsets(lengths), TOL_MIX["f32"])
    # the whole batch as one vector
    assert relerr(y, np.concatenate(ref)) < TOL_MIX["f32"]


def test_mixed_known_pairs(torch_cuda):
    """Delta -> ones and a pure tone -> a single bin, one member per
    schedule."""
    lengths = [16, 97, 4099, 8192, 1009]
    offs = _offsets(lengths)
    x = np.zeros(sum(lengths), np.complex64)
    x[offs[0] + 0] = 1           # delta at 0 -> all ones
    x[offs[1] + 5] = 1           # delta at 5 -> exp(-2 pi i 5 k / 97)
    k0 = 37                      # tone at bin 37 -> n at bin 37
    for j, (n, o) in enumerate(zip(lengths[2:], offs[2:])):
        t = np.arange(n)
        x[o:o + n] = np.exp(2j * np.pi * k0 * t / n)
    plan = P.MixedPlan(lengths, P.PLFFT_FORWARD, "f32")
    dy = torch_cuda.zeros(len(x), dtype=torch_cuda.complex64, device="cuda")
    plan.execute(to_dev(torch_cuda, x), dy)
    y = from_dev(torch_cuda, dy).astype(np.complex128)
    np.testing.assert_allclose(y[:16], np.ones(16), atol=1e-6)
    k = np.arange(97)
    np.testing.assert_allclose(y[offs[1]:offs[1] + 97], np.exp(-2j * np.pi * 5 * k / 97), atol=1e-5)
    for n, o in zip(lengths[2:], offs[2:]):
        want = np.zeros(n, np.complex128)
        want[k0] = n
        assert relerr(y[o:o + n], want) < 2e-6, n


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
def test_mixed_half_storage(orc, torch_cuda, dtype):
    """16-bit storage, FP32 arithmetic: compared with the FP64 transform of
    the rounded input."""
    lengths = rng.mixed_lengths(9, 60) + [1, 29, 4099]
    offs = _offsets(lengths)
    x = _inputs(lengths, 21)
    if dtype == "f16":
        xs = np.stack([x.real, x.imag], -1).astype(np.float16)
        xr = xs.astype(np.float64)
        dx = to_dev(torch_cuda, xs)
    else:
        xb = bf16_bits(np.stack([x.real, x.imag], -1).astype(np.float32))
        xr = bf16_to_f32(xb).astype(np.float64)
        dx = to_dev(torch_cuda, xb.view(np.int16))
    xin = xr[..., 0] + 1j * xr[..., 1]
    plan = P.MixedPlan(lengths, P.PLFFT_FORWARD, dtype)
    dy = torch_cuda.zeros_like(dx)
    plan.execute(dx, dy)
    yh = from_dev(torch_cuda, dy)
    yr = (yh.astype(np.float64) if dtype == "f16"
          else bf16_to_f32(yh.view(np.uint16)).astype(np.float64))
    y = yr[..., 0] + 1j * yr[..., 1]
    ref = _oracle_members(orc, xin, lengths, offs, P.PLFFT_FORWARD)
    _check(y, ref, lengths, offs, TOL_MIX[dtype], dtype)


def test_mixed_offsets_gaps_and_in_place(orc, torch_cuda):
    # every schedule: direct, one-block Bluestein, four-step, chirp-z, and
    # four-step with Bluestein rows (6054 = 3 * 2018, 16398 = 9 * 1822)
    lengths = [300, 17, 4096, 8192, 2053, 64, 6054, 16398]
    # input members reversed with gaps; output members packed differently
    in_offs, pos = [], 0
    for n in lengths:
        pos += 7
        in_offs.append(pos)
        pos += n
    in_total = pos + 3
    out_offs, pos = [], 0
    for n in reversed(lengths):
        pos += 5
        out_offs.append(pos)
        pos += n
    out_offs = out_offs[::-1]
    out_total = pos + 11
    x = rng.complex_uniform(31, in_total, np.complex64)
    plan = P.MixedPlan(lengths, P.PLFFT_FORWARD, "f32", in_offs, out_offs)
    assert plan.in_span == max(o + n for o, n in zip(in_offs, lengths))
    assert plan.out_span == max(o + n for o, n in zip(out_offs, lengths))
    sentinel = np.complex64(123.5 - 7.25j)
    dy = to_dev(torch_cuda, np.full(out_total, sentinel, np.complex64))
    plan.execute(to_dev(torch_cuda, x), dy)
    y = from_dev(torch_cuda, dy)
    ref = _oracle_members(orc, x, lengths, in_offs, P.PLFFT_FORWARD)
    _check(y, ref, lengths, out_offs, TOL_MIX["f32"], "gaps")
    mask = np.ones(out_total, bool)
    for n, o in zip(lengths, out_offs):
        mask[o:o + n] = False
    assert np.all(y[mask] == sentinel)
    # in place: same buffer, same offsets
    plan2 = P.MixedPlan(lengths, P.PLFFT_BACKWARD, "f32", in_offs, in_offs)
    dz = to_dev(torch_cuda, x)
    plan2.execute(dz, dz)
    z = from_dev(torch_cuda, dz)
    ref2 = _oracle_members(orc, x, lengths, in_offs, P.PLFFT_BACKWARD)
    _check(z, ref2, lengths, in_offs, TOL_MIX["f32"], "in place")


def test_mixed_host_buffers(orc, torch_cuda):
    lengths = [64, 97, 5000, 2053]
    x = _inputs(lengths, 41, np.complex64)
    y = np.zeros_like(x)
    plan = P.MixedPlan(lengths, P.PLFFT_FORWARD, "f32")
    plan.execute(x, y)
    ref = _oracle_members(orc, x, lengths, _offsets(lengths), P.PLFFT_FORWARD)
    _check(y, ref, lengths, _offsets(lengths), TOL_MIX["f32"], "host")
    # host in place
    z = x.copy()
    plan.execute(z, z)
    _check(z, ref, lengths, _offsets(lengths), TOL_MIX["f32"], "host in place")


def test_mixed_repeatable_async_workspace(torch_cuda):
    """Two executions give identical bits, with a caller workspace on a side
    stream."""
    lengths = rng.mixed_lengths(17, 120)
    plan = P.MixedPlan(lengths, P.PLFFT_FORWARD, "f32")
    x = to_dev(torch_cuda, _inputs(lengths, 51, np.complex64))
    ws = torch_cuda.empty(max(1, plan.workspace_size), dtype=torch_cuda.uint8, device="cuda")
    s = torch_cuda.cuda.Stream()
    y1 = torch_cuda.zeros_like(x)
    y2 = torch_cuda.zeros_like(x)
    s.wait_stream(torch_cuda.cuda.current_stream())
    with torch_cuda.cuda.stream(s):
        plan.execute_async(x, y1, stream=s, workspace=ws)
        plan.execute_async(x, y2, stream=s, workspace=ws)
    s.synchronize()
    assert torch_cuda.equal(y1, y2)


def test_mixed_graph_capture_replay(orc, torch_cuda):
    """The async execute captured in a CUDA/HIP graph: every replay matches
    the eager result, with eager work and input changes between replays."""
    torch = torch_cuda
    lengths = rng.mixed_lengths(23, 80) + [4099, 1009]
    plan = P.MixedPlan(lengths, P.PLFFT_FORWARD, "f32")
    x = to_dev(torch, _inputs(lengths, 61, np.complex64))
    y = torch.zeros_like(x)
    ws = torch.zeros(max(1, plan.workspace_size), dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # warm-up outside the capture
        plan.execute_async(x, y, stream=s, workspace=ws)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.execute_async(x, y, stream=torch.cuda.current_stream(), workspace=ws)
    torch.cuda.synchronize()
    for rep in range(4):
        x.copy_(to_dev(torch, _inputs(lengths, 70 + rep, np.complex64)))
        y.fill_(0)
        # eager work between replays
        eager = torch.zeros_like(x)
        plan.execute(x, eager)
        _ = (eager.abs().sum() + 1).item()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, eager), f"replay {rep}"
    xr = from_dev(torch, x)
    ref = _oracle_members(orc, xr, lengths, _offsets(lengths), P.PLFFT_FORWARD)
    _check(from_dev(torch, y), ref, lengths, _offsets(lengths), TOL_MIX["f32"], "graph")


def test_mixed_matches_single_length_plans(torch_cuda):
    """The fused launch agrees with one ordinary plan per length."""
    lengths = [128, 97, 5000, 16384, 2053, 4099]
    offs = _offsets(lengths)
    x = to_dev(torch_cuda, _inputs(lengths, 81, np.complex64))
    y = torch_cuda.zeros_like(x)
    P.MixedPlan(lengths, P.PLFFT_FORWARD, "f32").execute(x, y)
    for n, o in zip(lengths, offs):
        ys = torch_cuda.zeros(n, dtype=torch_cuda.complex64, device="cuda")
        P.Plan("c2c", n, P.PLFFT_FORWARD, "f32").execute(x[o:o + n], ys)
        assert relerr(from_dev(torch_cuda, y[o:o + n]), from_dev(torch_cuda, ys)) < 3e-6, n

"""C2C parity on the GPU: the HIP path through the C ABI vs the FP64 oracle.

Cases follow the reference's tests: round trips returning n*x
(test/advanced_test_utils.hpp:50-124), the c2c_f32 layout table
(test/c2c_f32.cpp:10-27), known pairs (test/delta_constant_test.hpp,
test/pure_tone_test.hpp) and MAY_ALIAS overlap (test/plfft_may_alias_test.cpp).
"""
import numpy as np
import pytest

import plfft_amd as P
from plfft_amd import rng
from gpu_util import CPLX_NP, TOL, default_tol, from_dev, relerr, span, to_dev

pytestmark = pytest.mark.gpu

POW2 = [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384]


def _batch_for(n):
    return max(1, min(64, (1 << 18) // n))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", POW2)
@pytest.mark.parametrize("direction", [-1, 1])
def test_c2c_contiguous_vs_oracle(torch_cuda, orc, prec, n, direction):
    torch = torch_cuda
    if prec == "f64" and n > 8192:
        pytest.skip("f64 single-kernel limit; covered by the large-N tests")
    hm = _batch_for(n) + 3
    x = rng.complex_normal(n, n * hm).astype(CPLX_NP[prec])
    xd = to_dev(torch, x)
    yd = torch.empty_like(xd)
    plan = P.Plan("c2c", n, direction, prec, howmany=hm)
    plan.execute(xd, yd)
    y = from_dev(torch, yd)
    ref = orc.c2c_batched(x.astype(np.complex128), n, direction, hm, 1, n, 1, n, n * hm)
    assert relerr(y, ref) < TOL[prec], plan.describe()


C2C_CASES = [
    (2, 1, 1, 2, 1, 2), (8, 1, 1, 8, 1, 8), (16, 4, 1, 16, 1, 16), (8, 3, 2, 20, 2, 20),
    (64, 2, 8, 600, 8, 600), (23, 49, 441, 9, 441, 9), (24, 4, 1, 32, 1, 32),
    (512, 1, 1, 512, 1, 512), (512, 3, 1, 512, 1, 512), (512, 3, 3, 1, 3, 1),
    (1024, 4, 1, 1024, 1, 1024), (1024, 4, 1, 1024, 1024, 1), (1058, 4, 1058, 1, 1058, 1),
]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", C2C_CASES)
def test_reference_layout_table_round_trip(torch_cuda, orc, prec, case):
    """test/c2c_f32.cpp:10-27 replayed: fwd (NO_ALIAS) then bwd (MAY_ALIAS),
    input (i+b+1)/n, result n*x within eps*4*n*log2(n); plus the forward
    output against the FP64 oracle."""
    torch = torch_cuda
    n, hm, is_, id_, os_, od_ = case
    isz, osz = span(n, is_, hm, id_), span(n, os_, hm, od_)
    x = np.zeros(isz, np.complex128)
    for b in range(hm):
        x[b * id_ + is_ * np.arange(n)] = (np.arange(n) + b + 1) / n
    x = x.astype(CPLX_NP[prec])
    fwd = P.Plan("c2c", n, -1, prec, hm, is_, id_, os_, od_)
    bwd = P.Plan("c2c", n, 1, prec, hm, os_, od_, is_, id_, may_alias=True)
    xd = to_dev(torch, x)
    yd = torch.zeros(osz, dtype=xd.dtype, device="cuda")
    zd = torch.zeros(isz, dtype=xd.dtype, device="cuda")
    fwd.execute(xd, yd)
    bwd.execute(yd, zd)
    y, z = from_dev(torch, yd), from_dev(torch, zd)
    ref = orc.c2c_batched(x.astype(np.complex128), n, -1, hm, is_, id_, os_, od_, osz)
    assert relerr(y, ref) < TOL[prec]
    tol = default_tol(prec, n)
    for b in range(hm):
        idx = b * id_ + is_ * np.arange(n)
        assert np.max(np.abs(z[idx].real / n - x[idx].real)) <= tol


@pytest.mark.parametrize("n", [16, 256, 4096])
def test_known_pairs(torch_cuda, n):
    torch = torch_cuda
    plan = P.Plan("c2c", n, -1, "f32", howmany=3)
    x = np.zeros((3, n), np.complex64)
    x[0, 0] = 1                      # delta -> constant
    x[1, :] = 1                      # constant -> n * delta
    k0 = n // 3
    x[2, k0] = 1                     # delta at k0 -> pure tone
    xd = to_dev(torch, x)
    yd = torch.empty_like(xd)
    plan.execute(xd, yd)
    y = from_dev(torch, yd)
    assert np.allclose(y[0], 1, atol=1e-6)
    expect = np.zeros(n)
    expect[0] = n
    assert np.allclose(y[1], expect, atol=1e-6 * n)
    tone = np.exp(-2j * np.pi * np.arange(n) * k0 / n)
    assert np.max(np.abs(y[2] - tone)) < 1e-5


@pytest.mark.parametrize("n", [64, 1024, 4096])
def test_in_place_and_partial_overlap(torch_cuda, orc, n):
    """MAY_ALIAS: exact overlap and +-1 element partial overlap of a
    single member (test/plfft_may_alias_test.cpp:53-79)."""
    torch = torch_cuda
    x = rng.complex_uniform(0, n)
    ref = orc.fft(x.astype(np.complex128))
    plan = P.Plan("c2c", n, -1, "f32", may_alias=True)
    for shift in (0, 1, -1):
        buf = np.zeros(n + 2, np.complex64)
        in_off = 1
        buf[in_off:in_off + n] = x
        d = to_dev(torch, buf)
        base = d.data_ptr()
        plan.execute(base + 8 * in_off, base + 8 * (in_off + shift))
        out = from_dev(torch, d)[in_off + shift:in_off + shift + n]
        assert relerr(out, ref) < TOL["f32"], shift


def test_host_pointers_are_staged(orc):
    """A reference caller passes host memory: the library stages it."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    n, hm = 1024, 5
    x = rng.complex_uniform(4, n * hm)
    y = np.zeros_like(x)
    P.Plan("c2c", n, -1, "f32", howmany=hm).execute(x, y)
    ref = orc.c2c_batched(x.astype(np.complex128), n, -1, hm, 1, n, 1, n, n * hm)
    assert relerr(y, ref) < TOL["f32"]


def test_async_on_torch_stream(torch_cuda, orc):
    torch = torch_cuda
    n, hm = 4096, 40
    x = rng.complex_uniform(9, n * hm)
    plan = P.Plan("c2c", n, -1, "f32", howmany=hm)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        xd = torch.from_numpy(x).to("cuda", non_blocking=False)
        yd = torch.empty_like(xd)
        plan.execute_async(xd, yd, stream=s)
    s.synchronize()
    ref = orc.c2c_batched(x.astype(np.complex128), n, -1, hm, 1, n, 1, n, n * hm)
    assert relerr(yd.cpu().numpy(), ref) < TOL["f32"]


def test_adjoint_c2c(torch_cuda):
    """<F x, y> == <x, F^H y> with F^H the adjoint execute."""
    torch = torch_cuda
    n, hm = 1024, 4
    x = rng.complex_normal(1, n * hm).astype(np.complex128)
    y = rng.complex_normal(2, n * hm).astype(np.complex128)
    plan = P.Plan("c2c", n, -1, "f64", howmany=hm)
    xd, yd = to_dev(torch, x), to_dev(torch, y)
    fx, ay = torch.empty_like(xd), torch.empty_like(yd)
    plan.execute(xd, fx)
    plan.adjoint(yd, ay)
    lhs = np.vdot(from_dev(torch, fx), y)
    rhs = np.vdot(x, from_dev(torch, ay))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


# Mixed-radix lengths served by the runtime-chain engine (prime factors
# <= 23), including 15015 = 3*5*7*11*13 of the prime/mixed-radix config.
MIXED = [3, 5, 6, 7, 9, 10, 12, 15, 17, 19, 23, 24, 25, 30, 36, 49, 60, 96, 100, 120,
         125, 243, 360, 529, 1000, 1001, 1058, 3072, 6000, 10240, 15015]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", MIXED)
def test_c2c_mixed_radix_vs_oracle(torch_cuda, orc, prec, n):
    torch = torch_cuda
    hm = max(1, min(24, (1 << 17) // n)) + 1
    x = rng.complex_normal(n + 1, n * hm).astype(CPLX_NP[prec])
    xd = to_dev(torch, x)
    yd = torch.empty_like(xd)
    for direction in (-1, 1):
        plan = P.Plan("c2c", n, direction, prec, howmany=hm)
        plan.execute(xd, yd)
        y = from_dev(torch, yd)
        ref = orc.c2c_batched(x.astype(np.complex128), n, direction, hm, 1, n, 1, n, n * hm)
        assert relerr(y, ref) < TOL[prec], plan.describe()


# Lengths beyond one LDS tile: the four-step column/row passes (or, for
# 32768-point f32, the single pass with the split LDS exchange).
LARGE = [(32768, "f32"), (65536, "f32"), (131072, "f32"), (1 << 20, "f32"),
         (16384, "f64"), (65536, "f64"), (1 << 20, "f64"), (3 * 65536, "f32"),
         (100000, "f32")]


@pytest.mark.parametrize("n,prec", LARGE)
def test_c2c_four_step_vs_oracle(torch_cuda, orc, n, prec):
    torch = torch_cuda
    hm = 3
    x = rng.complex_normal(n + 7, n * hm).astype(CPLX_NP[prec])
    xd = to_dev(torch, x)
    yd = torch.empty_like(xd)
    plan = P.Plan("c2c", n, -1, prec, howmany=hm)
    # 32768-point f32 members fit one pass with the split LDS exchange
    assert "fourstep" in plan.describe() or "_split" in plan.describe()
    plan.execute(xd, yd)
    y = from_dev(torch, yd)
    ref = orc.c2c_batched(x.astype(np.complex128), n, -1, hm, 1, n, 1, n, n * hm)
    assert relerr(y, ref) < TOL[prec], plan.describe()
    # in place, backward, with the round trip returning n*x
    plan_b = P.Plan("c2c", n, 1, prec, howmany=hm, may_alias=True)
    plan_b.execute(yd, yd)
    z = from_dev(torch, yd)
    assert relerr(z / n, x) < TOL[prec]


def test_four_step_strided_batch(torch_cuda, orc):
    """Interleaved members (istride = howmany) through the four-step path."""
    torch = torch_cuda
    n, hm = 65536, 4
    x = rng.complex_normal(5, n * hm).astype(np.complex64)
    xd = to_dev(torch, x)
    yd = torch.zeros_like(xd)
    plan = P.Plan("c2c", n, -1, "f32", hm, istride=hm, idist=1, ostride=1, odist=n)
    plan.execute(xd, yd)
    ref = orc.c2c_batched(x.astype(np.complex128), n, -1, hm, hm, 1, 1, n, n * hm)
    assert relerr(from_dev(torch, yd), ref) < TOL["f32"]


# Lengths with prime factors above 23: fused Rader (n prime, n-1 smooth),
# fused Bluestein, and chirp-z around the four-step transform; includes the
# prime configuration's 97, 1009 (Rader), 65537 and the reference's
# alias-test primes 509, 2053, 8191 (test/plfft_may_alias_test.cpp:40-52).
PRIMES = [29, 31, 37, 47, 53, 83, 97, 101, 509, 1009, 1399, 2053, 4099, 8191, 2 * 1009, 29 * 31]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", PRIMES)
def test_c2c_prime_lengths(torch_cuda, orc, prec, n):
    torch = torch_cuda
    hm = max(1, min(12, (1 << 16) // n)) + 1
    x = rng.complex_normal(n + 3, n * hm).astype(CPLX_NP[prec])
    xd = to_dev(torch, x)
    yd = torch.empty_like(xd)
    for direction in (-1, 1):
        plan = P.Plan("c2c", n, direction, prec, howmany=hm)
        plan.execute(xd, yd)
        ref = orc.c2c_batched(x.astype(np.complex128), n, direction, hm, 1, n, 1, n, n * hm)
        assert relerr(from_dev(torch, yd), ref) < TOL[prec], plan.describe()


def test_c2c_prime_plan_choice():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert "rader" in P.Plan("c2c", 97, -1, "f32").describe()
    assert "rader" in P.Plan("c2c", 1009, -1, "f32").describe()
    # 47 - 1 = 2*23 has an engine here (Rader); 59 - 1 = 2*29 does not
    assert "rader" in P.Plan("c2c", 47, -1, "f32").describe()
    assert "bluestein" in P.Plan("c2c", 59, -1, "f32").describe()


# primes past one LDS tile whose n - 1 has no compiled four-step split
# (65520 = 2^4 3^2 5 7 13) and a composite with a large prime factor
@pytest.mark.parametrize("n,prec", [(65521, "f32"), (65521, "f64"), (40009, "f32")])
def test_c2c_large_prime_chirp_z(torch_cuda, orc, n, prec):
    torch = torch_cuda
    hm = 3
    x = rng.complex_normal(n, n * hm).astype(CPLX_NP[prec])
    xd = to_dev(torch, x)
    yd = torch.empty_like(xd)
    plan = P.Plan("c2c", n, -1, prec, howmany=hm)
    assert "chirpz" in plan.describe()
    plan.execute(xd, yd)
    ref = orc.c2c_batched(x.astype(np.complex128), n, -1, hm, 1, n, 1, n, n * hm)
    assert relerr(from_dev(torch, yd), ref) < TOL[prec] * 2, plan.describe()


# Rader in three passes: primes whose n - 1 splits into compiled column and
# row sides (65537: 256 x 256; 786433: 1024 x 768 in FP32).
# (f16 / bf16 storage: tests/test_half_gpu.py, n = 65537)
@pytest.mark.parametrize("n,prec", [(65537, "f32"), (65537, "f64"), (786433, "f32")])
def test_c2c_rader_three_pass(torch_cuda, orc, n, prec):
    torch = torch_cuda
    hm = 3 if n < 100000 else 1
    x = rng.complex_normal(n + 7, n * hm).astype(CPLX_NP[prec])
    xd = to_dev(torch, x)
    for direction in (-1, 1):
        plan = P.Plan("c2c", n, direction, prec, howmany=hm)
        print(f"rader3 {prec} n={n} dir={direction}: {plan.describe()}", flush=True)
        assert "rader3" in plan.describe()
        yd = torch.empty_like(xd)
        plan.execute(xd, yd)
        ref = orc.c2c_batched(x.astype(np.complex128), n, direction, hm, 1, n, 1, n, n * hm)
        assert relerr(from_dev(torch, yd), ref) < TOL[prec] * 2, plan.describe()


def test_c2c_rader_three_pass_layouts(torch_cuda, orc):
    """In place, and strided input / output (istride 2, idist 2n + 64), of
    the three-pass Rader plan; plus delta and tone known pairs."""
    torch = torch_cuda
    n, hm = 65537, 2
    x = rng.complex_normal(n + 11, n * hm).astype(np.complex64)
    ref = orc.c2c_batched(x.astype(np.complex128), n, -1, hm, 1, n, 1, n, n * hm)
    xd = to_dev(torch, x)
    plan = P.Plan("c2c", n, -1, "f32", howmany=hm)
    assert "rader3" in plan.describe()
    plan.execute(xd, xd)
    assert relerr(from_dev(torch, xd), ref) < TOL["f32"] * 2
    idist = 2 * n + 64
    xs = np.zeros(idist * hm, np.complex64)
    for b in range(hm):
        xs[b * idist:b * idist + 2 * n:2] = x[b * n:(b + 1) * n]
    xsd = to_dev(torch, xs)
    ysd = torch.zeros_like(xsd)
    plan = P.Plan("c2c", n, -1, "f32", howmany=hm, istride=2, idist=idist, ostride=2,
                  odist=idist)
    plan.execute(xsd, ysd)
    ys = from_dev(torch, ysd)
    got = np.concatenate([ys[b * idist:b * idist + 2 * n:2] for b in range(hm)])
    assert relerr(got, ref) < TOL["f32"] * 2
    k0 = n // 3
    xp = np.zeros((2, n), np.complex64)
    xp[0, 0] = 1
    xp[1, k0] = 1
    xd = to_dev(torch, xp)
    yd = torch.empty_like(xd)
    P.Plan("c2c", n, -1, "f32", howmany=2).execute(xd, yd)
    y = from_dev(torch, yd)
    assert np.max(np.abs(y[0] - 1)) < 1e-5
    assert np.max(np.abs(y[1] - np.exp(-2j * np.pi * np.arange(n) * k0 / n))) < 2e-5


@pytest.mark.parametrize("n", [97, 509, 2053, 8191])
def test_prime_known_pairs(torch_cuda, n):
    """delta -> constant and pure tone at k0 = n//3 (prime configuration)."""
    torch = torch_cuda
    x = np.zeros((2, n), np.complex64)
    x[0, 0] = 1
    k0 = n // 3
    x[1, k0] = 1
    xd = to_dev(torch, x)
    yd = torch.empty_like(xd)
    P.Plan("c2c", n, -1, "f32", howmany=2).execute(xd, yd)
    y = from_dev(torch, yd)
    assert np.max(np.abs(y[0] - 1)) < 1e-5
    assert np.max(np.abs(y[1] - np.exp(-2j * np.pi * np.arange(n) * k0 / n))) < 1e-5


def test_size_limits_fail_cleanly(torch_cuda):
    """Lengths past the four-step range of a precision (2^22 points for f64,
    2^24 for f32) fail plan creation with PLFFT_PLAN_CREATION_FAILED instead of
    nesting chirp-z plans without end; the largest four-step lengths plan."""
    for prec, n_ok, n_bad in (("f64", 1 << 22, 1 << 23), ("f32", 1 << 24, 1 << 25)):
        print(f"size limit {prec}: {n_ok} plans, {n_bad} fails", flush=True)
        assert "fourstep" in P.Plan("c2c", n_ok, -1, prec, rigor=P.PLFFT_ESTIMATE).describe()
        with pytest.raises(P.PlfftError) as ei:
            P.Plan("c2c", n_bad, -1, prec, rigor=P.PLFFT_ESTIMATE)
        assert ei.value.status == P.PLFFT_PLAN_CREATION_FAILED
        assert "chirp-z" in P.last_error()
    # the real transforms of those lengths widen or fold onto the same plans
    with pytest.raises(P.PlfftError) as ei:
        P.Plan("r2c", (1 << 24) + 2, -1, "f64", rigor=P.PLFFT_ESTIMATE)
    assert ei.value.status == P.PLFFT_PLAN_CREATION_FAILED


@pytest.mark.parametrize("n", [65536, 65537, 40009])
def test_one_plan_from_four_host_threads(torch_cuda, orc, n):
    """include/plfft.h allows concurrent execute of one plan from several
    threads; multi-pass plans (four-step 65536, Rader 65537, chirp-z 40009)
    take one workspace per concurrent synchronous call from the plan's pool."""
    import threading
    torch = torch_cuda
    hm = 3
    plan = P.Plan("c2c", n, -1, "f32", howmany=hm)
    print(f"threads: {plan.describe()[:80]}", flush=True)
    assert plan.workspace_size > 0
    xs = [rng.complex_uniform(40 + t, n * hm) for t in range(4)]
    xd = [to_dev(torch, x) for x in xs]
    yd = [torch.empty_like(d) for d in xd]
    torch.cuda.synchronize()
    errors = []

    def work(t):
        try:
            for _ in range(3):
                plan.execute(xd[t], yd[t])  # ctypes releases the GIL
        except Exception as e:  # reported to the main thread
            errors.append(e)

    th = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors
    for t in range(4):
        ref = orc.c2c_batched(xs[t].astype(np.complex128), n, -1, hm, 1, n, 1, n, n * hm)
        assert relerr(from_dev(torch, yd[t]), ref) < TOL["f32"], t


# test/plfft_may_alias_test.cpp:39-42: every size of the reference's
# aliasing test (powers of two, primes for Rader / Bluestein / chirp-z,
# four-step 65536)
ALIAS_SIZES = [2, 3, 4, 5, 7, 8, 11, 16, 21, 32, 47, 64, 97, 128, 256, 509, 1024, 2053, 4096,
               8191, 16384, 65536]


@pytest.mark.parametrize("n", ALIAS_SIZES)
def test_alias_sizes_of_the_reference(torch_cuda, orc, n):
    """MAY_ALIAS with exact and +-1 element partial overlap at the
    reference's sizes (test/plfft_may_alias_test.cpp:44-79)."""
    torch = torch_cuda
    x = rng.complex_uniform(0, n)
    ref = orc.fft(x.astype(np.complex128))
    plan = P.Plan("c2c", n, -1, "f32", may_alias=True)
    for shift in (0, 1, -1):
        buf = np.zeros(n + 2, np.complex64)
        buf[1:1 + n] = x
        d = to_dev(torch, buf)
        base = d.data_ptr()
        plan.execute(base + 8, base + 8 * (1 + shift))
        out = from_dev(torch, d)[1 + shift:1 + shift + n]
        assert relerr(out, ref) < TOL["f32"], (shift, plan.describe())

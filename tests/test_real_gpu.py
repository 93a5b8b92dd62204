"""R2C / C2R parity on the GPU against the FP64 oracle.

Layouts replay the reference's tests: the strided R2C/C2R table of
test/r2c_c2r_1d_batched.cpp:40-116, the in-place batched round trip of
test/plfft_interface_test.cpp:360-412, the n = 2..128 sweep of
test/simple_test.cpp:14-49 (X[0] = sum x, round trip n*x), and the
differentiable-transform layout (n=4096, input stride 2, distance 2n+64).
"""
import numpy as np
import pytest

import plfft_amd as P
from plfft_amd import rng
from gpu_util import TOL, from_dev, relerr, span, to_dev

pytestmark = pytest.mark.gpu

RNP = {"f32": np.float32, "f64": np.float64}
CNP = {"f32": np.complex64, "f64": np.complex128}


def _r2c(torch, orc, prec, n, hm, is_, id_, os_, od_):
    isz, osz = span(n, is_, hm, id_), span(n // 2 + 1, os_, hm, od_)
    x = rng.normal_f64(n * 3 + hm, isz).astype(RNP[prec])
    xd = to_dev(torch, x)
    yd = torch.zeros(osz, dtype=getattr(torch, "complex64" if prec == "f32" else "complex128"),
                     device="cuda")
    plan = P.Plan("r2c", n, -1, prec, hm, is_, id_, os_, od_)
    plan.execute(xd, yd)
    ref = orc.r2c_batched(x.astype(np.float64), n, hm, is_, id_, os_, od_, osz)
    return x, from_dev(torch, yd), ref, plan


def _c2r(torch, orc, prec, n, hm, is_, id_, os_, od_, X):
    osz = span(n, os_, hm, od_)
    Xd = to_dev(torch, X.astype(CNP[prec]))
    yd = torch.zeros(osz, dtype=getattr(torch, "float32" if prec == "f32" else "float64"),
                     device="cuda")
    plan = P.Plan("c2r", n, 1, prec, hm, is_, id_, os_, od_)
    plan.execute(Xd, yd)
    ref = orc.c2r_batched(X.astype(np.complex128), n, hm, is_, id_, os_, od_, osz)
    return from_dev(torch, yd), ref, plan


SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 17, 30, 31, 46, 64, 77, 100, 128, 256, 500,
         1000, 1024, 2048, 4096, 6000, 8192, 16384]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", SIZES)
def test_r2c_c2r_contiguous(torch_cuda, orc, prec, n):
    torch = torch_cuda
    hm = max(1, min(16, (1 << 16) // n)) + 1
    m = n // 2 + 1
    x, y, ref, plan = _r2c(torch, orc, prec, n, hm, 1, n, 1, m)
    assert relerr(y, ref) < TOL[prec], plan.describe()
    # C2R of the spectrum: the round trip returns n*x
    z, zref, plan2 = _c2r(torch, orc, prec, n, hm, 1, m, 1, n, ref)
    assert relerr(z, zref) < TOL[prec], plan2.describe()
    assert relerr(z / n, x.astype(np.float64)) < TOL[prec]


def test_c2r_ignores_imaginary_dc_and_nyquist(torch_cuda, orc):
    torch = torch_cuda
    for n in (8, 9, 4096):
        m = n // 2 + 1
        X = rng.complex_normal(n, m * 2).astype(np.complex128)
        z, zref, _ = _c2r(torch, orc, "f64", n, 2, 1, m, 1, n, X)
        X2 = X.reshape(2, m).copy()
        X2[:, 0] = X2[:, 0].real
        if n % 2 == 0:
            X2[:, -1] = X2[:, -1].real
        z2, _, _ = _c2r(torch, orc, "f64", n, 2, 1, m, 1, n, X2.ravel())
        assert relerr(z, zref) < 1e-13
        assert np.allclose(z, z2, atol=1e-12 * n)


# test/r2c_c2r_1d_batched.cpp:40-60 (n, howmany, istride, idist, ostride, odist)
STRIDED = [(22, 3, 3, 88, 3, 90), (46, 4, 2, 100, 2, 92), (12, 5, 1, 12, 1, 12),
           (128, 6, 5, 650, 2, 256), (77, 7, 4, 308, 3, 231)]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", STRIDED)
def test_r2c_c2r_strided_round_trip(torch_cuda, orc, prec, case):
    torch = torch_cuda
    n, hm, is_, id_, os_, od_ = case
    x, y, ref, plan = _r2c(torch, orc, prec, n, hm, is_, id_, os_, od_)
    assert relerr(y, ref) < TOL[prec], plan.describe()
    # inverse with the layouts swapped, as the reference test does
    z, zref, _ = _c2r(torch, orc, prec, n, hm, os_, od_, is_, id_, ref)
    assert relerr(z, zref) < TOL[prec]
    for b in range(hm):
        idx = b * id_ + is_ * np.arange(n)
        assert relerr(z[idx] / n, x[idx].astype(np.float64)) < TOL[prec]


@pytest.mark.parametrize("n", [4, 5])
def test_in_place_batched_round_trip_host(n):
    """test/plfft_interface_test.cpp:360-412 on host memory (staged)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    hm, rspan, cspan = 2, 6, 3
    data = np.zeros(hm * rspan, np.float32)
    for b in range(hm):
        for e in range(n):
            data[b * rspan + e] = b * 10 + e + 1
    expected = data * n
    r2c = P.Plan("r2c", n, -1, "f32", hm, 1, rspan, 1, cspan, may_alias=True)
    c2r = P.Plan("c2r", n, 1, "f32", hm, 1, cspan, 1, rspan, may_alias=True)
    r2c.execute(data, data)
    c2r.execute(data, data)
    for b in range(hm):
        for e in range(n):
            i = b * rspan + e
            assert abs(data[i] - expected[i]) < 1e-4


@pytest.mark.parametrize("n", [4096, 1000])
def test_in_place_device(torch_cuda, orc, n):
    torch = torch_cuda
    hm = 9
    m = n // 2 + 1
    rdist = 2 * m  # padded real rows, FFTW-style in-place layout
    x = rng.normal_f64(1, hm * rdist).astype(np.float32)
    d = to_dev(torch, x)
    P.Plan("r2c", n, -1, "f32", hm, 1, rdist, 1, m, may_alias=True).execute(d, d)
    Y = from_dev(torch, d).view(np.complex64)
    ref = orc.r2c_batched(x.astype(np.float64), n, hm, 1, rdist, 1, m, hm * m)
    assert relerr(Y, ref) < TOL["f32"]
    P.Plan("c2r", n, 1, "f32", hm, 1, m, 1, rdist, may_alias=True).execute(d, d)
    z = from_dev(torch, d)
    for b in range(hm):
        assert relerr(z[b * rdist:b * rdist + n] / n, x[b * rdist:b * rdist + n]) < TOL["f32"]


def test_simple_sweep_first_output_is_sum(torch_cuda):
    """test/simple_test.hpp:16-95: X[0] = sum x, C2R[0] = Hermitian sum."""
    torch = torch_cuda
    for n in range(2, 129):
        x = (np.arange(n) % 7 - 3).astype(np.float32)
        X = torch.zeros(n // 2 + 1, dtype=torch.complex64, device="cuda")
        P.Plan("r2c", n, -1, "f32").execute(to_dev(torch, x), X)
        Xh = from_dev(torch, X)
        assert abs(Xh[0].real - x.sum()) <= 1e-5 * n and abs(Xh[0].imag) <= 1e-5
        y = torch.zeros(n, dtype=torch.float32, device="cuda")
        P.Plan("c2r", n, 1, "f32").execute(X, y)
        yh = from_dev(torch, y)
        herm = Xh[0].real + 2 * Xh[1:(n + 1) // 2].real.sum() + (Xh[n // 2].real if n % 2 == 0 else 0)
        assert abs(yh[0] - herm) <= 1e-4 * n


def test_differentiable_layout_r2c(torch_cuda, orc):
    """n=4096, input stride 2, distance 2n+64 (reduced batch)."""
    torch = torch_cuda
    n, hm, is_, id_ = 4096, 16, 2, 2 * 4096 + 64
    x, y, ref, plan = _r2c(torch, orc, "f32", n, hm, is_, id_, 1, n // 2 + 1)
    assert relerr(y, ref) < TOL["f32"], plan.describe()


# Long half lengths: f32 n = 65536 runs the one-pass split-exchange kernel
# (the 32768-point half transform held in registers); the others run the
# packed four-step passes (kernels_real4s.hpp): f64 65536 = 2 x 128 x 256,
# 2^20 = 2 x 512 x 1024 and 120000 = 2 x 240 x 250 (runtime radix chains).
# None of them widens to a full-length complex transform.
FOURSTEP = [65536, 1 << 20, 120000]


def _packed(plan):
    d = plan.describe()
    return "via_c2c" not in d and ("fourstep" in d or "_split" in d)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", FOURSTEP)
def test_r2c_c2r_fourstep(torch_cuda, orc, prec, n):
    torch = torch_cuda
    hm = 3 if n < (1 << 20) else 2
    m = n // 2 + 1
    print(f"r2c/c2r fourstep {prec} n={n} hm={hm}", flush=True)
    x, y, ref, plan = _r2c(torch, orc, prec, n, hm, 1, n, 1, m)
    assert _packed(plan), plan.describe()
    assert relerr(y, ref) < TOL[prec], plan.describe()
    z, zref, plan2 = _c2r(torch, orc, prec, n, hm, 1, m, 1, n, ref)
    assert _packed(plan2), plan2.describe()
    assert relerr(z, zref) < TOL[prec], plan2.describe()
    assert relerr(z / n, x.astype(np.float64)) < TOL[prec]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n", [65536, 1 << 20])
def test_r2c_c2r_fourstep_in_place(torch_cuda, orc, prec, n):
    torch = torch_cuda
    hm = 2
    m = n // 2 + 1
    rdist = 2 * m  # padded real rows, FFTW-style in-place layout
    print(f"r2c/c2r fourstep in place {prec} n={n}", flush=True)
    x = rng.normal_f64(5, hm * rdist).astype(RNP[prec])
    d = to_dev(torch, x)
    r2c = P.Plan("r2c", n, -1, prec, hm, 1, rdist, 1, m, may_alias=True)
    assert _packed(r2c), r2c.describe()
    r2c.execute(d, d)
    Y = from_dev(torch, d).view(CNP[prec])
    ref = orc.r2c_batched(x.astype(np.float64), n, hm, 1, rdist, 1, m, hm * m)
    assert relerr(Y, ref) < TOL[prec]
    P.Plan("c2r", n, 1, prec, hm, 1, m, 1, rdist, may_alias=True).execute(d, d)
    z = from_dev(torch, d)
    for b in range(hm):
        seg = slice(b * rdist, b * rdist + n)
        assert relerr(z[seg] / n, x[seg].astype(np.float64)) < TOL[prec]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_r2c_c2r_fourstep_strided(torch_cuda, orc, prec):
    """Unpaired real access: input stride 2 with an odd-padded distance,
    complex side strided too."""
    torch = torch_cuda
    n, hm = 65536, 3
    is_, id_, os_, od_ = 2, 2 * n + 63, 2, 2 * (n // 2 + 1) + 5
    x, y, ref, plan = _r2c(torch, orc, prec, n, hm, is_, id_, os_, od_)
    assert _packed(plan), plan.describe()
    assert relerr(y, ref) < TOL[prec], plan.describe()
    z, zref, _ = _c2r(torch, orc, prec, n, hm, os_, od_, is_, id_, ref)
    assert relerr(z, zref) < TOL[prec]
    for b in range(hm):
        idx = b * id_ + is_ * np.arange(n)
        assert relerr(z[idx] / n, x[idx].astype(np.float64)) < TOL[prec]


def test_r2c_c2r_fourstep_known_pairs(torch_cuda):
    """Delta -> constant and a pure tone -> one bin (test/delta_constant_test.hpp,
    test/pure_tone_test.hpp) at n = 2^20."""
    torch = torch_cuda
    n = 1 << 20
    m = n // 2 + 1
    x = np.zeros(n, np.float64)
    x[0] = 1.0
    X = torch.zeros(m, dtype=torch.complex128, device="cuda")
    P.Plan("r2c", n, -1, "f64").execute(to_dev(torch, x), X)
    Xh = from_dev(torch, X)
    assert np.max(np.abs(Xh - 1.0)) < 1e-12
    k0 = 12345
    # the phase reduced exactly mod n, so the tone itself carries no
    # frequency error (2*pi*k0*j/n in floating point leaks ~1e-6 into
    # neighbouring bins at this n)
    tone = np.cos(2 * np.pi * ((k0 * np.arange(n, dtype=np.int64)) % n) / n)
    P.Plan("r2c", n, -1, "f64").execute(to_dev(torch, tone), X)
    Xh = from_dev(torch, X)
    want = np.zeros(m, np.complex128)
    want[k0] = n / 2
    assert np.max(np.abs(Xh - want)) < 1e-9


@pytest.mark.parametrize("n", [16, 509, 4096, 65536])
def test_r2c_partial_overlap(torch_cuda, orc, n):
    """MAY_ALIAS R2C whose output starts one complex element before, at, or
    after its input (the C2C overlap cases of test/plfft_may_alias_test.cpp
    applied to the real transform)."""
    torch = torch_cuda
    x = rng.normal_f64(7, n).astype(np.float32)
    ref = orc.r2c_batched(x.astype(np.float64), n, 1, 1, n, 1, n // 2 + 1, n // 2 + 1)
    plan = P.Plan("r2c", n, -1, "f32", may_alias=True)
    m = n // 2 + 1
    for shift in (0, 1, -1):  # complex elements
        buf = np.zeros(2 * (m + 2) + 2 * ((n + 1) // 2), np.float32)  # whole complex elements
        buf[2:2 + n] = x  # input starts at float 2 = complex 1
        d = to_dev(torch, buf)
        base = d.data_ptr()
        plan.execute(base + 8, base + 8 * (1 + shift))
        out = from_dev(torch, d).view(np.complex64)[1 + shift:1 + shift + m]
        assert relerr(out, ref) < TOL["f32"], (shift, plan.describe())

"""DCT-II / DST-II / DCT-III / DST-III parity on the GPU against the oracle.

Mirrors test/r2r_test.cpp:11-43 (n = 2..128, first-element identity
DCT-II y0 = 2 sum x, round-trip factor 2n, test/r2r_test.hpp:60-132) and the
differentiable-transform layout (n=4096, input stride 2, distance 2n+64).
"""
import numpy as np
import pytest

import plfft_amd as P
from plfft_amd import rng
from gpu_util import TOL, from_dev, relerr, span, to_dev

pytestmark = pytest.mark.gpu

RNP = {"f32": np.float32, "f64": np.float64}
KINDS = {"dct2": 1, "dct3": 2, "dst2": 5, "dst3": 6}
INVERSE = {"dct2": "dct3", "dct3": "dct2", "dst2": "dst3", "dst3": "dst2"}


def _run(torch, kind, n, prec, x, hm=1, is_=1, id_=None, os_=1, od_=None):
    id_ = n * is_ if id_ is None else id_
    od_ = n * os_ if od_ is None else od_
    plan = P.Plan("r2r", n, -1, prec, hm, is_, id_, os_, od_, r2r=kind)
    xd = to_dev(torch, x.astype(RNP[prec]))
    yd = torch.zeros(span(n, os_, hm, od_), dtype=xd.dtype, device="cuda")
    plan.execute(xd, yd)
    return from_dev(torch, yd), plan


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 16, 17, 31, 64, 100, 127, 128, 1000,
                               4096, 6000, 8192])
def test_r2r_vs_oracle(torch_cuda, orc, prec, kind, n):
    torch = torch_cuda
    hm = max(1, min(8, (1 << 15) // n)) + 1
    x = rng.normal_f64(n + KINDS[kind], n * hm)
    y, plan = _run(torch, KINDS[kind], n, prec, x, hm)
    ref = orc.r2r_batched(KINDS[kind], x, n, hm, 1, n, 1, n, n * hm)
    assert relerr(y, ref) < TOL[prec] * 2, plan.describe()


@pytest.mark.parametrize("kind", list(KINDS))
def test_r2r_round_trip_factor(torch_cuda, kind):
    torch = torch_cuda
    for n in range(2, 129):
        x = rng.normal_f64(n, n)
        y, _ = _run(torch, KINDS[kind], n, "f64", x)
        z, _ = _run(torch, KINDS[INVERSE[kind]], n, "f64", y)
        assert relerr(z, 2 * n * x) < 1e-12, n


def test_dct2_first_element_is_twice_the_sum(torch_cuda):
    torch = torch_cuda
    for n in range(1, 129):
        x = (np.arange(n) % 5 - 2).astype(np.float32)
        y, _ = _run(torch, 1, n, "f32", x)
        assert abs(y[0] - 2 * x.sum()) <= 1e-5 * n


@pytest.mark.parametrize("kind", ["dct2", "dst2"])
def test_r2r_differentiable_layout(torch_cuda, orc, kind):
    torch = torch_cuda
    n, hm, is_, id_ = 4096, 12, 2, 2 * 4096 + 64
    x = rng.normal_f64(4, span(n, is_, hm, id_))
    y, plan = _run(torch, KINDS[kind], n, "f32", x, hm, is_, id_, 1, n)
    ref = orc.r2r_batched(KINDS[kind], x, n, hm, is_, id_, 1, n, n * hm)
    assert relerr(y, ref) < TOL["f32"] * 2, plan.describe()


# The other FFTW kinds (docs/plfft.md:186-230; test/r2r_test.cpp:11-43 runs
# every kind for n = 2..128): each through one complex DFT per member.
KINDS2 = {"dct1": 0, "dct4": 3, "dst1": 4, "dst4": 7, "dht": 8, "r2hc": 9, "hc2r": 10}
# kind -> (inverse kind, round-trip factor as a function of n)
ROUND = {"dct1": ("dct1", lambda n: 2 * (n - 1)), "dct4": ("dct4", lambda n: 2 * n),
         "dst1": ("dst1", lambda n: 2 * (n + 1)), "dst4": ("dst4", lambda n: 2 * n),
         "dht": ("dht", lambda n: n), "r2hc": ("hc2r", lambda n: n), "hc2r": ("r2hc", lambda n: n)}


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("kind", list(KINDS2))
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 16, 17, 31, 64, 100, 127, 128, 1000, 4096])
def test_r2r_other_kinds_vs_oracle(torch_cuda, orc, prec, kind, n):
    if kind == "dct1" and n == 1:
        pytest.skip("DCT-I is undefined for n = 1 (plan creation fails, test_api)")
    torch = torch_cuda
    # the oracle evaluates these kinds by their O(n^2) definitions
    hm = max(1, min(8, (1 << 15) // n)) + 1 if n <= 1000 else 2
    x = rng.normal_f64(n + 31 * KINDS2[kind], n * hm)
    y, plan = _run(torch, KINDS2[kind], n, prec, x, hm)
    ref = orc.r2r_batched(KINDS2[kind], x, n, hm, 1, n, 1, n, n * hm)
    assert relerr(y, ref) < TOL[prec] * 2, plan.describe()


@pytest.mark.parametrize("kind", list(KINDS2))
def test_r2r_other_kinds_round_trip(torch_cuda, kind):
    torch = torch_cuda
    inv, fac = ROUND[kind]
    for n in range(2, 129):
        x = rng.normal_f64(3 * n, n)
        y, _ = _run(torch, KINDS2[kind], n, "f64", x)
        z, _ = _run(torch, KINDS2[inv], n, "f64", y)
        assert relerr(z, fac(n) * x) < 1e-12, (kind, n)


def test_r2r_other_kinds_strided(torch_cuda, orc):
    """Input stride 2 with padded distance, output stride 3 (the batched
    layout rules of docs/plfft.md:136-170), gaps untouched."""
    torch = torch_cuda
    n, hm, is_, id_, os_, od_ = 300, 5, 2, 2 * 300 + 64, 3, 3 * 300 + 5
    for kind in KINDS2.values():
        x = rng.normal_f64(kind + 101, span(n, is_, hm, id_))
        plan = P.Plan("r2r", n, -1, "f64", hm, is_, id_, os_, od_, r2r=kind)
        xd = to_dev(torch, x)
        yd = torch.full((span(n, os_, hm, od_),), 7.5, dtype=torch.float64, device="cuda")
        plan.execute(xd, yd)
        y = from_dev(torch, yd)
        ref = orc.r2r_batched(kind, x, n, hm, is_, id_, os_, od_, span(n, os_, hm, od_))
        mask = np.zeros(y.size, bool)
        for b in range(hm):
            mask[b * od_ + np.arange(n) * os_] = True
        assert relerr(y[mask], ref[mask]) < 1e-12, kind
        assert np.all(y[~mask] == 7.5), kind

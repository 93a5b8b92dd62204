"""FP16 and BF16 storage with FP32 arithmetic (A30, and the bf16 extension).

Inputs are rounded to the storage type first; the FP64 oracle transforms
those rounded values, so the error measured is the GPU path's own (FP32
butterflies plus one rounding of the output to the storage type). Bounds:
f16 relative L2 <= 1e-3 (2^-11 output rounding), bf16 <= 8e-3 (2^-8).
"""
import numpy as np
import pytest

import plfft_amd as P
from plfft_amd import rng
from gpu_util import bf16_bits, bf16_to_f32, from_dev, relerr, to_dev

pytestmark = pytest.mark.gpu

HTOL = {"f16": 1e-3, "bf16": 8e-3}


def _encode(x, prec):
    """float array -> (storage uint16 array, exactly representable float64 values)"""
    x = np.asarray(x, np.float32)
    if prec == "f16":
        h = x.astype(np.float16)
        return h.view(np.uint16), h.astype(np.float64)
    b = bf16_bits(x)
    return b, bf16_to_f32(b).astype(np.float64)


def _decode(u, prec):
    if prec == "f16":
        return u.view(np.float16).astype(np.float64)
    return bf16_to_f32(u).astype(np.float64)


def _run(torch, plan, xs, out_reals):
    xd = to_dev(torch, xs)
    yd = torch.zeros(out_reals, dtype=torch.int16, device="cuda")
    plan.execute(xd, yd)
    return from_dev(torch, yd).view(np.uint16)


@pytest.mark.parametrize("prec", ["f16", "bf16"])
@pytest.mark.parametrize("n", [16, 32, 256, 1000, 4096, 15015, 65536, 97, 1009, 2053, 40009,
                               65537])
def test_c2c_half(torch_cuda, orc, prec, n):
    torch = torch_cuda
    hm = max(1, min(8, (1 << 16) // n)) + 1
    xs, xv = _encode(rng.normal_f64(n, 2 * n * hm), prec)
    for direction in (-1, 1):
        plan = P.Plan("c2c", n, direction, prec, howmany=hm)
        y = _decode(_run(torch, plan, xs, 2 * n * hm), prec)
        yc = y[0::2] + 1j * y[1::2]
        ref = orc.c2c_batched(xv[0::2] + 1j * xv[1::2], n, direction, hm, 1, n, 1, n, n * hm)
        assert relerr(yc, ref) < HTOL[prec], plan.describe()


@pytest.mark.parametrize("prec", ["f16", "bf16"])
@pytest.mark.parametrize("n", [64, 1000, 4096])
def test_r2c_c2r_half(torch_cuda, orc, prec, n):
    torch = torch_cuda
    hm = 5
    m = n // 2 + 1
    xs, xv = _encode(rng.normal_f64(n + 1, n * hm), prec)
    plan = P.Plan("r2c", n, -1, prec, howmany=hm)
    y = _decode(_run(torch, plan, xs, 2 * m * hm), prec)
    ref = orc.r2c_batched(xv, n, hm, 1, n, 1, m, m * hm)
    assert relerr(y[0::2] + 1j * y[1::2], ref) < HTOL[prec], plan.describe()
    # C2R of the (rounded) spectrum
    Xs, Xv = _encode(np.stack([ref.real, ref.imag], -1).ravel() / n, prec)
    plan2 = P.Plan("c2r", n, 1, prec, howmany=hm)
    z = _decode(_run(torch, plan2, Xs, n * hm), prec)
    zref = orc.c2r_batched(Xv[0::2] + 1j * Xv[1::2], n, hm, 1, m, 1, n, n * hm)
    assert relerr(z, zref) < HTOL[prec], plan2.describe()


@pytest.mark.parametrize("prec", ["f16", "bf16"])
@pytest.mark.parametrize("kind", ["dct2", "dst2", "dct3", "dst3"])
def test_r2r_half(torch_cuda, orc, prec, kind):
    torch = torch_cuda
    n, hm = 4096, 3
    k = {"dct2": 1, "dct3": 2, "dst2": 5, "dst3": 6}[kind]
    xs, xv = _encode(rng.normal_f64(k, n * hm) / 8, prec)
    plan = P.Plan("r2r", n, -1, prec, howmany=hm, r2r=kind)
    y = _decode(_run(torch, plan, xs, n * hm), prec)
    ref = orc.r2r_batched(k, xv, n, hm, 1, n, 1, n, n * hm)
    assert relerr(y, ref) < HTOL[prec], plan.describe()


@pytest.mark.parametrize("direction", [-1, 1])
def test_c2c_f16_matrix_core_stage(torch_cuda, orc, monkeypatch, direction):
    """The 4096-point f16 kernel whose first radix-16 stage runs on the
    matrix cores (f16 operands, DFT matrix split into f16 high and low
    parts, f32 accumulation), forced: it matches the FP64 oracle to the f16
    tolerance and the plain f32-butterfly kernel to within output rounding."""
    torch = torch_cuda
    n, hm = 4096, 37
    xs, xv = _encode(rng.normal_f64(11, 2 * n * hm), "f16")
    ref = orc.c2c_batched(xv[0::2] + 1j * xv[1::2], n, direction, hm, 1, n, 1, n, n * hm)
    outs = {}
    for name in ("_mfma", "n4096_r16x16x16_t256_b1_p5_split"):
        monkeypatch.setenv("PLFFT_KERNEL_FILTER", name)
        plan = P.Plan("c2c", n, direction, "f16", howmany=hm)
        assert name in plan.describe(), plan.describe()
        y = _decode(_run(torch, plan, xs, 2 * n * hm), "f16")
        outs[name] = y[0::2] + 1j * y[1::2]
        assert relerr(outs[name], ref) < HTOL["f16"], plan.describe()
    # the two kernels differ only in f32 rounding before the f16 store
    assert relerr(outs["_mfma"], outs["n4096_r16x16x16_t256_b1_p5_split"]) < 3e-4


@pytest.mark.parametrize("direction", [-1, 1])
def test_c2c_f16_n16_matrix_core(torch_cuda, orc, monkeypatch, direction):
    """n = 16 as one matrix-core product per 32 packed members, forced: it
    matches the FP64 oracle and the cross-lane kernel (f32 butterflies) to
    within output rounding, for a batch that leaves a partial group and more
    groups than resident waves; a strided batch does not take it."""
    torch = torch_cuda
    n, hm = 16, 300007
    xs, xv = _encode(rng.normal_f64(12, 2 * n * hm), "f16")
    ref = orc.c2c_batched(xv[0::2] + 1j * xv[1::2], n, direction, hm, 1, n, 1, n, n * hm)
    outs = {}
    for name in ("_mfma", "_xlane"):
        monkeypatch.setenv("PLFFT_KERNEL_FILTER", name)
        plan = P.Plan("c2c", n, direction, "f16", howmany=hm)
        assert name in plan.describe(), plan.describe()
        y = _decode(_run(torch, plan, xs, 2 * n * hm), "f16")
        outs[name] = y[0::2] + 1j * y[1::2]
        assert relerr(outs[name], ref) < HTOL["f16"], plan.describe()
    assert relerr(outs["_mfma"], outs["_xlane"]) < 3e-4
    monkeypatch.delenv("PLFFT_KERNEL_FILTER")
    plan = P.Plan("c2c", n, direction, "f16", howmany=8, idist=20, odist=20)
    assert "_mfma" not in plan.describe()

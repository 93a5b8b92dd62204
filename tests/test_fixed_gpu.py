"""Fixed-point Q0.15 / Q0.7 transforms on the GPU, bit-exact against the
oracle's restatement (oracle.c orc_fixed_c2c), through the C ABI.

The oracle itself is pinned on the CPU (tests/test_oracle.py) by the
reference's fixed-point checks: delta <-> constant known pairs and DFT/n
within 6 log2(n) LSB (test/delta_constant_test.hpp:40-62,
test/test_utils.hpp:132-151), and the generated device codelets are pinned
there against the oracle's interpreter of the reference's op lists. What
stays unpinned is bit parity with the reference's emitted NEON code after
its peephole rewrites (print_algo.cpp:236-973), which cannot be run here.
"""
import numpy as np
import pytest

import plfft_amd as P
from gpu_util import from_dev, to_dev

pytestmark = pytest.mark.gpu

DT = {15: ("q15", np.int16), 7: ("q7", np.int8)}
NS = [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768]


def _rand(seed, shape, bits):
    rng = np.random.default_rng(seed)
    hi = 1 << bits
    return rng.integers(-hi, hi, size=shape).astype(DT[bits][1])


def _oracle_batch(orc, x, n, count, direction, bits):
    """x: (count, n, 2) storage ints."""
    return np.stack([orc.fixed_c2c(x[b].astype(np.int32), direction, bits) for b in range(count)])


@pytest.mark.parametrize("bits", [15, 7])
@pytest.mark.parametrize("n", NS)
def test_fixed_c2c_bit_exact(orc, torch_cuda, bits, n):
    count = 3 if n <= 4096 else 2
    x = _rand(n * bits, (count, n, 2), bits)
    name, npdt = DT[bits]
    for direction in (P.PLFFT_FORWARD, P.PLFFT_BACKWARD):
        plan = P.Plan("c2c", n, direction, name, howmany=count)
        dx = to_dev(torch_cuda, x.reshape(-1))
        dy = torch_cuda.zeros_like(dx)
        plan.execute(dx, dy)
        y = from_dev(torch_cuda, dy).reshape(count, n, 2)
        ref = _oracle_batch(orc, x, n, count, direction, bits)
        assert np.array_equal(y.astype(np.int32), ref), (n, direction)


@pytest.mark.parametrize("bits", [15, 7])
def test_fixed_known_pairs(torch_cuda, bits):
    """Constant -> delta of amplitude 1 and delta -> constant 1/n (the
    reference's delta_constant tests for fixed point)."""
    name, npdt = DT[bits]
    one = (1 << bits) - 1
    for n in (16, 256, 4096):
        tol = 6 * int(np.log2(n))
        x = np.zeros((2, n, 2), npdt)
        x[0, :, 0] = one            # member 0: constant
        x[1, 0, 0] = one            # member 1: delta
        plan = P.Plan("c2c", n, P.PLFFT_FORWARD, name, howmany=2)
        dx = to_dev(torch_cuda, x.reshape(-1))
        dy = torch_cuda.zeros_like(dx)
        plan.execute(dx, dy)
        y = from_dev(torch_cuda, dy).reshape(2, n, 2).astype(np.int64)
        want0 = np.zeros((n, 2))
        want0[0, 0] = one
        assert np.abs(y[0] - want0).max() <= tol
        assert np.abs(y[1] - np.array([one / n, 0.0])).max() <= tol


@pytest.mark.parametrize("bits", [15, 7])
def test_fixed_strided_and_in_place(orc, torch_cuda, bits):
    name, npdt = DT[bits]
    n, count, istride, idist, ostride, odist = 256, 5, 2, 2 * 256 + 7, 3, 3 * 256 + 1
    isz = (count - 1) * idist + (n - 1) * istride + 1
    osz = (count - 1) * odist + (n - 1) * ostride + 1
    xin = _rand(77 + bits, (isz, 2), bits)
    sentinel = np.array([11, -13], npdt)
    plan = P.Plan("c2c", n, P.PLFFT_FORWARD, name, howmany=count, istride=istride, idist=idist,
                  ostride=ostride, odist=odist)
    dx = to_dev(torch_cuda, xin.reshape(-1))
    dy = to_dev(torch_cuda, np.tile(sentinel, osz))
    plan.execute(dx, dy)
    y = from_dev(torch_cuda, dy).reshape(osz, 2)
    touched = np.zeros(osz, bool)
    for b in range(count):
        xm = xin[b * idist + np.arange(n) * istride].astype(np.int32)
        ref = orc.fixed_c2c(xm, -1, bits)
        idx = b * odist + np.arange(n) * ostride
        assert np.array_equal(y[idx].astype(np.int32), ref)
        touched[idx] = True
    assert np.all(y[~touched] == sentinel)
    # in place, contiguous
    x = _rand(91 + bits, (4, 1024, 2), bits)
    plan = P.Plan("c2c", 1024, P.PLFFT_BACKWARD, name, howmany=4, may_alias=True)
    dz = to_dev(torch_cuda, x.reshape(-1))
    plan.execute(dz, dz)
    z = from_dev(torch_cuda, dz).reshape(4, 1024, 2).astype(np.int32)
    assert np.array_equal(z, _oracle_batch(orc, x, 1024, 4, P.PLFFT_BACKWARD, bits))


@pytest.mark.parametrize("bits", [15, 7])
@pytest.mark.parametrize("n", [2, 16, 256, 4096])
def test_fixed_r2c_c2r(orc, torch_cuda, bits, n):
    name, npdt = DT[bits]
    count = 3
    rng = np.random.default_rng(n + bits)
    hi = 1 << bits
    xr = rng.integers(-hi, hi, size=(count, n)).astype(npdt)
    plan = P.Plan("r2c", n, P.PLFFT_FORWARD, name, howmany=count)
    h = n // 2 + 1
    dx = to_dev(torch_cuda, xr.reshape(-1))
    dy = torch_cuda.zeros(count * h * 2, dtype=dx.dtype, device="cuda")
    plan.execute(dx, dy)
    y = from_dev(torch_cuda, dy).reshape(count, h, 2).astype(np.int32)
    for b in range(count):
        wide = np.stack([xr[b].astype(np.int32), np.zeros(n, np.int32)], -1)
        assert np.array_equal(y[b], orc.fixed_c2c(wide, -1, bits)[:h])
    # C2R of those bins
    plan = P.Plan("c2r", n, P.PLFFT_BACKWARD, name, howmany=count)
    dz = torch_cuda.zeros(count * n, dtype=dx.dtype, device="cuda")
    plan.execute(dy, dz)
    z = from_dev(torch_cuda, dz).reshape(count, n).astype(np.int32)
    for b in range(count):
        X = y[b]
        full = np.zeros((n, 2), np.int32)
        full[:h] = X
        full[0, 1] = 0
        if n % 2 == 0:
            full[n // 2, 1] = 0
        for k in range(h, n):
            full[k, 0] = X[n - k, 0]
            full[k, 1] = min(hi - 1, -int(X[n - k, 1]))
        assert np.array_equal(z[b], orc.fixed_c2c(full, 1, bits)[:, 0])


def test_fixed_adjoint_is_backward(orc, torch_cuda):
    n, count = 512, 2
    x = _rand(5, (count, n, 2), 15)
    plan = P.Plan("c2c", n, P.PLFFT_FORWARD, "q15", howmany=count)
    dx = to_dev(torch_cuda, x.reshape(-1))
    dy = torch_cuda.zeros_like(dx)
    plan.adjoint(dx, dy)
    y = from_dev(torch_cuda, dy).reshape(count, n, 2).astype(np.int32)
    assert np.array_equal(y, _oracle_batch(orc, x, n, count, P.PLFFT_BACKWARD, 15))


def test_fixed_host_buffers(orc, torch_cuda):
    n, count = 2048, 3
    x = _rand(8, (count, n, 2), 15)
    y = np.zeros_like(x)
    P.Plan("c2c", n, P.PLFFT_FORWARD, "q15", howmany=count).execute(x, y)
    assert np.array_equal(y.astype(np.int32), _oracle_batch(orc, x, n, count, -1, 15))


@pytest.mark.parametrize("n,msg", [(12, "power-of-two"), (97, "power-of-two"),
                                   (65536, "32768")])
def test_fixed_unsupported_lengths(torch_cuda, n, msg):
    with pytest.raises(P.PlfftError) as e:
        P.Plan("c2c", n, P.PLFFT_FORWARD, "q15")
    assert e.value.status == P.PLFFT_PLAN_CREATION_FAILED
    assert msg in str(e.value)


@pytest.mark.parametrize("bits", [15, 7])
def test_fixed_mixed_batch(orc, torch_cuda, bits):
    """A fixed-point mixed-length batch (power-of-two lengths, the
    reference's fixed-point domain) equals one fixed-point plan per member,
    bit for bit, in both directions: one gathered launch per kernel tier,
    blocks of one length each, lengths interleaved in the batch."""
    name, npdt = DT[bits]
    rngl = np.random.default_rng(3)
    lengths = [int(v) for v in 2 ** rngl.integers(0, 16, size=40)] + [1, 2, 32768, 1, 16, 16]
    offs = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(int)
    x = _rand(200 + bits, (sum(lengths), 2), bits)
    for direction in (P.PLFFT_FORWARD, P.PLFFT_BACKWARD):
        plan = P.MixedPlan(lengths, direction, name)
        d = plan.describe()
        assert "mixed_fixed" in d and "launches=5" in d, d
        dx = to_dev(torch_cuda, x.reshape(-1))
        dy = torch_cuda.zeros_like(dx)
        plan.execute(dx, dy)
        y = from_dev(torch_cuda, dy).reshape(-1, 2).astype(np.int32)
        for n, o in zip(lengths, offs):
            ref = orc.fixed_c2c(x[o:o + n].astype(np.int32), direction, bits)
            assert np.array_equal(y[o:o + n], ref), (n, direction)
    with pytest.raises(P.PlfftError):
        P.MixedPlan([16, 12], P.PLFFT_FORWARD, name)

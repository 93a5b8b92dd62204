"""Every compile-time C2C kernel variant the planner may choose, forced one
at a time (PLFFT_KERNEL_FILTER picks the first candidate whose name holds
the string) and checked against the FP64 oracle in both directions.

Under the default MEASURE rigor the planner times the candidates of a
length and keeps the fastest, so which variant runs can differ from box to
box; this runs each of them. The batch sizes leave a partial last block
(several members per block) and give the persistent (pipelined) kernels
more members than resident blocks.
"""
import numpy as np
import pytest

import plfft_amd as P
from plfft_amd import rng
from gpu_util import TOL, from_dev, relerr, to_dev

pytestmark = pytest.mark.gpu

# (n, name substring of the variant)
VARIANTS = [
    (16, "n16_r2x2x2x2_t1_b256_p5_xlane"),
    (32, "n32_r2x2x2x2x2_t2_b128_p5_xlane"),
    (16, "n16_r16_t1_b256"),
    (5, "n5_r5_t1_b256_p5_staged"),
    (17, "n17_r17_t1_b256_p5_staged"),
    (23, "n23_r23_t1_b256_p5_staged"),
    (32, "n32_r8x4_t4_b64"),
    (64, "n64_r16x4_t4_b64"),
    (256, "n256_r16x16_t16_b16_p4"),
    (256, "n256_r16x16_t8_b32_p4_pair"),
    (1024, "n1024_r16x16x4_t64_b4_p4"),
    (1024, "n1024_r16x16x4_t32_b8_p4_pair"),
    (2048, "n2048_r8x8x8x4_t128_b2_p4_pair"),
    (4096, "n4096_r16x16x16_t256_b1_p4"),
    (4096, "n4096_r16x16x16_t256_b1_p5"),
    (4096, "n4096_r16x16x16_t128_b2_p4_pair"),
    (4096, "n4096_r8x8x8x8_t256_b1_p4_pair"),
    (4096, "n4096_r8x8x8x8_t256_b2_p4_pair"),
    (4096, "_pipe4"),
    (4096, "_pipe3"),
    (4096, "n4096_r16x16x16_t256_b1_p5_split"),
    (8192, "n8192_r16x16x8x4_t512_b1_p5_split"),
    (16384, "n16384_r16x16x16x4_t1024_b1_p4"),
    (16384, "n16384_r16x16x16x4_t512_b1_p4_pair"),
    (16384, "n16384_r16x16x16x4_t1024_b1_p5_split"),
    (16384, "n16384_r4x16x16x16_t1024_b1_p5_split"),
    (16384, "n16384_r16x16x16x4_t1024_b1_p5_split_w8"),
    (16384, "n16384_r4x16x16x16_t1024_b1_p5_split_w8"),
    (32768, "n32768_r16x16x16x8_t1024_b1_p5_split"),
]


@pytest.mark.parametrize("direction", [-1, 1])
@pytest.mark.parametrize("n,name", VARIANTS)
def test_forced_variant_vs_oracle(torch_cuda, orc, monkeypatch, n, name, direction):
    torch = torch_cuda
    monkeypatch.setenv("PLFFT_KERNEL_FILTER", name)
    hm = 1100 if n <= 4096 else 37
    x = rng.complex_normal(n + hm, n * hm).astype(np.complex64)
    plan = P.Plan("c2c", n, direction, "f32", howmany=hm)
    assert name in plan.describe(), plan.describe()
    xd = to_dev(torch, x)
    yd = torch.empty_like(xd)
    plan.execute(xd, yd)
    y = from_dev(torch, yd)
    ref = orc.c2c_batched(x.astype(np.complex128), n, direction, hm, 1, n, 1, n, n * hm)
    assert relerr(y, ref) < TOL["f32"], plan.describe()


@pytest.mark.parametrize("n", [16, 32])
def test_cross_lane_kernel_layouts(torch_cuda, orc, n):
    """The cross-lane kernels take packed members only: a batch at distance
    n (and in place) runs them, a wider distance falls back to the LDS
    kernels; all match the oracle."""
    torch = torch_cuda
    hm = 4099  # a partial last wave
    x = rng.complex_normal(7 * n + hm, (n + 3) * hm).astype(np.complex64)
    for dist in (n, n + 3):
        plan = P.Plan("c2c", n, -1, "f32", howmany=hm, idist=dist, odist=dist)
        assert ("xlane" in plan.describe()) == (dist == n), plan.describe()
        xd = to_dev(torch, x[:dist * hm])
        yd = torch.zeros_like(xd)
        plan.execute(xd, yd)
        y = from_dev(torch, yd)
        ref = orc.c2c_batched(x[:dist * hm].astype(np.complex128), n, -1, hm, 1, dist, 1, dist,
                              dist * hm)
        assert relerr(y, ref) < TOL["f32"], plan.describe()
    plan = P.Plan("c2c", n, 1, "f32", howmany=hm, may_alias=True)
    xd = to_dev(torch, x[:n * hm])
    plan.execute(xd, xd)
    ref = orc.c2c_batched(x[:n * hm].astype(np.complex128), n, 1, hm, 1, n, 1, n, n * hm)
    assert relerr(from_dev(torch, xd), ref) < TOL["f32"]

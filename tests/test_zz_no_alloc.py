"""plfft_plan_execute_noalloc (include/plfft_gpu.h) against plfft_plan_execute_async.

The no-alloc entry point runs the same kernels as plfft_plan_execute_async with
the plan's scratch taken from the library's fixed device array; four-step plans
shrink their batch chunk to fit it. The transform of each member does not
depend on the chunk, so the two entry points must agree bit for bit. Two calls
in a row on one device, the second smaller, check that the shared array's
leftover contents do not leak into a result.
"""
import numpy as np
import pytest

from gpu_util import relerr, to_dev, from_dev

pytestmark = pytest.mark.gpu


def _case(torch, P, n, howmany, direction, seed, in_place=False):
    from plfft_amd import rng
    print(f"noalloc case: c2c f32 n={n} howmany={howmany} dir={direction} "
          f"in_place={in_place}", flush=True)
    plan = P.Plan("c2c", n, direction, "f32", howmany=howmany, may_alias=in_place,
                  rigor=P.PLFFT_ESTIMATE)
    x = rng.complex_uniform(seed, n * howmany)
    xd = to_dev(torch, x)
    ref = torch.empty_like(xd)
    plan.execute_async(xd, ref)  # the plan's own workspace
    if in_place:
        yd = xd.clone()
        plan.execute_noalloc(yd, yd)
    else:
        yd = torch.empty_like(xd)
        plan.execute_noalloc(xd, yd)
        assert np.array_equal(from_dev(torch, xd), x)  # input untouched
    torch.cuda.synchronize()
    return plan, x, from_dev(torch, ref), from_dev(torch, yd)


def test_noalloc_matches_async_two_calls(torch_cuda, orc):
    import plfft_amd as P
    torch = torch_cuda
    # 2^20 points x 12: the plan's own workspace (12 transforms, 96 MiB) is
    # larger than the 64 MiB array, so the no-alloc call runs chunks of 8 + 4
    plan, x, ref, y = _case(torch, P, 1 << 20, 12, P.PLFFT_FORWARD, 11)
    assert plan.workspace_size > P.PLFFT_NOALLOC_SCRATCH_BYTES
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32)), "first call differs"
    # the second call is smaller (65536 points x 5, backward, in place)
    n, hm = 65536, 5
    plan2, x2, ref2, y2 = _case(torch, P, n, hm, P.PLFFT_BACKWARD, 12, in_place=True)
    assert plan2.workspace_size > 0  # a four-step plan: it uses the scratch
    assert np.array_equal(y2.view(np.uint32), ref2.view(np.uint32)), "second call differs"
    want = orc.c2c_batched(x2.astype(np.complex128), n, +1, hm, 1, n, 1, n, n * hm)
    assert relerr(y2, want) < 2e-6


def test_noalloc_rader_three_pass_chunks(torch_cuda, orc):
    """Rader in three passes (n = 65537): 200 members need 105 MB of
    workspace with their (x_0, A_0) slots, so the no-alloc call runs chunks
    of 128 + 72 members, whose slots follow the chunk's workspace."""
    import plfft_amd as P
    torch = torch_cuda
    n, hm = 65537, 200
    plan, x, ref, y = _case(torch, P, n, hm, P.PLFFT_FORWARD, 14)
    assert "rader3" in plan.describe()
    assert plan.workspace_size > P.PLFFT_NOALLOC_SCRATCH_BYTES
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))
    k = 3  # a few members against the oracle
    want = orc.c2c_batched(x[:n * k].astype(np.complex128), n, -1, k, 1, n, 1, n, n * k)
    assert relerr(y[:n * k], want) < 4e-6


def test_noalloc_single_kernel_plan(torch_cuda):
    import plfft_amd as P
    torch = torch_cuda
    plan, x, ref, y = _case(torch, P, 4096, 7, P.PLFFT_FORWARD, 13)
    assert plan.workspace_size == 0
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))


def test_noalloc_rejects_plans_beyond_the_array(torch_cuda):
    import plfft_amd as P
    # one 2^24-point f32 transform needs 128 MiB of four-step scratch
    print("noalloc case: c2c f32 n=2^24 rejection (nothing launched)", flush=True)
    plan = P.Plan("c2c", 1 << 24, P.PLFFT_FORWARD, "f32", rigor=P.PLFFT_ESTIMATE)
    lib = P.load()
    # real buffers of the plan's size, so that nothing could fault even if a
    # launch happened
    buf = torch_cuda.zeros(1 << 25, dtype=torch_cuda.float32, device="cuda")
    st = lib.plfft_plan_execute_noalloc(plan.h, buf.data_ptr(), buf.data_ptr(), 0)
    torch_cuda.cuda.synchronize()
    assert st == P.PLFFT_INVALID_SIZE
    assert bool((buf == 0).all())  # nothing ran
    assert "PLFFT_NOALLOC_SCRATCH_BYTES" in P.last_error()
    assert lib.plfft_plan_execute_noalloc(None, 1, 1, 0) == P.PLFFT_NULL_ARGUMENT

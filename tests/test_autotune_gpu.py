"""The timed planner on the GPU: wisdom reproduces a kernel choice, and
ESTIMATE ignores wisdom (registry order, deterministic)."""
import os

import pytest

import plfft_amd as P

pytestmark = pytest.mark.gpu


def _device_tag(torch):
    p = torch.cuda.get_device_properties(0)
    return f"{p.gcnArchName}/{p.multi_processor_count}"


def test_plan_reproduces_from_wisdom(torch_cuda, tmp_path):
    torch = torch_cuda
    n = 4096
    default = P.Plan("c2c", n, P.PLFFT_FORWARD, "f32", 64, rigor=P.PLFFT_ESTIMATE).describe()
    # a wisdom file naming another compiled candidate of the same family
    other = "c2c_unit_f32_fwd_n4096_r16x16x16_t256_b1_p5"
    if other in default:
        other = "c2c_unit_f32_fwd_n4096_r16x16x16_t256_b1_p4"
    path = tmp_path / "wisdom.txt"
    path.write_text(f"c2c_unit:{n}:f32:f:{_device_tag(torch)} {other}\n")
    old = os.environ.get("PLFFT_WISDOM")
    os.environ["PLFFT_WISDOM"] = str(path)
    try:
        print(f"wisdom case: {path} -> {other}", flush=True)
        measured = P.Plan("c2c", n, P.PLFFT_FORWARD, "f32", 64, rigor=P.PLFFT_MEASURE).describe()
        assert other in measured, measured
        assert P.Plan("c2c", n, P.PLFFT_FORWARD, "f32", 64,
                      rigor=P.PLFFT_ESTIMATE).describe() == default
    finally:
        if old is None:
            del os.environ["PLFFT_WISDOM"]
        else:
            os.environ["PLFFT_WISDOM"] = old

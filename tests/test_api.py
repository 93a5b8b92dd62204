"""The C-ABI boundary without a GPU: exports, defaults, setters, validation
order and status strings, mirroring test/plfft_interface_test.cpp of the
reference (defaults :278-299, setters :301-330, null handling :462-489,
size/batch errors :492-555, enum errors :557-643, status strings :645-660).
"""
import ctypes
import os
import re

import pytest

import plfft_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for h in ("plfft.h", "plfft_gpu.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(plfft_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(lib):
    declared = _declared_functions()
    assert len(declared) >= 23 + 6
    assert sorted(P.EXPORTS) == declared
    for name in declared:
        assert getattr(lib, name) is not None


def test_every_host_kernel_has_device_code():
    """Each kernel the host side can launch (a __device_stub__) has a gfx950
    kernel descriptor (<name>.kd) in the embedded code object; a stale or
    half-rebuilt object would otherwise fail only at launch time."""
    data = open(P.LIB_PATH, "rb").read()
    prefix = b"__device_stub__"
    kds = set(re.findall(rb"(_ZN3plf[A-Za-z0-9_]+)\.kd\x00", data))
    stubs = set(re.findall(rb"_ZN3plf(\d+)__device_stub__([A-Za-z0-9_]+)\x00", data))
    assert len(stubs) > 100
    missing = []
    for length, body in stubs:
        k = int(length) - len(prefix)  # identifier length of the kernel's own name
        kernel = b"_ZN3plf" + str(k).encode() + body
        if kernel not in kds:
            missing.append(kernel.decode())
    assert not missing, missing[:5]


def test_reference_abi_function_count():
    text = open(os.path.join(ROOT, "include", "plfft.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    assert len(set(re.findall(r"\b(plfft_[a-z0-9_]+)\s*\(", text))) == 23


def _new(lib, kind="c2c", n=8, arg=P.PLFFT_FORWARD):
    h = ctypes.c_void_p()
    if kind == "c2c":
        st = lib.plfft_config_create(ctypes.byref(h), n, arg)
    elif kind == "r2c":
        st = lib.plfft_config_create_r2c(ctypes.byref(h), n)
    elif kind == "c2r":
        st = lib.plfft_config_create_c2r(ctypes.byref(h), n)
    else:
        st = lib.plfft_config_create_r2r(ctypes.byref(h), n, arg)
    return st, h


@pytest.mark.parametrize("kind,arg", [("c2c", -1), ("c2c", 1), ("r2c", 0), ("c2r", 0), ("r2r", 9)])
def test_config_defaults(lib, kind, arg):
    st, h = _new(lib, kind, 8, arg)
    assert st == P.PLFFT_OK and h.value
    assert lib.plfft_config_get_data_type(h) == P.DATA_TYPES["f32"]
    assert lib.plfft_config_get_io_alias(h) == P.PLFFT_IO_NO_ALIAS
    assert lib.plfft_config_get_sme_mode(h) == P.PLFFT_SME_DISABLED
    assert lib.plfft_config_get_rigor(h) == P.PLFFT_MEASURE
    assert lib.plfft_config_get_howmany(h) == 1
    assert lib.plfft_config_get_istride(h) == 1
    assert lib.plfft_config_get_ostride(h) == 1
    assert lib.plfft_config_get_idist(h) == 8
    assert lib.plfft_config_get_odist(h) == 8
    lib.plfft_config_destroy(h)


def test_setters_store_without_validation(lib):
    st, h = _new(lib)
    lib.plfft_config_set_data_type(h, 123)
    lib.plfft_config_set_io_alias(h, 77)
    lib.plfft_config_set_sme_mode(h, 55)
    lib.plfft_config_set_rigor(h, 99)
    lib.plfft_config_set_batch(h, -3, 0, -1, 7, 9)
    assert lib.plfft_config_get_data_type(h) == 123
    assert lib.plfft_config_get_io_alias(h) == 77
    assert lib.plfft_config_get_sme_mode(h) == 55
    assert lib.plfft_config_get_rigor(h) == 99
    assert [lib.plfft_config_get_howmany(h), lib.plfft_config_get_istride(h),
            lib.plfft_config_get_idist(h), lib.plfft_config_get_ostride(h),
            lib.plfft_config_get_odist(h)] == [-3, 0, -1, 7, 9]
    lib.plfft_config_destroy(h)


def test_null_handling(lib):
    assert lib.plfft_config_create(None, 8, -1) == P.PLFFT_NULL_ARGUMENT
    assert lib.plfft_config_create_r2c(None, 8) == P.PLFFT_NULL_ARGUMENT
    assert lib.plfft_config_create_c2r(None, 8) == P.PLFFT_NULL_ARGUMENT
    assert lib.plfft_config_create_r2r(None, 8, 9) == P.PLFFT_NULL_ARGUMENT
    assert lib.plfft_plan_create_from_config(None, None) == P.PLFFT_NULL_ARGUMENT
    plan = ctypes.c_void_p(12345)
    assert lib.plfft_plan_create_from_config(ctypes.byref(plan), None) == P.PLFFT_NULL_ARGUMENT
    assert not plan.value  # failed creation leaves the plan null
    lib.plfft_config_destroy(None)
    lib.plfft_plan_destroy(None)


def test_invalid_kind_and_direction_leave_config_null(lib):
    h = ctypes.c_void_p(1)
    assert lib.plfft_config_create(ctypes.byref(h), 8, 123) == P.PLFFT_INVALID_TRANSFORM_KIND
    assert not h.value
    h = ctypes.c_void_p(1)
    assert lib.plfft_config_create_r2r(ctypes.byref(h), 8, 123) == P.PLFFT_INVALID_TRANSFORM_KIND
    assert not h.value
    h = ctypes.c_void_p(1)
    assert lib.plfft_config_create(ctypes.byref(h), 8, 0) == P.PLFFT_INVALID_TRANSFORM_KIND


def _plan_status(lib, h):
    plan = ctypes.c_void_p(999)
    st = lib.plfft_plan_create_from_config(ctypes.byref(plan), h)
    if st != P.PLFFT_OK:
        assert not plan.value
    else:
        lib.plfft_plan_destroy(plan)
    return st


def test_size_and_batch_errors(lib):
    for n in (0, -1):
        st, h = _new(lib, n=n)
        assert _plan_status(lib, h) == P.PLFFT_INVALID_SIZE
        lib.plfft_config_destroy(h)
    st, h = _new(lib, n=1)
    for batch in [(0, 1, 1, 1, 1), (-1, 1, 1, 1, 1), (2, 0, 2, 1, 2), (2, 1, 0, 1, 2),
                  (2, 1, 1, 0, 2), (2, 1, 1, 1, 0), (1, 0, 5, 1, 5), (1, 1, 5, -1, 5)]:
        lib.plfft_config_set_batch(h, *batch)
        assert _plan_status(lib, h) == P.PLFFT_UNSUPPORTED_BATCH, batch
    lib.plfft_config_destroy(h)


def test_enum_errors_in_validation_order(lib):
    st, h = _new(lib)
    lib.plfft_config_set_io_alias(h, 123)
    assert _plan_status(lib, h) == P.PLFFT_INVALID_IO_ALIAS
    lib.plfft_config_set_io_alias(h, P.PLFFT_IO_NO_ALIAS)
    lib.plfft_config_set_sme_mode(h, 123)
    assert _plan_status(lib, h) == P.PLFFT_INVALID_SME_MODE
    lib.plfft_config_set_sme_mode(h, P.PLFFT_SME_DISABLED)
    lib.plfft_config_set_data_type(h, 123)
    assert _plan_status(lib, h) == P.PLFFT_INVALID_DATA_TYPE
    # rigor is checked before the data type (src/plfft.cpp:484-513)
    lib.plfft_config_set_rigor(h, 9)
    assert _plan_status(lib, h) == P.PLFFT_INVALID_PLANNER_RIGOR
    # size beats everything after the null checks
    lib.plfft_config_set_batch(h, 0, 1, 1, 1, 1)
    assert _plan_status(lib, h) == P.PLFFT_UNSUPPORTED_BATCH
    lib.plfft_config_destroy(h)


def test_sme_statuses(lib):
    st, h = _new(lib, "r2r", 8, 9)
    lib.plfft_config_set_sme_mode(h, P.PLFFT_SME_ENABLED)
    assert _plan_status(lib, h) == P.PLFFT_SME_UNSUPPORTED_R2R
    lib.plfft_config_destroy(h)
    for dt in ("q7", "q15"):
        st, h = _new(lib)
        lib.plfft_config_set_data_type(h, P.DATA_TYPES[dt])
        lib.plfft_config_set_sme_mode(h, P.PLFFT_SME_ENABLED)
        assert _plan_status(lib, h) == P.PLFFT_SME_UNSUPPORTED_FIXED_POINT
        lib.plfft_config_destroy(h)
    st, h = _new(lib)
    lib.plfft_config_set_sme_mode(h, P.PLFFT_SME_ENABLED)
    assert _plan_status(lib, h) == P.PLFFT_SME_NOT_ENABLED  # built without SME
    lib.plfft_config_destroy(h)


def test_fixed_point_r2r_fails_plan_creation(lib):
    st, h = _new(lib, "r2r", 8, 1)
    lib.plfft_config_set_data_type(h, P.DATA_TYPES["q15"])
    assert _plan_status(lib, h) == P.PLFFT_PLAN_CREATION_FAILED
    lib.plfft_config_destroy(h)


def test_status_strings(lib):
    expect = {0: "success", 1: "null argument", 2: "invalid transform size",
              3: "Unsupported batch or stride configuration", 4: "invalid transform kind",
              5: "invalid data type", 6: "invalid I/O alias mode", 7: "plan creation failed",
              8: "invalid SME mode", 14: "invalid planning rigor level"}
    for k, v in expect.items():
        assert lib.plfft_status_to_string(k).decode() == v
    for k in range(15):
        assert lib.plfft_status_to_string(k)
    assert lib.plfft_status_to_string(999).decode() == "unknown status"
    assert lib.plfft_status_to_string(-5).decode() == "unknown status"


def test_plan_without_gpu_fails_cleanly(lib):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    st, h = _new(lib, n=1024)
    assert _plan_status(lib, h) == P.PLFFT_PLAN_CREATION_FAILED
    assert "device" in P.last_error() or P.last_error()
    lib.plfft_config_destroy(h)


def _mixed_status(lib, lengths, direction=-1, dtype=0, rigor=1, in_offs=None, out_offs=None,
                  null_plan=False, null_n=False):
    count = len(lengths)
    arr = (ctypes.c_int64 * max(count, 1))(*lengths)
    io = (ctypes.c_int64 * max(count, 1))(*in_offs) if in_offs is not None else None
    oo = (ctypes.c_int64 * max(count, 1))(*out_offs) if out_offs is not None else None
    h = ctypes.c_void_p()
    st = lib.plfft_mixed_plan_create(
        None if null_plan else ctypes.byref(h), count,
        None if null_n else ctypes.cast(arr, ctypes.c_void_p),
        ctypes.cast(io, ctypes.c_void_p) if io is not None else None,
        ctypes.cast(oo, ctypes.c_void_p) if oo is not None else None,
        direction, dtype, rigor)
    if st == P.PLFFT_OK:
        lib.plfft_mixed_plan_destroy(h)
    else:
        assert not h.value
    return st


def test_mixed_plan_argument_validation(lib):
    """Argument errors of plfft_mixed_plan_create come before any device
    work, in this order: null pointers, sizes/offsets, direction, data
    type, rigor."""
    assert _mixed_status(lib, [8], null_plan=True) == P.PLFFT_NULL_ARGUMENT
    assert _mixed_status(lib, [8], null_n=True) == P.PLFFT_NULL_ARGUMENT
    assert _mixed_status(lib, []) == P.PLFFT_INVALID_SIZE
    assert _mixed_status(lib, [8, 0]) == P.PLFFT_INVALID_SIZE
    assert _mixed_status(lib, [8, -3]) == P.PLFFT_INVALID_SIZE
    assert _mixed_status(lib, [8, 4], in_offs=[0, -1]) == P.PLFFT_INVALID_SIZE
    assert _mixed_status(lib, [8, 4], out_offs=[-8, 0]) == P.PLFFT_INVALID_SIZE
    assert _mixed_status(lib, [0], direction=7) == P.PLFFT_INVALID_SIZE
    assert _mixed_status(lib, [8], direction=0) == P.PLFFT_INVALID_TRANSFORM_KIND
    assert _mixed_status(lib, [8], direction=2, dtype=99) == P.PLFFT_INVALID_TRANSFORM_KIND
    assert _mixed_status(lib, [8], dtype=99) == P.PLFFT_INVALID_DATA_TYPE
    assert _mixed_status(lib, [8], dtype=-1) == P.PLFFT_INVALID_DATA_TYPE
    # fixed point: power-of-two lengths only (plan creation fails for 12)
    assert _mixed_status(lib, [8, 12], dtype=P.DATA_TYPES["q15"]) == P.PLFFT_PLAN_CREATION_FAILED
    assert _mixed_status(lib, [8], dtype=99, rigor=9) == P.PLFFT_INVALID_DATA_TYPE
    assert _mixed_status(lib, [8], rigor=4) == P.PLFFT_INVALID_PLANNER_RIGOR
    assert _mixed_status(lib, [8], rigor=-1) == P.PLFFT_INVALID_PLANNER_RIGOR


def test_mixed_plan_without_gpu_fails_cleanly(lib):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    assert _mixed_status(lib, [8, 97, 5000]) == P.PLFFT_PLAN_CREATION_FAILED
    assert "device" in P.last_error()
    # null-safe queries
    assert lib.plfft_mixed_plan_get_workspace_size(None) == 0
    assert lib.plfft_mixed_plan_get_span(None, 0) == 0
    assert lib.plfft_mixed_plan_execute(None, None, None) == P.PLFFT_NULL_ARGUMENT
    assert lib.plfft_mixed_plan_execute_async(None, None, None, None, None) == P.PLFFT_NULL_ARGUMENT
    lib.plfft_mixed_plan_destroy(None)


def test_planner_budget_and_margin_follow_the_reference(lib):
    """Rigor -> planning budget (seconds) and decision margin, the reference's
    get_planner_target_secs_total / get_planner_margin (src/plfft.cpp:149-184):
    ESTIMATE plans without timing; n < 1500 shares one curve; MEASURE,
    PATIENT and EXHAUSTIVE grow with n on their own curves. The library's
    internal C++ functions are reached by their mangled names (no GPU)."""
    budget = getattr(lib, "_ZN3plf19planner_budget_secsEil")
    budget.restype, budget.argtypes = ctypes.c_double, [ctypes.c_int, ctypes.c_long]
    margin = getattr(lib, "_ZN3plf14planner_marginEi")
    margin.restype, margin.argtypes = ctypes.c_double, [ctypes.c_int]
    # values of the reference's formulas at these lengths
    want = {
        (1, 64): 2.27690966431604e-05, (1, 1499): 0.000179036427437457,
        (1, 1500): 0.00019068291013525106, (1, 4096): 0.0004982782054794112,
        (1, 65536): 0.007060759501285816, (1, 1 << 20): 0.10005319154393118,
        (2, 1499): 0.000179036427437457, (2, 1500): 0.00882967218749666,
        (2, 4096): 0.017889485134077232, (2, 1 << 20): 0.8817586184232925,
        (3, 1024): 0.00013954679201450457, (3, 1500): 0.31054855123728353,
        (3, 65536): 2.3375085797605277, (3, 1 << 20): 10.28572555746398,
    }
    for (r, n), v in want.items():
        assert budget(r, n) == pytest.approx(v, rel=1e-12), (r, n)
    for n in (1, 64, 4096, 1 << 20):
        assert budget(P.PLFFT_ESTIMATE, n) == 0.0
    assert [margin(r) for r in range(4)] == [0.01, 0.02, 0.05, 0.5]

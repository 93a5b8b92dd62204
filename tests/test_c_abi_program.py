"""A C program (tests/c_abi/c_abi_main.c) that calls every function of
include/plfft.h, compiled with gcc as plain C99 and linked against the
library, as the reference's test/c_compat does for its C interface. Without a
GPU it checks configs, defaults, setters and status strings, and that plan
creation fails cleanly; on the GPU it also runs transforms on host buffers."""
import os
import subprocess

import pytest

import plfft_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path):
    exe = tmp_path / "c_abi_main"
    libdir = os.path.dirname(P.LIB_PATH)
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-O1",
           os.path.join(ROOT, "tests", "c_abi", "c_abi_main.c"),
           "-I", os.path.join(ROOT, "include"), "-L", libdir, "-lplfft_amd", "-lm",
           "-Wl,-rpath," + libdir, "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_c_program_without_gpu(tmp_path):
    exe = _build(tmp_path)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] in ("no device", "ok")


@pytest.mark.gpu
def test_c_program_on_gpu(tmp_path, torch_cuda):
    exe = _build(tmp_path)
    print("c program: all 23 functions, host buffers", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "ok", r.stdout

"""The C-ABI used from a C program (tests/c/abi_smoke.c): it compiles
warning-free against include/vt.h and libvt.so here; on a GPU it runs and
checks a render, the centre pick, an edit and the error statuses."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "unnamed-voxel-tracer_amd", "lib")
SRC = os.path.join(ROOT, "tests", "c", "abi_smoke.c")


def build(out):
    cmd = ["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", SRC, "-L", LIB,
           "-lvt", "-L", "/opt/rocm/lib", "-lamdhip64", "-lm", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_c_program_compiles(tmp_path):
    assert os.path.exists(build(tmp_path / "abi_smoke"))


@pytest.mark.gpu
def test_c_program_runs_on_gpu(tmp_path):
    from conftest import _has_gpu
    if not _has_gpu():
        pytest.skip("no HIP device")
    exe = build(tmp_path / "abi_smoke")
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = LIB + ":/opt/rocm/lib:" + env.get("LD_LIBRARY_PATH", "")
    print("[c-abi] running abi_smoke", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0 and "ABI SMOKE OK" in r.stdout, r.stdout + r.stderr

"""The reference-side adapter (vedo_amd/adapter/GpuSimMediator.h) compiles
against the reference's own headers with its build macros, links with
libvedo_hip.so and the reference compiled in place (oracle/_ref), builds a
DOWorld with the reference constructors and hands it to the library; without
a GPU it fails with vg_create's error instead of crashing or falling back.
Needs the reference sources and oracle/_ref (the build container only)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libvedo_ref.a")
LIB_DIR = os.path.join(ROOT, "vedo_amd")


@pytest.mark.skipif(not (os.path.isdir(REF_SRC) and os.path.exists(REF_LIB)
                         and os.path.exists(os.path.join(LIB_DIR, "libvedo_hip.so"))),
                    reason="needs the reference sources, oracle/_ref and the built library")
def test_adapter_compiles_links_and_fails_cleanly_without_gpu(tmp_path):
    exe = str(tmp_path / "adapter_main")
    cmd = ["g++", "-std=c++11", "-O1", "-D_UNIX_LIKE", "-D_INT32", "-D_FLOAT64", "-D_STD_CPP_11", "-w",
           "-I", REF_SRC, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "vedo_amd", "adapter"),
           os.path.join(ROOT, "tests", "adapter", "adapter_main.cpp"), REF_LIB,
           "-L", LIB_DIR, "-lvedo_hip", f"-Wl,-rpath,{LIB_DIR}", "-o", exe]
    subprocess.run(cmd, check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    print(out.stdout)
    gpu = os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK)
    if gpu:
        assert out.stdout.startswith("stepped 0.0002"), out.stdout
    else:
        assert out.stdout.startswith("refused: ") and len(out.stdout.strip()) > len("refused: "), out.stdout

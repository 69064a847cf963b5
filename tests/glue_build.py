"""Build of tests/abi_glue.cpp (the C++ caller of the C ABI) against
include/semtex_amd.h and semtex_amd/libsemtex_amd.so, with g++ (host code
only: it calls the library and the HIP runtime API)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "abi_glue.cpp")
BIN = os.path.join(ROOT, "tests", "_bin", "abi_glue")


def build_glue(force=False):
    if os.path.exists(BIN) and not force and os.path.getmtime(BIN) >= os.path.getmtime(SRC):
        return BIN
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", SRC,
           "-o", BIN + ".tmp", "-L", os.path.join(ROOT, "semtex_amd"), "-lsemtex_amd",
           "-L", "/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../semtex_amd", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    os.replace(BIN + ".tmp", BIN)
    return BIN

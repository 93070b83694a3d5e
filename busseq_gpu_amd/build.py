"""Build recipe of libbusseq (gfx950) and the BUSseq_gpu executable.

    python -m busseq_gpu_amd.build        # or busseq_gpu_amd.build.build()

Outputs, in-tree:
    busseq_gpu_amd/lib/libbusseq.so    C ABI of include/busseq.h
    busseq_gpu_amd/bin/BUSseq_gpu      drop-in executable of the reference CLI

All GPU code is compiled for gfx950 with floating-point contraction off,
denormals preserved and packed-FP32 instructions disabled; host code with the
same IEEE settings, so that host and device evaluate the shared numerics
header bit for bit.
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "busseq_gpu_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
OBJ = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BUSSEQ_ARCH", "gfx950")

IEEE = ["-ffp-contract=off", "-fno-fast-math"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Xarch_host", "-mfma",
    "-Wall", "-Wno-unused-function",
] + IEEE
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-mfma", "-Wall", "-Wno-unused-function"] + IEEE


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=ROOT)


def build(verbose=True):
    for d in (LIB, BIN, OBJ):
        os.makedirs(d, exist_ok=True)
    hdrs = [os.path.join(ROOT, "include", "busseq.h"),
            os.path.join(ROOT, "include", "busseq", "numerics.h"),
            os.path.join(ROOT, "include", "busseq", "model.h"),
            os.path.join(CSRC, "engine.h")]
    eng_o = os.path.join(OBJ, "engine.o")
    host_o = os.path.join(OBJ, "host.o")
    init_o = os.path.join(OBJ, "init.o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "engine.hip"), "-o", eng_o])
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "init.hip"), "-o", init_o])
    cxx = shutil.which("g++") or "g++"
    _run([cxx] + CXX_FLAGS + ["-c", os.path.join(CSRC, "host.cpp"), "-o", host_o])
    so = os.path.join(LIB, "libbusseq.so")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", so + ".tmp", eng_o, init_o, host_o, "-lgomp"])
    os.replace(so + ".tmp", so)
    cli_o = os.path.join(OBJ, "cli.o")
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-c", os.path.join(CSRC, "cli.cpp"), "-o", cli_o])
    exe = os.path.join(BIN, "BUSseq_gpu")
    _run([cxx, "-o", exe, cli_o, f"-L{LIB}", "-lbusseq", "-Wl,-rpath,$ORIGIN/../lib"])
    del hdrs
    return so


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B"])
    return os.path.join(ROOT, "oracle", "liboracle.so")


if __name__ == "__main__":
    build_oracle()
    build()
    sys.exit(0)

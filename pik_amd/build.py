"""Builds libpikgpu.so (the gfx950 HIP library) in-tree with hipcc.

Every HIP translation unit is compiled for gfx950 with
  -ffp-contract=off           operation order of the reference's scalar path
                              (no fused multiply-add unless written as fmaf)
  -target-feature -packed-fp32-ops
                              no packed FP32 instructions
and linked into pik_amd/libpikgpu.so.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpikgpu.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PIKGPU_ARCH", "gfx950")

SOURCES = ["recon_lf.hip", "recon_blocks.hip", "filter.hip", "adjoint.hip",
           "encode.hip", "search.hip", "api.hip", "noalloc.hip",
           "host_tables.cpp", "quant_tables.cpp", "entropy.cpp", "dc_groups.cpp",
           "frame_parse.cpp"]
HEADERS = ["common.h", "device.h", "kernels.h", "xform.h", "dct1d.h",
           "dct_consts.h", "host_api.h", "entropy.h"]
HOST_OUT = os.path.join(HERE, "libpikgpu_host.so")
HOST_SOURCES = ["host_check.cpp", "host_tables.cpp", "entropy.cpp",
                "dc_groups.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
          "-fvisibility=hidden", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(HERE, "..", "include")]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src + ".o")
        cmd = [HIPCC] + COMMON
        if src.endswith(".hip"):
            cmd += ["-x", "hip"] + DEVICE
        else:
            cmd += ["-x", "c++"]
        cmd += ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] +
         objs + ["-Wl,--no-undefined", "-Wl,-soname,libpikgpu.so"])
    os.replace(tmp, OUT)
    # Host build of the transforms and tables for the CPU test suite.
    tmp = HOST_OUT + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
          "-Wno-unknown-pragmas",
          "-fvisibility=hidden", "-Wall", "-o", tmp] +
         [os.path.join(CSRC, s) for s in HOST_SOURCES])
    os.replace(tmp, HOST_OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

"""Builds liblbhip.so (the HIP kernels + C ABI) in-tree for gfx950 with hipcc."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "lbhip.hip")
OUT = os.path.join(HERE, "liblbhip.so")
PROBE_SRC = os.path.join(HERE, "csrc", "probe.hip")
PROBE_OUT = os.path.join(HERE, "liblbprobe.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-munsafe-fp-atomics",
               # no packed FP32 instructions (v_pk_*_f32): plain scalar FP32 arithmetic only
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-value", "-Wno-unused-result"]


def _compile(src, out, verbose, libs=()):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + HIPCC_FLAGS + ["-I", INCLUDE, "-o", out + ".tmp", src] + list(libs)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stderr[-4000:])
    os.replace(out + ".tmp", out)
    if verbose:
        print(" ".join(cmd))


CERT_SRC = os.path.join(HERE, "csrc", "lbcert.cpp")
CERT_OUT = os.path.join(HERE, "liblbcert.so")
VERIFY_SRC = os.path.join(HERE, "csrc", "lbverify.cpp")
VERIFY_OUT = os.path.join(HERE, "liblbverify.so")
HOST_SRC = os.path.join(HERE, "csrc", "lbhost.cpp")
HOST_OUT = os.path.join(HERE, "liblbhost.so")


def _compile_host(src, out, verbose):
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", INCLUDE, "-o", out + ".tmp", src]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{cxx} failed:\n" + res.stderr[-4000:])
    os.replace(out + ".tmp", out)
    if verbose:
        print(" ".join(cmd))


def build(verbose: bool = False) -> str:
    """Compile every native source of the package: liblbhip.so (HIP kernels + C ABI), the bandwidth probe
    liblbprobe.so used by the benchmark's traffic calibration, the host part of the sparse certificate liblbcert.so,
    the independent verifier's factorization liblbverify.so and the search's host bookkeeping liblbhost.so; always
    recompiles."""
    _compile(SRC, OUT, verbose, libs=["-ldl"])
    _compile(PROBE_SRC, PROBE_OUT, verbose)
    _compile_host(CERT_SRC, CERT_OUT, verbose)
    _compile_host(VERIFY_SRC, VERIFY_OUT, verbose)
    _compile_host(HOST_SRC, HOST_OUT, verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

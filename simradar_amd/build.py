"""Build the engine library in-tree: simradar_amd/libsimradar_amd.so.

The RS_* host runtime is plain C (gcc), the device layer is HIP for gfx950
(hipcc); both are linked into one shared library whose exports are the C-ABI
declared in include/rs.h and include/rsd.h.

    python -m simradar_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsimradar_amd.so")
CLI = os.path.join(HERE, "bin", "simradar")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
ARCH = "gfx950"

# -ffp-contract=off everywhere: a*b+c stays two roundings unless written fma.
HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC,
]
C_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-pthread", "-ffp-contract=off", "-fno-fast-math",
           "-Wall", "-Wextra", "-Wno-unused-parameter", "-I" + INCLUDE]

HIP_SOURCES = ["rsd_engine.hip", "rsd_mlp.hip"]
C_SOURCES = ["rs_host.c", "rs_providers.c"]


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), out))
    if out.strip():
        print(out)


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in C_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        _run([CC] + C_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lm", "-lpthread"])
    os.replace(tmp, LIB)
    # command-line driver (simradar.c equivalent), linked against the library
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    _run([CC] + C_FLAGS + [os.path.join(CSRC, "simradar_cli.c"), "-o", CLI + ".tmp", "-L" + HERE, "-lsimradar_amd",
                           "-Wl,-rpath,$ORIGIN/..", "-lm", "-lpthread"])
    os.replace(CLI + ".tmp", CLI)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose=True):
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-s"])
    lib = os.path.join(odir, "_build", "liboracle.so")
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build()
    if "--no-oracle" not in sys.argv:
        build_oracle()

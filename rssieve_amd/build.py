"""Build the native parts of rssieve_amd in-tree.

    python -m rssieve_amd.build        (or __graft_entry__.build())

Produces, next to this file:
  librssieve.so   HIP kernels for gfx950 + the C-ABI (include/rssieve.h)
  rssieve         the command-line sieve (the reference's -p/-P/-i/-f interface)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librssieve.so")
CLI = os.path.join(HERE, "rssieve")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are disabled for every HIP translation unit
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include")]

LIB_SOURCES = ["rs_kernels.hip", "rs_mlp.hip", "rs_capi.cpp", "rs_host.cpp", "rs_io.cpp", "rs_model.cpp", "rs_peak.hip"]
CLI_SOURCES = ["rs_cli.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode:
        raise RuntimeError("command failed (%d): %s" % (r.returncode, " ".join(cmd)))


def build(verbose=True):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in LIB_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        cmd = [HIPCC, "--offload-arch=" + ARCH] + COMMON
        if src.endswith(".hip"):
            cmd += DEVICE_FLAGS + ["-x", "hip"]
        cmd += ["-c", path, "-o", obj]
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    cli_tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include")]
         + [os.path.join(CSRC, s) for s in CLI_SOURCES]
         + ["-o", cli_tmp, "-L" + HERE, "-lrssieve", "-Wl,-rpath,$ORIGIN"])
    os.replace(cli_tmp, CLI)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)

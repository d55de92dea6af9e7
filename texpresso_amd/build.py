"""Build the gfx950 library libtexpresso_hip.so in-tree with hipcc.

Flags that matter for bit-exactness with the reference encoder:
  -ffp-contract=off                     no a*b+c fusion into FMA
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                        no packed FP32 (v_pk_*_f32) code
HIP's default correctly rounded f32 division and square root are kept (no
-ffast-math, no -fno-hip-fp32-correctly-rounded-divide-sqrt).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtexpresso_hip.so")
SOURCES = ["bcn_kernels.hip", "tx_batch.hip", "tx_capi.hip", "tx_context.hip", "tx_dds.hip", "tx_mips.hip",
           "tx_pipeline.hip", "tx_png.hip", "tx_probe.hip", "tx_server.hip"]
# command-line tool over the C ABI, linked against the library with an
# $ORIGIN run path so it finds it beside itself
CLI_SOURCE = os.path.join(CSRC, "cli", "texpresso_cli.cpp")
CLI = os.path.join(HERE, "texpresso")
ARCH = "gfx950"


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def flags():
    return [
        "--offload-arch=" + ARCH,
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-I" + os.path.join(ROOT, "include"),
        "-I" + CSRC,
    ]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + flags() + ["-shared", "-o", tmp] + srcs + ["-lz"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l and "hip-link" not in l)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out)
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


def build_cli(verbose=False):
    tmp = CLI + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
           CLI_SOURCE, "-o", tmp, "-L" + HERE, "-l:libtexpresso_hip.so", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError("g++ failed:\n" + res.stdout)
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)

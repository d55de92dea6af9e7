"""Build the gfx950 library libtexpresso_hip.so in-tree with hipcc.

Flags that matter for bit-exactness with the reference encoder:
  -ffp-contract=off                     no a*b+c fusion into FMA
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                        no packed FP32 (v_pk_*_f32) code
HIP's default correctly rounded f32 division and square root are kept (no
-ffast-math, no -fno-hip-fp32-correctly-rounded-divide-sqrt).

Each source compiles to its own object in parallel (each carries its own
gfx950 code object), then one link.  Beside the library: the command-line
tool (texpresso) and the benchmark's VALU probe (libtx_bench_probe.so, which
the library does not link).
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtexpresso_hip.so")
SOURCES = ["bcn_kernels.hip", "tx_batch.hip", "tx_capi.hip", "tx_context.hip", "tx_dds.hip", "tx_mips.hip",
           "tx_pipeline.hip", "tx_png.hip", "tx_server.hip"]
# command-line tool over the C ABI, linked against the library with an
# $ORIGIN run path so it finds it beside itself
CLI_SOURCE = os.path.join(CSRC, "cli", "texpresso_cli.cpp")
CLI = os.path.join(HERE, "texpresso")
PROBE_SOURCE = os.path.join(CSRC, "bench", "valu_probe.hip")
PROBE = os.path.join(HERE, "libtx_bench_probe.so")
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


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l and "hip-link" not in l)
    return res.returncode, out


def _jobs():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    with tempfile.TemporaryDirectory(prefix="tx_build_") as tmpdir:
        objs = [os.path.join(tmpdir, os.path.basename(s) + ".o") for s in srcs]

        def compile_one(pair):
            src, obj = pair
            return src, _run([hipcc()] + flags() + ["-c", src, "-o", obj], verbose)

        with ThreadPoolExecutor(max_workers=_jobs()) as ex:
            results = list(ex.map(compile_one, zip(srcs, objs)))
        for src, (rc, out) in results:
            if rc != 0:
                raise RuntimeError("hipcc failed on %s:\n%s" % (src, out))
            if verbose and out.strip():
                print(out)
        tmp = LIB + ".tmp"
        rc, out = _run([hipcc()] + flags() + ["-shared", "-o", tmp] + objs + ["-lz"], verbose)
        if rc != 0:
            raise RuntimeError("hipcc link failed:\n" + out)
    os.replace(tmp, LIB)
    build_cli(verbose)
    build_probe(verbose)
    return LIB


def build_cli(verbose=False):
    tmp = CLI + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
           CLI_SOURCE, "-o", tmp, "-L" + HERE, "-l:libtexpresso_hip.so", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath-link,/opt/rocm/lib"]
    rc, out = _run(cmd, verbose)
    if rc != 0:
        raise RuntimeError("g++ failed:\n" + out)
    os.replace(tmp, CLI)
    return CLI


def build_probe(verbose=False):
    tmp = PROBE + ".tmp"
    rc, out = _run([hipcc()] + flags() + ["-shared", "-o", tmp, PROBE_SOURCE], verbose)
    if rc != 0:
        raise RuntimeError("hipcc failed on the VALU probe:\n" + out)
    os.replace(tmp, PROBE)
    return PROBE


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)

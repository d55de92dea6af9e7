#!/usr/bin/env python3
"""Throughput benchmark of the MI355X BC1-BC5 encoder.

Default workload (one step = one pass of the encoder over one texture):
BC3 with Algorithm::IterativeClusterFit on a seeded 2048x2048 RGBA8 texture
(RGB: 4-octave value noise + gradient + rectangles; alpha: smooth noise with
10% of blocks binarised), default perceptual weights, input and output
resident in HBM.  `value` = unpadded pixels / mean kernel time per step.

Also measured in the same run: BC5 on a 2048x2048 seeded normal map, a
byte-for-byte parity check of the full BC3 output against the CPU oracle,
and the oracle itself on the host cores as the CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BC1/BC3 ClusterFit Mpix/s on 1 MI355X (bit-exact) vs host CPU; p99 job latency"
VALU_PEAK_TFLOPS = 157.3     # MI355X FP32 vector peak (FMA = 2 flops, packed)
VALU_PLAIN_TOPS = 39.3       # one FP32 op per lane per clock: 256 CU x 64 lanes x 2.4 GHz
HBM_PEAK_GBS = 8000.0


def split_counts(n):
    """Splits searched per ordering for n distinct colours (cluster.rs loops)."""
    c4 = sum(1 for i in range(n) for j in range(i, n + 1) for k in range(max(j, 1), n + 1))
    c3 = sum(1 for i in range(n) for j in range(max(i, 1), n + 1))
    return c3, c4


C3 = np.array([split_counts(n)[0] for n in range(17)], np.float64)
C4 = np.array([split_counts(n)[1] for n in range(17)], np.float64)


def algorithmic_flops(stats):
    """Counted FP32 operations of the colour fits (per-split costs of the
    reference expression tree: 138 for a 3-colour split, 159 for a 4-colour
    split, plus ~1.4k per ClusterFit block and 0.45k per ordering; RangeFit
    ~1.5k per block)."""
    n = stats["count"].clip(0, 16)
    cl = stats["fit"] == 2
    orders = stats["searched3"] + stats["searched4"]
    f = np.where(cl, C3[n] * 138 * stats["searched3"] + C4[n] * 159 * stats["searched4"] + 1400 + 450 * orders, 0.0)
    f += np.where(stats["fit"] == 1, 1500.0, 0.0)
    return float(f.sum())


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_device(tx, fmt, d_in, w, h, params, d_out, out_len, steps, warmup):
    """Mean kernel milliseconds per step from HIP events recorded on the
    library stream around `steps` launches, plus the wall time around them."""
    if warmup:
        tx.time_compress(fmt, d_in.ptr, w, h, params, d_out.ptr, out_len, warmup)
    t0 = time.perf_counter()
    ms = tx.time_compress(fmt, d_in.ptr, w, h, params, d_out.ptr, out_len, steps)
    wall = time.perf_counter() - t0
    return ms / steps, wall


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the full parity check")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import texpresso_amd as tx
        n = args.copy_calibration << 20
        a, b = tx.DeviceBuffer(n), tx.DeviceBuffer(n)
        for _ in range(3):
            tx.lib().tx_util_copy(b.ptr, a.ptr, n)
        print(json.dumps({"copy_calibration_bytes": n, "copies": 3}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the encoder shards by blocks/textures with no collective, "
              "and multi-GPU runs are not part of this build (requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import texpresso_amd as tx
    from texpresso_amd import texgen

    if tx.lib().tx_device_arch() is None:
        raise SystemExit("no gfx950 device usable by libtexpresso_hip.so")

    W = H = args.size
    seed = texgen.SEED_BASE + 1
    img = texgen.with_alpha_field(texgen.rgb_texture(W, H, seed), seed)
    nmap = texgen.normal_map(W, H, seed)
    fmt = tx.Format.Bc3
    params = tx.Params(tx.Algorithm.IterativeClusterFit)
    out_len = fmt.compressed_size(W, H)
    d_in = tx.DeviceBuffer.from_array(img.reshape(-1))
    d_out = tx.DeviceBuffer(out_len)
    print("bench: BC3 IterativeClusterFit %dx%d, %d warmup + %d steps" % (W, H, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    ms, wall = time_device(tx, fmt, d_in, W, H, params, d_out, out_len, args.steps, args.warmup)
    mpix = W * H / 1e6
    value = mpix / (ms / 1e3)
    gpu_out = d_out.download()

    # BC5 on the normal map (second half of the workload), same protocol
    n5 = tx.Format.Bc5.compressed_size(W, H)
    d_nm = tx.DeviceBuffer.from_array(nmap.reshape(-1))
    d_nm_out = tx.DeviceBuffer(n5)
    ms5, _ = time_device(tx, tx.Format.Bc5, d_nm, W, H, tx.Params(), d_nm_out, n5, args.steps, args.warmup)
    print("bench: BC3 %.3f ms/step (%.1f Mpix/s), BC5 %.3f ms/step" % (ms, value, ms5), file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded value-noise/gradient/rectangle RGB with smooth alpha (10%% binary blocks), seed 0x%X"
                % seed,
        "config": {"workload": "bc3-iterative-clusterfit-%dx%d-rgba8" % (W, H), "format": "BC3",
                   "algorithm": "IterativeClusterFit", "width": W, "height": H, "weights": "perceptual",
                   "input_sha256": hashlib.sha256(img.tobytes()).hexdigest()[:16]},
        "wall_s": round(wall, 4),
        "secondary": {"bc5_normal_map_mpix_s": round(mpix / (ms5 / 1e3), 1), "bc5_ms_per_step": round(ms5, 4)},
    }

    import ctypes
    probe = ctypes.c_double()
    if tx.lib().tx_util_valu_probe(ctypes.byref(probe)) == 0:
        line["secondary"]["valu_plain_fp32_measured_tops"] = round(probe.value / 1e12, 2)

    if not args.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads()
        op = O.params(int(params.algorithm), params.weights, False)
        want, stats = O.compress(int(fmt), img, W, H, op, threads=threads, stats=True)
        bad = int(np.count_nonzero(np.any(gpu_out.reshape(-1, 16) != want.reshape(-1, 16), axis=1)))
        line["parity"] = {"blocks": int(want.size // 16), "mismatched_blocks": bad}
        flops = algorithmic_flops(stats)
        achieved = flops / (ms / 1e3) / 1e12
        line["roofline"] = {
            "bound": "valu", "achieved": round(achieved, 3), "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / VALU_PEAK_TFLOPS, 4), "traffic": None,
            "plain_fp32_peak_tops": VALU_PLAIN_TOPS, "frac_of_plain": round(achieved / VALU_PLAIN_TOPS, 4),
            "flops_per_launch": flops,
            "hbm_gbs": round((4 * W * H + out_len) / (ms / 1e3) / 1e9, 2), "hbm_peak_gbs": HBM_PEAK_GBS,
        }
        # CPU baseline: the oracle on the same texture over `threads` host cores
        rows = H // 4
        sample_rows = rows
        times = []
        for _ in range(args.cpu_runs):
            t0 = time.perf_counter()
            O.compress(int(fmt), img[:4 * sample_rows], W, 4 * sample_rows, op, threads=threads)
            times.append(time.perf_counter() - t0)
            if times[-1] > 30:
                break
        t_cpu = float(np.median(times))
        line["cpu_baseline"] = {
            "value": round(W * 4 * sample_rows / 1e6 / t_cpu, 3), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": "full %dx%d BC3 IterativeClusterFit texture, median of %d runs, %d threads on %s"
                      % (W, 4 * sample_rows, len(times), threads, cpu_model()),
        }
        line["gpu_over_cpu"] = round(value / line["cpu_baseline"]["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

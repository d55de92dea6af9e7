#!/usr/bin/env python3
"""Throughput benchmark of the MI355X BC1-BC5 encoder.

Default workload (one step = one pass of the encoder over one texture):
BC3 with Algorithm::IterativeClusterFit on a seeded 2048x2048 RGBA8 texture
(RGB: 4-octave value noise + gradient + rectangles; alpha: smooth noise with
10% of blocks binarised), default perceptual weights, input and output
resident in HBM.  `value` = unpadded pixels / mean kernel time per step.

Also measured in the same run: BC5 on a 2048x2048 seeded normal map, a
byte-for-byte parity check of the full BC3 output against the CPU oracle
(blocks that differ are counted, with their RGB/alpha MSE delta against the
source), and the oracle itself on the host cores as the CPU baseline.

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
VALU_PLAIN_TOPS = 78.6       # plain FP32 mul/add with vector-register operands: 1024 SIMDs x 32 lanes x 2.4 GHz
HBM_PEAK_GBS = 8000.0


def split_counts(n):
    """Splits searched per ordering for n distinct colours (cluster.rs loops)."""
    c4 = sum(1 for i in range(n) for j in range(i, n + 1) for k in range(max(j, 1), n + 1))
    c3 = sum(1 for i in range(n) for j in range(max(i, 1), n + 1))
    return c3, c4


C3 = np.array([split_counts(n)[0] for n in range(17)], np.float64)
C4 = np.array([split_counts(n)[1] for n in range(17)], np.float64)


def search_flops(stats):
    """The part of algorithmic_flops spent in the split searches (per-split
    costs plus the per-ordering work), split by pass: (3-colour, 4-colour)."""
    n = stats["count"].clip(0, 16)
    cl = stats["fit"] == 2
    f3 = np.where(cl, C3[n] * 138 * stats["searched3"] + 450 * stats["searched3"], 0.0)
    f4 = np.where(cl, C4[n] * 159 * stats["searched4"] + 450 * stats["searched4"], 0.0)
    return float(f3.sum()), float(f4.sum())


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 summary (calibrated FETCH_SIZE + WRITE_SIZE), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.md")
    try:
        with open(path) as f:
            for line in f:
                parts = [x.strip() for x in line.strip().strip("|").split("|")]
                if len(parts) >= 5 and parts[0] == kernel and parts[1] == workload:
                    return int(float(parts[4]))
    except (OSError, ValueError):
        pass
    return None


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


def cpu_counts():
    """(threads used, CPUs online, CPUs in this process's affinity mask).

    The CPU baseline runs on every CPU this process may use, unless
    OMP_NUM_THREADS names a smaller share (a host that runs one job per GPU
    sets it to the cores that belong to each GPU's job)."""
    online = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = online
    n = affinity
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        share = 0
    if share > 0:
        n = min(n, share)
    return max(1, n), online, affinity


def cpu_threads():
    return cpu_counts()[0]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def valu_probe(waves_per_simd):
    """Plain FP32 lane-operations per second measured by the benchmark's
    probe library (texpresso_amd/libtx_bench_probe.so, built beside the
    encoder library and not part of it); None when it is absent or fails."""
    import ctypes
    path = os.path.join(ROOT, "texpresso_amd", "libtx_bench_probe.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    L.txb_valu_probe.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.txb_valu_probe.restype = ctypes.c_int
    rate = ctypes.c_double()
    return rate.value if L.txb_valu_probe(waves_per_simd, ctypes.byref(rate)) == 0 else None


def op_rates():
    """Rows {mix, wave_instrs_per_us_per_simd, lanes_per_clk_at_2400mhz} from
    the probe library's instruction-mix kernels (8 waves per SIMD)."""
    import ctypes
    path = os.path.join(ROOT, "texpresso_amd", "libtx_bench_probe.so")
    if not os.path.exists(path):
        sys.exit("libtx_bench_probe.so is not built")
    L = ctypes.CDLL(path)
    L.txb_op_mix_name.restype = ctypes.c_char_p
    L.txb_op_mix_name.argtypes = [ctypes.c_int]
    L.txb_op_rate.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    rows = []
    for m in range(L.txb_op_mixes()):
        rate = ctypes.c_double()
        if L.txb_op_rate(m, ctypes.byref(rate)) != 0:
            sys.exit("op-rate probe failed on mix %d" % m)
        rows.append({"mix": L.txb_op_mix_name(m).decode(), "wave_instrs_per_us_per_simd": round(rate.value / 1e6, 1),
                     "lanes_per_clk_at_2400mhz": round(rate.value * 64 / 2.4e9, 2)})
    return rows


def parity_report(O, fmt, img, w, h, got, want):
    """Blocks whose bytes differ from the oracle's, and over those blocks the
    RGB and alpha mean squared error against the source pixels of the GPU
    block's decode minus that of the oracle block's decode (decoded with the
    oracle's decoder; 0 when no block differs)."""
    bs = 8 if fmt in (0, 3) else 16
    g = np.asarray(got, np.uint8).reshape(-1, bs)
    r = np.asarray(want, np.uint8).reshape(-1, bs)
    diff = np.flatnonzero(np.any(g != r, axis=1))
    rep = {"blocks": int(r.shape[0]), "mismatched_blocks": int(diff.size), "rgb_mse_delta": 0.0,
           "alpha_mse_delta": 0.0}
    if diff.size:
        src = np.asarray(img, np.uint8).reshape(h, w, 4).astype(np.float64)
        bw = (w + 3) // 4
        d_rgb, d_a = [], []
        for b in diff:
            bx, by = int(b) % bw, int(b) // bw
            px = [(4 * by + y, 4 * bx + x, 4 * y + x) for y in range(4) for x in range(4)
                  if 4 * by + y < h and 4 * bx + x < w]
            if not px:
                continue
            dg = O.decompress_block(fmt, g[b]).reshape(16, 4).astype(np.float64)
            dr = O.decompress_block(fmt, r[b]).reshape(16, 4).astype(np.float64)
            s = np.array([src[yy, xx] for yy, xx, _ in px])
            ig = dg[[i for _, _, i in px]]
            ir = dr[[i for _, _, i in px]]
            d_rgb.append(np.mean((ig[:, :3] - s[:, :3]) ** 2) - np.mean((ir[:, :3] - s[:, :3]) ** 2))
            d_a.append(np.mean((ig[:, 3] - s[:, 3]) ** 2) - np.mean((ir[:, 3] - s[:, 3]) ** 2))
        rep["rgb_mse_delta"] = round(float(np.mean(d_rgb)), 6) if d_rgb else 0.0
        rep["alpha_mse_delta"] = round(float(np.mean(d_a)), 6) if d_a else 0.0
    return rep


def time_device(tx, fmt, d_in, w, h, params, d_out, out_len, steps, warmup):
    """Mean kernel milliseconds per step from HIP events recorded on the
    library stream around `steps` launches, plus the wall time around them."""
    if warmup:
        tx.time_compress(fmt, d_in.ptr, w, h, params, d_out.ptr, out_len, warmup)
    t0 = time.perf_counter()
    ms = tx.time_compress(fmt, d_in.ptr, w, h, params, d_out.ptr, out_len, steps)
    wall = time.perf_counter() - t0
    return ms / steps, wall


def _upload_set(tx, items):
    """items: (format, params, image) -> device buffers and batch descriptors."""
    bufs, descs = [], []
    for fmt, p, img in items:
        h, w = img.shape[:2]
        d_in = tx.DeviceBuffer.from_array(img.reshape(-1))
        n = tx.Format(fmt).compressed_size(w, h)
        d_out = tx.DeviceBuffer(max(n, 16))
        bufs.append((d_in, d_out, n))
        descs.append((fmt, p, d_in.ptr, d_out.ptr, w, h))
    return bufs, descs


def _download(bufs):
    return [d_out.download(n) for _, d_out, n in bufs]


def bc1_texture(texgen, size):
    """The seeded size x size RGB texture of the BC1 cases (alpha 255)."""
    return texgen.rgb_texture(size, size, texgen.SEED_BASE)


def bc1_cluster_gpu(tx, texgen, size, steps, warmup, img=None):
    """BC1 ClusterFit on the seeded size x size RGB texture (alpha 255,
    default perceptual weights), device-resident: the timed GPU half of a
    bc1 case (512 is the CPU-runnable configuration)."""
    img = bc1_texture(texgen, size) if img is None else img
    p = tx.Params()
    d_in = tx.DeviceBuffer.from_array(img.reshape(-1))
    n = tx.Format.Bc1.compressed_size(size, size)
    d_out = tx.DeviceBuffer(n)
    ms, _ = time_device(tx, tx.Format.Bc1, d_in, size, size, p, d_out, n, steps, warmup)
    got = d_out.download()
    d_in.free()
    d_out.free()
    return {"size": size, "img": img, "got": got, "ms": ms, "weights": p.weights}


def bc1_cluster_cpu(O, g, threads):
    """The CPU half of a bc1 case: byte parity with the oracle and the
    oracle's own rate on the host CPUs."""
    W = H = g["size"]
    op = O.params(1, g["weights"], False)
    t0 = time.perf_counter()
    want, stats = O.compress(0, g["img"], W, H, op, threads=threads, stats=True)
    t_cpu = time.perf_counter() - t0
    par = parity_report(O, 0, g["img"], W, H, g["got"], want)
    flops = algorithmic_flops(stats)
    ms = g["ms"]
    gpu = W * H / 1e6 / (ms / 1e3)
    cpu = W * H / 1e6 / t_cpu
    return {"config": "bc1-%d: BC1 ClusterFit %dx%d" % (W, W, H), "gpu_mpix_s": round(gpu, 1),
            "ms": round(ms, 4), "cpu_mpix_s": round(cpu, 3), "cpu_threads": threads,
            "gpu_over_cpu": round(gpu / cpu, 1), "mismatched_blocks": par["mismatched_blocks"],
            "rgb_mse_delta": par["rgb_mse_delta"], "alpha_mse_delta": par["alpha_mse_delta"], "blocks": W * H // 16,
            "input_sha256": hashlib.sha256(g["img"].tobytes()).hexdigest()[:16],
            "valu_tflops": round(flops / (ms / 1e3) / 1e12, 2)}


def bc1_cluster_case(tx, texgen, O, size, threads, steps, warmup):
    return bc1_cluster_cpu(O, bc1_cluster_gpu(tx, texgen, size, steps, warmup), threads)


def suite_bc1_512(tx, texgen, O, threads, steps, warmup):
    return bc1_cluster_case(tx, texgen, O, 512, threads, steps, warmup)


def chain_4096(texgen):
    """Full mip chain (13 levels down to 1x1) of the seeded 4096x4096 texture."""
    seed = texgen.SEED_BASE + 2
    base = texgen.with_alpha_field(texgen.rgb_texture(4096, 4096, seed), seed)
    return texgen.mip_chain(base)


def chain_gpu(tx, chain, name, fmt, alg, steps, warmup, tuning_path, retune):
    """One format/algorithm over the whole chain, device-resident: the chain
    as one batch (the tuning the autotuner table holds for it) and as one
    plain launch per level; the timed GPU half of a chain case."""
    p = tx.Params(alg)
    items = [(fmt, p, l) for l in chain]
    bufs, descs = _upload_set(tx, items)
    best, _ = tx.tuned("chain4096-%s-%s" % (name, alg.name), descs, table_path=tuning_path, retune=retune)
    batch = tx.Batch(descs, best)
    batch.run_timed(max(1, warmup))
    ms_batch = batch.run_timed(steps) / steps
    got = _download(bufs)
    batch.close()
    tx.time_per_texture(descs, max(2, warmup // 4))
    ms_plain = tx.time_per_texture(descs, steps) / steps
    got_plain = _download(bufs)
    # the same per-level launches with the binned pipeline off: every level
    # through the one-pass lockstep kernel (the plainest launch there is)
    old = tx.set_pipeline_threshold(2 ** 64 - 1)
    try:
        tx.time_per_texture(descs, 2)
        ms_onepass = tx.time_per_texture(descs, max(3, steps // 4)) / max(3, steps // 4)
        got_onepass = _download(bufs)
    finally:
        tx.set_pipeline_threshold(old)
    for d_in, d_out, _ in bufs:
        d_in.free()
        d_out.free()
    return {"name": name, "fmt": fmt, "alg": alg, "weights": p.weights, "got": got, "got_plain": got_plain,
            "got_onepass": got_onepass, "ms_onepass": ms_onepass,
            "ms_batch": ms_batch, "ms_plain": ms_plain, "tuning": best.as_dict(),
            "block_size": tx.Format(fmt).block_size()}


def chain_cpu(O, chain, g, threads):
    """The CPU half of a chain case: every level of both GPU outputs
    compared with the oracle, which is also timed on the host CPUs over the
    same levels."""
    px = sum(l.shape[0] * l.shape[1] for l in chain)
    op = O.params(int(g["alg"]), g["weights"], False)
    t0 = time.perf_counter()
    want, flops = [], 0.0
    for l in chain:
        w_, st = O.compress(g["fmt"], l, l.shape[1], l.shape[0], op, threads=threads, stats=True)
        want.append(w_)
        flops += algorithmic_flops(st)
    t_cpu = time.perf_counter() - t0
    bs = g["block_size"]

    def bad_blocks(outs):
        return sum(int(np.count_nonzero(np.any(o.reshape(-1, bs) != w.reshape(-1, bs), axis=1)))
                   for o, w in zip(outs, want))

    def mse_deltas(outs):
        reps = [parity_report(O, g["fmt"], l, l.shape[1], l.shape[0], o, w) for l, o, w in zip(chain, outs, want)
                if np.any(o != w)]
        n = sum(r["mismatched_blocks"] for r in reps)
        if not n:
            return 0.0, 0.0
        return (round(sum(r["rgb_mse_delta"] * r["mismatched_blocks"] for r in reps) / n, 6),
                round(sum(r["alpha_mse_delta"] * r["mismatched_blocks"] for r in reps) / n, 6))

    traffic = 4 * px + sum(len(w) for w in want)
    ms_batch, ms_plain = g["ms_batch"], g["ms_plain"]
    return {"config": "chains: %s %s 4096 chain (%d levels)" % (g["name"], g["alg"].name, len(chain)),
            "mpix": round(px / 1e6, 3), "blocks": int(sum(len(w) // bs for w in want)),
            "batch_ms": round(ms_batch, 4), "batch_mpix_s": round(px / 1e6 / (ms_batch / 1e3), 1),
            "per_level_ms": round(ms_plain, 4), "per_level_mpix_s": round(px / 1e6 / (ms_plain / 1e3), 1),
            "batch_over_per_level": round(ms_plain / ms_batch, 3),
            "per_level_onepass_ms": round(g["ms_onepass"], 4),
            "batch_over_per_level_onepass": round(g["ms_onepass"] / ms_batch, 3),
            "cpu_mpix_s": round(px / 1e6 / t_cpu, 3), "cpu_threads": threads,
            "gpu_over_cpu": round((px / ms_batch / 1e3) / (px / t_cpu / 1e6), 1),
            "mismatched_blocks": bad_blocks(g["got"]), "mismatched_blocks_per_level": bad_blocks(g["got_plain"]),
            "mismatched_blocks_per_level_onepass": bad_blocks(g["got_onepass"]),
            "rgb_alpha_mse_delta_batched": mse_deltas(g["got"]),
            "tuning": g["tuning"],
            "valu_tflops": round(flops / (ms_batch / 1e3) / 1e12, 3),
            "valu_frac_of_157": round(flops / (ms_batch / 1e3) / 1e12 / VALU_PEAK_TFLOPS, 4),
            "hbm_gbs": round(traffic / (ms_batch / 1e3) / 1e9, 1),
            "hbm_frac": round(traffic / (ms_batch / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}


def chain_case(tx, O, chain, name, fmt, alg, threads, steps, warmup, tuning_path, retune):
    return chain_cpu(O, chain, chain_gpu(tx, chain, name, fmt, alg, steps, warmup, tuning_path, retune), threads)


CHAIN_CASES = [("BC1", 0, "RangeFit"), ("BC1", 0, "ClusterFit"), ("BC2", 1, "ClusterFit"),
               ("BC3", 2, "IterativeClusterFit"), ("BC4", 3, "ClusterFit"), ("BC1", 0, "IterativeClusterFit")]


def suite_chains(tx, texgen, O, threads, steps, warmup, tuning_path, retune=True):
    """Full mip chain of a seeded 4096x4096 texture, per format;
    the chain is one batch with the autotuned tuning for that format
    (re-tuned first unless `retune` is false)."""
    chain = chain_4096(texgen)
    rows = []
    for name, fmt, alg in CHAIN_CASES:
        rows.append(chain_case(tx, O, chain, name, fmt, tx.Algorithm[alg], threads, steps, warmup, tuning_path,
                               retune=retune))
        print("suite: " + json.dumps(rows[-1]), file=sys.stderr, flush=True)
    return rows


def batch_items(tx, texgen, count=1000):
    """`count` textures, sides log-uniform in [3, 2048], full mip
    chains, formats uniform over BC1/BC3/BC4/BC5, Params::default()."""
    rng = np.random.Generator(np.random.PCG64(texgen.SEED_BASE + 3))
    ws = texgen.log_uniform_sides(rng, count, 3, 2048)
    hs = texgen.log_uniform_sides(rng, count, 3, 2048)
    fmts = rng.choice([0, 2, 3, 4], size=count)
    items = []
    for i in range(count):
        img = texgen.with_alpha_field(texgen.rgb_texture(int(ws[i]), int(hs[i]), texgen.SEED_BASE + 1000 + i),
                                      texgen.SEED_BASE + 1000 + i)
        for lv in texgen.mip_chain(img):
            items.append((int(fmts[i]), tx.Params(), lv))
    return items


def suite_batch(tx, texgen, O, threads, steps, warmup, tuning_path, count=1000):
    items = batch_items(tx, texgen, count)
    px = sum(im.shape[0] * im.shape[1] for _, _, im in items)
    bufs, descs = _upload_set(tx, items)
    sig = "batch-%d" % count
    best, tune_ms = tx.tuned(sig, descs, table_path=tuning_path, retune=True)
    batch = tx.Batch(descs, best)
    batch.run_timed(max(1, warmup))
    ms_batch = batch.run_timed(steps) / steps
    got = _download(bufs)
    default = tx.Batch(descs)
    default.run_timed(max(1, warmup))
    ms_default = default.run_timed(steps) / steps
    tx.time_per_texture(descs, 2)
    ms_plain = tx.time_per_texture(descs, steps) / steps
    plain = _download(bufs)
    mismatch_vs_plain = sum(int(np.count_nonzero(a != b)) for a, b in zip(got, plain))
    # oracle check on a seeded sample of textures
    rng = np.random.default_rng(7)
    sample = rng.choice(len(items), size=min(200, len(items)), replace=False)
    bad = 0
    for i in sample:
        fmt, p, im = items[i]
        want = O.compress(fmt, im, im.shape[1], im.shape[0], O.params(int(p.algorithm), p.weights, False),
                          threads=threads)
        bad += int(np.count_nonzero(got[i] != want))
    return {"config": "batch: %d textures + mip chains (%d levels), BC1/BC3/BC4/BC5 default params" % (count, len(items)),
            "mpix": round(px / 1e6, 2), "batch_ms": round(ms_batch, 3),
            "batch_mpix_s": round(px / 1e6 / (ms_batch / 1e3), 1), "batch_default_tuning_ms": round(ms_default, 3),
            "per_texture_ms": round(ms_plain, 3), "per_texture_mpix_s": round(px / 1e6 / (ms_plain / 1e3), 1),
            "batch_speedup_vs_per_texture": round(ms_plain / ms_batch, 2), "tuning": best.as_dict(),
            "autotune_best_ms": round(tune_ms, 3), "mismatched_bytes_vs_per_texture": mismatch_vs_plain,
            "oracle_sample_textures": int(len(sample)), "oracle_sample_mismatched_bytes": bad}


def serving_pool(tx, texgen, count=24):
    """Serving job inputs: seeded textures, side log-uniform in [64, 4096],
    format uniform over BC1/BC3/BC5, IterativeClusterFit."""
    rng = np.random.Generator(np.random.PCG64(texgen.SEED_BASE + 4))
    sides = texgen.log_uniform_sides(rng, count, 64, 4096)
    fmts = rng.choice([0, 2, 4], size=count)
    pool = []
    for i in range(count):
        s = int(sides[i])
        img = texgen.with_alpha_field(texgen.rgb_texture(s, s, texgen.SEED_BASE + 4000 + i), texgen.SEED_BASE + 4000 + i)
        img = texgen.normal_map(s, s, texgen.SEED_BASE + 4000 + i) if fmts[i] == 4 else img
        pool.append((int(fmts[i]), np.ascontiguousarray(img.reshape(-1)), s))
    return pool


def _pcts(lat_ms):
    a = np.sort(np.asarray(lat_ms))
    return {"p50_ms": round(float(np.percentile(a, 50)), 3), "p99_ms": round(float(np.percentile(a, 99)), 3),
            "mean_ms": round(float(a.mean()), 3)}


def serve_run(tx, pool, arrivals, picks, params, server_kw):
    """Jobs arrive at `arrivals` (s, relative); returns per-job latencies and
    the wall time from first arrival to last completion."""
    import threading
    srv = tx.Server(**server_kw)
    n = len(arrivals)
    ids = [None] * n
    t_arr = [0] * n
    ready = [threading.Event() for _ in range(n)]
    done_dev, done_host, launch_jobs = [0] * n, [0] * n, [0] * n
    phases = [None] * n
    t0 = time.monotonic_ns() + 20_000_000

    def submitter():
        for k in range(n):
            target = t0 + int(arrivals[k] * 1e9)
            while True:
                d = target - time.monotonic_ns()
                if d <= 0:
                    break
                time.sleep(min(d / 1e9, 0.001) if d > 200_000 else 0)
            fmt, img, s = pool[picks[k]]
            t_arr[k] = time.monotonic_ns()
            while True:
                try:
                    ids[k] = srv.submit(fmt, img, s, s, params)
                    break
                except tx.TexpressoError as e:
                    if e.code != -9:
                        raise
                    time.sleep(0.0002)
            ready[k].set()

    def collector():
        for k in range(n):
            ready[k].wait()
            info = srv.wait(ids[k])
            fmt, img, s = pool[picks[k]]
            srv.fetch(ids[k], tx.Format(fmt).compressed_size(s, s))
            done_host[k] = time.monotonic_ns()
            done_dev[k] = info.t_done_ns
            launch_jobs[k] = info.launch_jobs
            phases[k] = ((info.t_admit_ns - info.t_submit_ns) / 1e6, (info.t_launch_ns - info.t_admit_ns) / 1e6,
                         (info.t_done_ns - info.t_launch_ns) / 1e6)
            srv.release(ids[k])

    ts = threading.Thread(target=submitter)
    tc = threading.Thread(target=collector)
    ts.start()
    tc.start()
    ts.join()
    tc.join()
    srv.close()
    lat_dev = [(done_dev[k] - t_arr[k]) / 1e6 for k in range(n)]
    lat_host = [(done_host[k] - t_arr[k]) / 1e6 for k in range(n)]
    wall = (max(done_host) - min(t_arr)) / 1e9
    ph = np.array(phases)
    breakdown = {"queue_p50_ms": round(float(np.median(ph[:, 0])), 3),
                 "admit_to_launch_p50_ms": round(float(np.median(ph[:, 1])), 3),
                 "launch_to_done_p50_ms": round(float(np.median(ph[:, 2])), 3)}
    return lat_dev, lat_host, wall, float(np.mean(launch_jobs)), breakdown


def sequential_run(tx, pool, arrivals, picks, params):
    """The plain path: one tx_compress (upload, encode, download) per job,
    in arrival order, one at a time."""
    n = len(arrivals)
    outs = {}
    lat = []
    t0 = time.monotonic_ns() + 20_000_000
    first = None
    last = 0
    for k in range(n):
        target = t0 + int(arrivals[k] * 1e9)
        now = time.monotonic_ns()
        if now < target:
            time.sleep((target - now) / 1e9)
        fmt, img, s = pool[picks[k]]
        f = tx.Format(fmt)
        if fmt not in outs or outs[fmt][1] != s:
            outs[fmt] = (np.zeros(f.compressed_size(s, s), np.uint8), s)
        f.compress(img, s, s, params, outs[fmt][0])
        last = time.monotonic_ns()
        first = target if first is None else first
        lat.append((last - target) / 1e6)
    return lat, (last - first) / 1e9


def suite_serving(tx, texgen, rates, pool_size=24, seconds=2.0):
    pool = serving_pool(tx, texgen, pool_size)
    params = tx.Params(tx.Algorithm.IterativeClusterFit)
    # untimed warm-up: every pool texture once through a server and once
    # sequentially (kernel loading, scratch growth)
    warm = np.zeros(len(pool))
    serve_run(tx, pool, warm, np.arange(len(pool)), params,
              dict(hbm_budget=16 << 30, max_inflight=3, max_jobs_per_launch=64))
    sequential_run(tx, pool, warm, np.arange(len(pool)), params)
    rows = []
    for rate in rates:
        n = int(max(300, min(2000, rate * seconds)))
        rng = np.random.default_rng(int(rate))
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
        picks = rng.integers(0, len(pool), size=n)
        mpix = float(sum(pool[p][2] ** 2 for p in picks)) / 1e6
        lat_dev, lat_host, wall, avg_batch, breakdown = serve_run(
            tx, pool, arrivals, picks, params,
            dict(hbm_budget=16 << 30, max_inflight=3, max_jobs_per_launch=64))
        lat_seq, wall_seq = sequential_run(tx, pool, arrivals, picks, params)
        row = {"config": "serving: Poisson %g jobs/s, %d jobs, BC1/BC3/BC5 IterativeClusterFit, sides 64-4096" %
                         (rate, n),
               "offered_mpix_s": round(mpix / arrivals[-1], 1),
               "server_mpix_s": round(mpix / wall, 1), "server_device_latency": _pcts(lat_dev),
               "server_latency_incl_fetch": _pcts(lat_host), "avg_jobs_per_launch": round(avg_batch, 2),
               "server_phases": breakdown,
               "sequential_mpix_s": round(mpix / wall_seq, 1), "sequential_latency": _pcts(lat_seq)}
        rows.append(row)
        print("suite: " + json.dumps(row), file=sys.stderr, flush=True)
    return rows


def run_suite(args):
    import texpresso_amd as tx
    from oracle import oracle as O
    from texpresso_amd import texgen
    threads = cpu_threads()
    which = args.suite.split(",")
    out = []
    if "bc1-512" in which:
        out.append(suite_bc1_512(tx, texgen, O, threads, args.steps, args.warmup))
        print(json.dumps(out[-1]), flush=True)
    if "bc1-2048" in which:
        out.append(bc1_cluster_case(tx, texgen, O, 2048, threads, args.steps, args.warmup))
        print(json.dumps(out[-1]), flush=True)
    if "chains" in which:
        for r in suite_chains(tx, texgen, O, threads, args.steps, args.warmup, args.tuning_table,
                              retune=not args.keep_tuning):
            out.append(r)
            print(json.dumps(r), flush=True)
    if "batch" in which:
        out.append(suite_batch(tx, texgen, O, threads, args.steps, args.warmup, args.tuning_table, args.batch_count))
        print(json.dumps(out[-1]), flush=True)
    if "serving" in which:
        for r in suite_serving(tx, texgen, [float(x) for x in args.qps.split(",")]):
            out.append(r)
            print(json.dumps(r), flush=True)
    if args.suite_out:
        with open(args.suite_out, "w") as f:
            json.dump(out, f, indent=1)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # the GPU's clocks settle over the first ~30 launches (1.33 -> 1.11 ms per
    # launch of the split search on the default workload); time the steady state
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the full parity check")
    ap.add_argument("--op-rates", action="store_true",
                    help="issue rate of single instruction forms and two-form mixes (benchmark probe library)")
    ap.add_argument("--valu-sweep", action="store_true",
                    help="print the plain FP32 VALU issue rate at 1..8 waves per SIMD and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (calibrates HBM byte counters) and exit")
    ap.add_argument("--suite", default="", help="comma list of bc1-512,bc1-2048,chains,batch,serving: per-workload lines instead of the "
                                                 "default line")
    ap.add_argument("--suite-out", default="", help="write the suite results to this JSON file")
    ap.add_argument("--tuning-table", default=os.path.join(ROOT, "texpresso_amd", "tuning.json"))
    ap.add_argument("--keep-tuning", action="store_true",
                    help="chains suite: use the tunings the table holds instead of re-running the autotuner")
    ap.add_argument("--batch-count", type=int, default=1000)
    ap.add_argument("--qps", default="100,250,500,1000,2000", help="serving arrival-rate sweep (jobs/s)")
    ap.add_argument("--pipeline-threshold", type=int, default=-1,
                    help="blocks from which one texture goes through the binned ClusterFit pipeline "
                         "(-1: library default; 0: always; a huge value: never)")
    args = ap.parse_args()
    if args.pipeline_threshold >= 0:
        import texpresso_amd as tx
        tx.set_pipeline_threshold(args.pipeline_threshold)
    if args.suite:
        if args.gpus != 1:
            sys.exit(2)
        run_suite(args)
        return
    if args.op_rates:
        for row in op_rates():
            print(json.dumps(row), flush=True)
        return
    if args.valu_sweep:
        for w in range(1, 9):
            rate = valu_probe(w)
            if rate is None:
                sys.exit("valu probe failed")
            print(json.dumps({"waves_per_simd": w, "plain_fp32_tops": round(rate / 1e12, 2)}), flush=True)
        return
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

    # inputs first (generated on the CPU and uploaded), then the GPU phases
    # back to back, then the CPU phases (oracle parity, CPU baselines): the
    # GPU's clocks drop while it idles (the split search runs ~1.3 ms per
    # launch for its first launches after an idle stretch, ~1.11 ms once
    # they settle), so no CPU work sits between the timed GPU phases
    W = H = args.size
    seed = texgen.SEED_BASE + 1
    img = texgen.with_alpha_field(texgen.rgb_texture(W, H, seed), seed)
    nmap = texgen.normal_map(W, H, seed)
    fmt = tx.Format.Bc3
    params = tx.Params(tx.Algorithm.IterativeClusterFit)
    out_len = fmt.compressed_size(W, H)
    d_in = tx.DeviceBuffer.from_array(img.reshape(-1))
    d_out = tx.DeviceBuffer(out_len)
    n5 = tx.Format.Bc5.compressed_size(W, H)
    d_nm = tx.DeviceBuffer.from_array(nmap.reshape(-1))
    d_nm_out = tx.DeviceBuffer(n5)
    pending = []
    chain = None
    if not args.no_cpu:
        # secondaries: BC1 ClusterFit (the metric's other format) at 512^2 and
        # 2048^2, and the 4096^2 mip chain in BC1 and BC3
        # IterativeClusterFit, batched and per level
        bc1_imgs = {size: bc1_texture(texgen, size) for size in (512, 2048)}
        chain = chain_4096(texgen)
        for size in (512, 2048):
            pending.append(("bc1_clusterfit_%d" % size, "bc1", bc1_cluster_gpu(tx, texgen, size, args.steps,
                                                                             args.warmup, bc1_imgs[size])))
            print("bench: BC1 ClusterFit %d^2 %.4f ms/step" % (size, pending[-1][2]["ms"]), file=sys.stderr,
                  flush=True)
        for name, f, alg in (("BC1", 0, "IterativeClusterFit"), ("BC3", 2, "IterativeClusterFit")):
            g = chain_gpu(tx, chain, name, f, tx.Algorithm[alg], args.steps, args.warmup, args.tuning_table,
                          retune=False)
            pending.append(("chain4096_%s_%s" % (name.lower(), alg.lower()), "chain", g))
            print("bench: 4096^2 chain %s %s batched %.3f ms, per level %.3f ms" %
                  (name, alg, g["ms_batch"], g["ms_plain"]), file=sys.stderr, flush=True)

    # BC5 on the normal map (second half of the workload), same protocol
    ms5, _ = time_device(tx, tx.Format.Bc5, d_nm, W, H, tx.Params(), d_nm_out, n5, args.steps, args.warmup)
    print("bench: BC3 IterativeClusterFit %dx%d, %d warmup + %d steps" % (W, H, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    ms, wall = time_device(tx, fmt, d_in, W, H, params, d_out, out_len, args.steps, args.warmup)
    mpix = W * H / 1e6
    value = mpix / (ms / 1e3)
    gpu_out = d_out.download()
    # per-kernel split of the same launches (events between the pipeline's kernels)
    phases = tx.time_phases(fmt, d_in.ptr, W, H, params, d_out.ptr, out_len, args.steps)
    # job latency: every job (this texture, input resident in HBM) timed on
    # its own, events around its launches and a wait for it to finish
    lat_dev = [tx.time_compress(fmt, d_in.ptr, W, H, params, d_out.ptr, out_len, 1) for _ in range(max(50, args.steps))]

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
    if phases is not None:
        line["secondary"]["pipeline_ms_per_step"] = {k: round(v / args.steps, 4) for k, v in phases.items()}
    # host buffers in and out (tx_compress: PCIe upload, encode, download), wall clock
    host_out = np.zeros(out_len, np.uint8)
    flat = np.ascontiguousarray(img.reshape(-1))
    fmt.compress(flat, W, H, params, host_out)
    t_host = []
    for _ in range(20):
        t0 = time.perf_counter()
        fmt.compress(flat, W, H, params, host_out)
        t_host.append(time.perf_counter() - t0)
    line["secondary"]["host_buffers_ms"] = round(1e3 * float(np.median(t_host)), 4)
    line["secondary"]["host_buffers_mpix_s"] = round(mpix / float(np.median(t_host)), 1)
    line["latency"] = {"job": "one %dx%d BC3 IterativeClusterFit texture" % (W, H),
                       "device_resident": dict(_pcts(lat_dev), max_ms=round(float(max(lat_dev)), 3),
                                               samples=len(lat_dev)),
                       "host_buffers_incl_pcie": dict(_pcts([1e3 * t for t in t_host]), samples=len(t_host))}

    probe = valu_probe(8)
    if probe is not None:
        line["secondary"]["valu_plain_fp32_measured_tops"] = round(probe / 1e12, 2)

    if not args.no_cpu:
        from oracle import oracle as O
        threads, online, affinity = cpu_counts()
        op = O.params(int(params.algorithm), params.weights, False)
        want, stats = O.compress(int(fmt), img, W, H, op, threads=threads, stats=True)
        line["parity"] = parity_report(O, int(fmt), img, W, H, gpu_out, want)
        flops = algorithmic_flops(stats)
        workload = line["config"]["workload"]
        if phases is not None:
            # dominant kernel: the 4-colour split search of the binned pipeline
            f3, f4 = search_flops(stats)
            kernel, k_ms, k_flops = "k_pipe_search<4>", phases["search4"] / args.steps, f4
        else:
            kernel, k_ms, k_flops = "k_encode<BC3,IterativeClusterFit>", ms, flops
        achieved = k_flops / (k_ms / 1e3) / 1e12
        line["roofline"] = {
            "bound": "valu", "kernel": kernel, "achieved": round(achieved, 3), "peak": VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / VALU_PEAK_TFLOPS, 4),
            "traffic": profiled_traffic(kernel, workload), "kernel_ms": round(k_ms, 4),
            "algorithmic_flops_per_launch": k_flops, "path_flops_per_step": flops,
            "path_tflops": round(flops / (ms / 1e3) / 1e12, 3),
            "plain_fp32_peak_tops": VALU_PLAIN_TOPS,
            "algorithmic_bytes_per_step": 4 * W * H + out_len,
            "hbm_gbs": round((4 * W * H + out_len) / (ms / 1e3) / 1e9, 2), "hbm_peak_gbs": HBM_PEAK_GBS,
        }
        # CPU baseline: the oracle on the same texture over `threads` host
        # cores, median of --cpu-runs timed runs (the parity run above is
        # their warm-up)
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
        # one core on a bounded band of the same texture (the first 256 rows)
        band = min(H, 256)
        t0 = time.perf_counter()
        O.compress(int(fmt), img[:band], W, band, op, threads=1)
        t_one = time.perf_counter() - t0
        line["cpu_baseline"] = {
            "value": round(W * 4 * sample_rows / 1e6 / t_cpu, 3), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": "full %dx%d BC3 IterativeClusterFit texture, median of %d runs, %d threads on %s"
                      % (W, 4 * sample_rows, len(times), threads, cpu_model()),
            "cpu_model": cpu_model(), "cpus_online": online, "cpus_in_affinity": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "single_core_mpix_s": round(W * band / 1e6 / t_one, 3),
            "single_core_sample": "rows 0-%d of the same texture, 1 thread" % (band - 1),
        }
        line["gpu_over_cpu"] = round(value / line["cpu_baseline"]["value"], 1)

        # CPU halves of the secondaries
        sec = line["secondary"]
        for key, kind, g in pending:
            row = bc1_cluster_cpu(O, g, threads) if kind == "bc1" else chain_cpu(O, chain, g, threads)
            sec[key] = {k: v for k, v in row.items() if k != "config"}
            print("bench: %s %s" % (key, json.dumps(sec[key])), file=sys.stderr, flush=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

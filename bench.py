#!/usr/bin/env python3
"""Walk-throughput benchmark on one MI355X.

Workload (the DeepWalk configuration the headline metric is quoted on):
DeepWalk with static alias sampling on a seeded Chung-Lu power-law graph,
2^22 vertices, exponent 2.1, minimum degree 2, maximum degree 65,536,
average degree 16 (8 * 2^22 undirected edges), symmetrised, int32 CSR with
FP64 edge weights i.i.d. U[1,5) mirrored on both directions; walk length 80,
one walker per non-isolated vertex (generate_all_source).  One bench "step"
is one launch of the fused walk kernel over every walker (79 moves each).

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  value = walk steps (moves) per second with the graph,
tables, sources and output buffers resident in HBM; timed with HIP events on
the stream the kernel runs on.  roofline.achieved = algorithmic bytes per
launch (steps x 132 B: two random 64-byte sectors plus the 4-byte path entry)
over the average launch time.  cpu_baseline runs the reference's own DeepWalk
binary (oracle/_ref/deepwalk.out, built from the reference sources) on all
host cores on the same graph.
"""
import argparse
import ctypes
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import thunderrw_amd as T  # noqa: E402

METRIC = "walk steps/sec + % of HBM gather roofline; batched vs plain; PPR p99 latency"
HBM_PEAK_GBS = 8000.0
SECTOR = 64
# Algorithmic bytes per step: distinct 64-B sectors of data-dependent reads + streamed bytes.
B_STEP = {
    "alias": SECTOR * 2 + 4,          # offset pair + AliasSlot sectors, 4-B path write
    "alias_replay": SECTOR * 2 + 12,  # + two 4-B replay draws
    "uniform": SECTOR * 2 + 4,        # offset pair + adj sectors, path write
}
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "deepwalk.out")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_walk_fused_alias_deepwalk_2p22.json")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22, help="log2 of the vertex count")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline measurement only")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also launch the copy kernel over this many MiB (byte-counter calibration)")
    return ap.parse_args()


class Batch:
    """Device buffers for one walker batch."""

    def __init__(self, ctx, sources, length, paths=True):
        self.ctx = ctx
        self.n = len(sources)
        self.length = length
        self.src = T.DeviceBuffer.from_array(ctx, sources)
        self.paths = T.DeviceBuffer(ctx, self.n * length * 4) if paths else None
        self.counters = T.DeviceBuffer.from_array(ctx, np.zeros(4, np.uint64))

    def io(self, draws=None, K=0):
        return T.make_io(self.src, self.n, self.length, 0, draws, K, None, 0, self.paths, counters=self.counters)

    def steps(self):
        return int(self.counters.to_numpy(np.uint64, 4)[0])

    def reset(self):
        self.counters.zero()

    def free(self):
        for b in (self.src, self.paths, self.counters):
            if b is not None:
                b.free()


def time_launches(ctx, graph, params, io, warmup, steps):
    for _ in range(warmup):
        T.walk_async(graph, params, io)
    T.synchronize(ctx)
    timer = T.DeviceTimer(ctx)
    timer.begin()
    for _ in range(steps):
        T.walk_async(graph, params, io)
    ms = timer.end()
    T.synchronize(ctx)
    return ms


def reference_cpu_baseline(hg, length):
    """The reference's DeepWalk (static alias, ring interleaving on) on every host core."""
    cores = len(os.sched_getaffinity(0))
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    if not os.path.exists(REF_BIN):
        return None
    allowed = os.sched_getaffinity(0)
    if any(c not in allowed for c in range(cores)):
        log("reference pins threads to CPUs 0..n-1, which are not all available here; skipping it")
        return None
    tmp = tempfile.mkdtemp(prefix="trw_ref_", dir="/tmp")
    try:
        t0 = time.time()
        hg.store(tmp)
        log("wrote reference CSR files in %.1fs; running the reference on %d cores" % (time.time() - t0, cores))
        r = subprocess.run([REF_BIN, "-f", tmp, "-ew", "-n", str(cores), "-l", str(length)], capture_output=True,
                           text=True, timeout=900)
        text = r.stdout + r.stderr
        if r.returncode != 0:
            log("reference exited with %d" % r.returncode)
            return None
        # Totals are logged by compute() (creeper.h:801-803); per-thread lines carry "Thread".
        totals = "\n".join(ln for ln in text.splitlines() if "func: compute)" in ln)
        thr = re.findall(r"Throughput \(steps per second\): ([0-9.]+)", totals)
        steps = re.findall(r"Num of steps: (\d+)", totals)
        walk_s = re.findall(r"Walking time \(seconds\): ([0-9.]+)", totals)
        pre = re.findall(r"Preprocessing time: ([0-9.]+)", text)
        if not thr:
            log("could not find the reference's throughput line")
            return None
        return {"value": float(thr[-1]), "unit": "steps/s", "cores": cores, "kind": "reference",
                "sample": "the reference's deepwalk.out (static alias, step interleaving on, -n %d) on the full "
                          "workload: the same graph as reference CSR files, one walker per non-isolated vertex, "
                          "length %d; %s steps in %ss of 'Walking time' (graph load and the %ss alias-table "
                          "build excluded, as the reference times it); CPU: %s"
                          % (cores, length, steps[-1] if steps else "?", walk_s[-1] if walk_s else "?",
                             pre[-1] if pre else "?", cpu)}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def load_traffic():
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return float(t["fetch_bytes_per_launch"]) + float(t["write_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def main():
    args = parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; this build has no multi-GPU path (walkers would shard by "
              "contiguous ranges with the graph replicated, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    nv = 1 << args.scale
    length = 80
    t0 = time.time()
    hg = T.HostGraph.chung_lu(nv, 8 * nv, 2.1, 2.0, 65536.0, args.seed, True, 0)
    deg, adj, w, _ = hg.arrays()
    log("graph: %d vertices, %d directed edges, max degree %d (%.1fs)" % (nv, len(adj), deg.max(), time.time() - t0))

    ctx = T.Context(0)
    dg = T.DeviceGraph(ctx, deg, adj, w)
    dg.build_tables("alias", "edge")
    alias_build_ms = dg.last_build_ms()
    log("alias tables built on the device in %.1f ms" % alias_build_ms)

    src = T.sources_all(deg)
    n = len(src)
    batch = Batch(ctx, src, length)
    prm = T.make_params("deepwalk", execution="static", sampling="alias", length=length, rng="philox", seed=2024,
                        form="fused")
    ms = time_launches(ctx, dg, prm, batch.io(), args.warmup, args.steps)
    total = batch.steps()
    per_launch = (length - 1) * n
    if total != per_launch * (args.warmup + args.steps):
        raise RuntimeError("step count %d != %d launches x %d" % (total, args.warmup + args.steps, per_launch))
    avg_s = ms / 1e3 / args.steps
    value = per_launch / avg_s
    achieved = per_launch * B_STEP["alias"] / avg_s / 1e9
    log("fused alias (philox): %.3f ms/launch, %.3e steps/s, %.1f GB/s algorithmic" % (avg_s * 1e3, value, achieved))

    details = {"alias_table_build_ms": round(alias_build_ms, 3), "walkers": n, "steps_per_launch": per_launch,
               "b_step_bytes": B_STEP["alias"]}
    if not args.quick:
        # Replay mode: the same walk consuming the host SFMT stream (bit-exact with the oracle).
        K = 2 * (length - 1)
        t1 = time.time()
        draws = T.replay_fill(1234, n, K)
        d_draws = T.DeviceBuffer.from_array(ctx, draws)
        del draws
        log("replay stream %d x %d generated and uploaded in %.1fs" % (K, n, time.time() - t1))
        batch.reset()
        prm_r = T.make_params("deepwalk", execution="static", sampling="alias", length=length, rng="replay",
                              form="fused")
        ms_r = time_launches(ctx, dg, prm_r, batch.io(d_draws, K), 1, args.steps)
        d_draws.free()
        v_r = per_launch / (ms_r / 1e3 / args.steps)
        details["alias_replay_steps_per_s"] = v_r
        details["alias_replay_roofline_frac"] = v_r * B_STEP["alias_replay"] / 1e9 / HBM_PEAK_GBS
        log("fused alias (replay): %.3e steps/s" % v_r)

        dg.build_tables("its", "edge")
        details["its_table_build_ms"] = round(dg.last_build_ms(), 3)
        prm_i = T.make_params("deepwalk", execution="static", sampling="its", length=length, rng="philox", seed=7,
                              form="fused")
        ms_i = time_launches(ctx, dg, prm_i, batch.io(), 1, args.steps)
        details["its_steps_per_s"] = per_launch / (ms_i / 1e3 / args.steps)
        log("fused ITS (philox): %.3e steps/s" % details["its_steps_per_s"])

    if not args.quick:
        # Measured ceiling for this access pattern: dependent random 16-byte gathers at full
        # occupancy over a table the size of the alias walk records (and a 32 MiB cache-resident one).
        for label, nbytes in (("probe_gathers_per_s_4GiB", 4 << 30), ("probe_gathers_per_s_32MiB", 32 << 20)):
            v = ctypes.c_double()
            T._abi.check(T.lib().trw_probe_random_gather(ctx.h, nbytes, 4, 64, ctypes.byref(v)))
            details[label] = v.value
        log("random-gather ceiling: %.3e gathers/s (4 GiB table)" % details["probe_gathers_per_s_4GiB"])

    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        a = T.DeviceBuffer(ctx, nbytes)
        b = T.DeviceBuffer(ctx, nbytes)
        a.zero()
        T.lib().trw_device_copy(ctx.h, b.ptr, a.ptr, nbytes, None)
        T.synchronize(ctx)
        a.free()
        b.free()

    cpu = None if args.no_cpu_baseline else reference_cpu_baseline(hg, length)
    if cpu:
        log("reference CPU baseline: %.3e steps/s on %d cores" % (cpu["value"], cpu["cores"]))

    traffic = load_traffic()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": avg_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "deepwalk-alias: chung-lu 2^%d V, avg deg 16, gamma 2.1, dmin 2, dmax 65536, "
                               "FP64 U[1,5) weights, length 80, one walker per non-isolated vertex, philox"
                               % args.scale,
                   "vertices": nv, "directed_edges": int(len(adj)), "walkers": n, "walk_length": length,
                   "kernel": "walk_fused_multi_kernel<alias, philox, 1 walker per lane>"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "details": details,
    }
    batch.free()
    dg.close()
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

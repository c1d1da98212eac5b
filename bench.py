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
import collections
import csv
import ctypes
import glob
import hashlib
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
# The sources the headline kernel's HBM traffic depends on (the kernel, its records, the
# table builds that lay them out); a traffic record taken on other sources is not used.
TRAFFIC_SOURCES = ("thunderrw_amd/csrc/walk.hip", "thunderrw_amd/csrc/engine.h", "thunderrw_amd/csrc/tables.hip")
# The headline kernel's instantiation: walk_fused_multi_kernel<alias Move (1), philox, walkers per lane>.
HEADLINE_KERNEL = "walk_fused_multi_kernel<1, false, %d>"
COPY_KERNEL = "copy_u128_kernel"


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
    ap.add_argument("--qps", type=float, default=10000.0, help="offered load of the PPR serving run")
    ap.add_argument("--queries", type=int, default=20000, help="queries in the PPR serving run")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-ppr-l1", dest="ppr_l1", action="store_false",
                    help="skip the served PPR estimates' L1 check against an FP64 power iteration")
    ap.add_argument("--sweep", type=float, nargs="*", default=[10000.0, 25000.0, 50000.0, 100000.0, 200000.0, 400000.0],
                    help="offered loads (queries/s) of the serving sweep")
    ap.add_argument("--only", choices=["its", "mix", "serving", "c1", "c3", "functor"],
                    help="run one secondary measurement alone and print its figures (for profiling)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also launch the copy kernel over this many MiB (byte-counter calibration)")
    ap.add_argument("--record-traffic", metavar="DIR",
                    help="write the headline traffic record from DIR/fetch and DIR/write (rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE -f csv outputs of a --quick --calibrate-copy run) and "
                         "DIR/bench.json (that run's output line); no GPU is used")
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

    def io(self, draws=None, K=0, unpadded=False, kind_rows=None):
        """kind_rows=(split, La, Lb): walkers [0, split) in rows of La ints, the rest in rows of
        Lb ints after them (io.path_stride_b), within the same buffer."""
        if kind_rows:
            split, La, Lb = kind_rows
            assert split * La + (self.n - split) * Lb <= self.n * self.length
            return T.make_io(self.src, self.n, La, 0, draws, K, None, 0, self.paths, counters=self.counters,
                             unpadded=unpadded, path_stride_b=Lb)
        return T.make_io(self.src, self.n, self.length, 0, draws, K, None, 0, self.paths, counters=self.counters,
                         unpadded=unpadded)

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


FUNCTOR_APP = os.path.join(ROOT, "thunderrw_amd", "dropin", "bin", "customwalk")


def functor_engine(ctx, dg, hg, deg, args):
    """The drop-in functor engine -- app functors the library has no kernel for, their weight()
    and update() run on the device -- on the same graph, through tests/dropin_apps/customwalk.cpp
    built against thunderrw_amd/dropin (its walk time is HIP events around the walk alone):
      * a Node2Vec functor under another name (max-weight rejection, a lane per walker, p = 2,
        q = 0.5), one walker per non-isolated vertex, beside the library's own Node2Vec kernel on
        the same walkers and Philox stream (the two walk identically);
      * DeepWalk's weights by dynamic ITS, and a second-order walker-dependent functor by dynamic
        alias: a wavefront per walker and a gather of every neighbour's weight per move, as the
        reference's Dynamic mode does, on 2^14 walkers from random sources."""
    if not os.path.exists(FUNCTOR_APP):
        log("functor engine: %s not built; skipped" % FUNCTOR_APP)
        return None
    tmp = tempfile.mkdtemp(prefix="trw_functor_", dir="/tmp")
    try:
        hg.store(tmp)
        env = dict(os.environ, TRW_RNG="philox", TRW_SEED="11")

        def run(extra):
            r = subprocess.run([FUNCTOR_APP, "-f", tmp, "-ew", "-l", "80"] + extra, env=env, capture_output=True,
                               text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError("functor app %s failed: %s" % (extra, r.stderr[-500:]))
            steps = int(re.search(r"Num of steps: (\d+)", r.stderr).group(1))
            secs = float(re.search(r"Walking time \(seconds\): ([0-9.]+)", r.stderr).group(1))
            return steps, secs

        out = {}
        steps, secs = run(["-kind", "4", "-em", "2", "-sm", "4"])
        out["node2vec_functor_steps_per_s"] = steps / secs
        src = T.sources_all(deg)
        b = Batch(ctx, src, 80)
        prm = T.make_params("node2vec", length=80, rng="philox", seed=11)
        ms = time_launches(ctx, dg, prm, b.io(), 1, 3)
        lib_steps = b.steps() // 4
        b.free()
        if lib_steps != steps:
            raise RuntimeError("library Node2Vec walked %d steps, the functor %d" % (lib_steps, steps))
        out["node2vec_library_steps_per_s"] = lib_steps / (ms / 1e3 / 3)
        # Dynamic mode: DeepWalk's weights by dynamic ITS (kind 3) and a second-order functor by
        # dynamic alias (kind 2 sampler, kind 1 functor), on 2^14 and 2^17 walkers from random
        # sources; beside them the reference's own engine (customwalk.out, the same functor built
        # against the reference's sources) on the same graph, walkers and flags on the host cores.
        for key, flags in (("dynamic_its", ["-kind", "3", "-em", "2", "-sm", "1"]),
                           ("second_order_dynamic_alias", ["-kind", "1", "-em", "2", "-sm", "2"])):
            for wn, tag in ((1 << 14, "sample"), (1 << 17, "2p17")):
                steps, secs = run(flags + ["-wn", str(wn)])
                out["%s_%s_steps_per_s" % (key, tag)] = steps / secs
            if not args.no_cpu_baseline:
                ref = reference_run(hg, "customwalk", ["-ew", "-l", "80"] + flags + ["-wn", str(1 << 14)])
                if ref:
                    out["reference_%s_sample_steps_per_s" % key] = ref["value"]
                    out["reference_cores"] = ref["cores"]
        log("functor engine: Node2Vec functor %.3e steps/s (library kernel %.3e); dynamic ITS %.3e (2^14 walkers), "
            "%.3e (2^17), reference %s; second-order dynamic alias %.3e (2^14), %.3e (2^17), reference %s"
            % (out["node2vec_functor_steps_per_s"], out["node2vec_library_steps_per_s"],
               out["dynamic_its_sample_steps_per_s"], out["dynamic_its_2p17_steps_per_s"],
               "%.3e" % out["reference_dynamic_its_sample_steps_per_s"]
               if "reference_dynamic_its_sample_steps_per_s" in out else "n/a",
               out["second_order_dynamic_alias_sample_steps_per_s"], out["second_order_dynamic_alias_2p17_steps_per_s"],
               "%.3e" % out["reference_second_order_dynamic_alias_sample_steps_per_s"]
               if "reference_second_order_dynamic_alias_sample_steps_per_s" in out else "n/a"))
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def host_cpu():
    """The host CPU as /proc/cpuinfo describes it: model name (which a virtualised host may cut
    short), CPUID family / model / stepping, and the logical CPUs of the host and of this
    process's affinity (the cores the reference is given)."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping", "vendor_id") and k not in info:
                    info[k] = v
    except OSError:
        pass
    return ("%s (%s, CPUID family %s, model %s, stepping %s; %d logical CPUs on the host, %d in this "
            "process's affinity)" % (info.get("model name", "unknown"), info.get("vendor_id", "?"),
                                     info.get("cpu family", "?"), info.get("model", "?"), info.get("stepping", "?"),
                                     os.cpu_count() or 0, len(os.sched_getaffinity(0))))


def reference_run(hg, app, flags):
    """Runs the reference's own app binary (oracle/_ref/<app>.out, built from its sources) on
    every host core, on `hg` written as reference CSR files; returns its compute() totals."""
    binary = os.path.join(ROOT, "oracle", "_ref", app + ".out")
    if not os.path.exists(binary):
        return None
    cores = len(os.sched_getaffinity(0))
    if any(c not in os.sched_getaffinity(0) for c in range(cores)):
        log("reference pins threads to CPUs 0..n-1, which are not all available here; skipping it")
        return None
    tmp = tempfile.mkdtemp(prefix="trw_ref_", dir="/tmp")
    try:
        hg.store(tmp)
        r = subprocess.run([binary, "-f", tmp, "-n", str(cores)] + flags, capture_output=True, text=True,
                           timeout=900)
        text = r.stdout + r.stderr
        if r.returncode != 0:
            log("reference %s exited with %d" % (app, r.returncode))
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
        return {"value": float(thr[-1]), "steps": int(steps[-1]) if steps else None,
                "walk_s": float(walk_s[-1]) if walk_s else None, "pre_s": float(pre[-1]) if pre else None,
                "cores": cores, "flags": " ".join(flags)}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def reference_cpu_baseline(hg, length):
    """The reference's DeepWalk (static alias, ring interleaving on) on every host core."""
    r = reference_run(hg, "deepwalk", ["-ew", "-l", str(length)])
    if r is None:
        return None
    return {"value": r["value"], "unit": "steps/s", "cores": r["cores"], "kind": "reference",
            "sample": "the reference's deepwalk.out (static alias, step interleaving on, -n %d) on the full "
                      "workload: the same graph as reference CSR files, one walker per non-isolated vertex, "
                      "length %d; %s steps in %ss of 'Walking time' (graph load and the %ss alias-table "
                      "build excluded, as the reference times it); CPU: %s"
                      % (r["cores"], length, r["steps"], r["walk_s"], r["pre_s"], host_cpu())}


def c1_longest_walk(ctx, dg, prm, src):
    """The batch's longest walk timed alone: walks are keyed by walker id, so the walker with the
    most steps, launched by itself under its own id, repeats its walk.  Its time is the floor of a
    batch launch (a chain of dependent gathers that no other walker can shorten)."""
    n = len(src)
    lengths = T.DeviceBuffer(ctx, n * 4)
    all_src = T.DeviceBuffer.from_array(ctx, src)
    io = T.make_io(all_src, n, lengths=lengths)
    T.walk_async(dg, prm, io)
    T.synchronize(ctx)
    lens = lengths.to_numpy(np.int32, n)
    w = int(np.argmax(lens))
    one_src = T.DeviceBuffer.from_array(ctx, src[w:w + 1].copy())
    one_id = T.DeviceBuffer.from_array(ctx, np.array([w], np.uint64))
    one_len = T.DeviceBuffer(ctx, 4)
    counters = T.DeviceBuffer.from_array(ctx, np.zeros(4, np.uint64))
    io1 = T.make_io(one_src, 1, lengths=one_len, counters=counters, walker_ids=one_id)
    T.walk_async(dg, prm, io1)
    T.synchronize(ctx)
    if int(one_len.to_numpy(np.int32, 1)[0]) != int(lens[w]):
        raise RuntimeError("the longest walker alone walked %d positions, in the batch %d"
                           % (int(one_len.to_numpy(np.int32, 1)[0]), int(lens[w])))
    reps = 20
    timer = T.DeviceTimer(ctx)
    timer.begin()
    for _ in range(reps):
        T.walk_async(dg, prm, io1)
    ms = timer.end() / reps
    T.synchronize(ctx)
    for b in (lengths, all_src, one_src, one_id, one_len, counters):
        b.free()
    return {"walker": w, "steps": int(lens[w]) - 1, "alone_ms": ms,
            "us_per_step_alone": ms * 1e3 / max(1, int(lens[w]) - 1)}


def c1_ppr(ctx, args):
    """PPR with uniform Moves and termination 0.2 on a 2^18-vertex Chung-Lu graph (exponent 2.1,
    dmin 2, dmax 4096, ~1.5M undirected edges, unweighted): one walker per non-isolated vertex,
    and the reference's default of |V| walkers from the max-degree vertex (ppr.cpp:53-56,76)."""
    hg = T.HostGraph.chung_lu(1 << 18, 1500000, 2.1, 2.0, 4096.0, 1, False, 0)
    deg, adj, _, _ = hg.arrays()
    dg = T.DeviceGraph(ctx, deg, adj)
    out = {"graph": "chung-lu 2^18 V, 1.5M undirected edges, gamma 2.1, dmin 2, dmax 4096, unweighted",
           "directed_edges": int(len(adj))}
    prm = T.make_params("ppr", stop_probability=0.2, rng="philox", seed=1)
    # The queue form's chunk and blocks per CU for this graph, tuned on a sample the size of the batch
    # (a batch of one walker per vertex: the grid's residency sets how many walkers a lane runs).
    out["autotune"] = tune_record(dg.autotune(prm, sample_walkers=1 << 18, seed=1, with_paths=False))
    log("C1 autotune: %s" % out["autotune"])
    for label, src in (("all_sources", T.sources_all(deg)), ("single_source", T.sources_single(deg, len(deg)))):
        b = Batch(ctx, src, 1, paths=False)
        io = b.io()
        T.walk_async(dg, prm, io)
        T.synchronize(ctx)
        b.reset()
        reps = 20
        timer = T.DeviceTimer(ctx)
        timer.begin()
        for _ in range(reps):
            T.walk_async(dg, prm, io)
        ms = timer.end() / reps
        T.synchronize(ctx)
        out[label + "_steps_per_s"] = b.steps() / reps / (ms / 1e3)
        out[label + "_ms"] = ms
        b.free()
        out[label + "_longest_walk"] = c1_longest_walk(ctx, dg, prm, src)
    dg.close()
    ref = None if args.no_cpu_baseline else reference_run(hg, "ppr", ["-sp", "0.2"])
    if ref:
        out["reference_single_source_steps_per_s"] = ref["value"]
        out["reference_cores"] = ref["cores"]
    hg.close()
    log("C1 PPR: %.3e steps/s (all sources), %.3e (single source), reference %s"
        % (out["all_sources_steps_per_s"], out["single_source_steps_per_s"],
           "%.3e" % ref["value"] if ref else "n/a"))
    return out


def c3_node2vec(ctx, args):
    """Node2Vec p=2, q=0.5 (max-weight rejection) on a 2^22-vertex Chung-Lu graph (exponent 2.0,
    average degree 32, dmin 2, dmax 65536, FP64 U[1,5) weights mirrored), length 80, one walker per
    non-isolated vertex: the reference's functor (node2vec.cpp, which ignores the weights) and
    Node2Vec over the edge weights (weight w_e times the factor, bound max w_e times max_)."""
    t0 = time.time()
    hg = T.HostGraph.chung_lu(1 << 22, 16 << 22, 2.0, 2.0, 65536.0, 3, True, 0)
    deg, adj, w, _ = hg.arrays()
    log("C3 graph: %d directed edges, max degree %d (%.1fs)" % (len(adj), deg.max(), time.time() - t0))
    dg = T.DeviceGraph(ctx, deg, adj, w)
    src = T.sources_all(deg)
    b = Batch(ctx, src, 80)
    reps = max(1, min(args.steps, 3))
    per_launch = 79 * len(src)
    out = {"graph": "chung-lu 2^22 V, avg deg 32, gamma 2.0, dmin 2, dmax 65536, FP64 U[1,5) weights",
           "directed_edges": int(len(adj)), "walkers": len(src), "steps_per_launch": per_launch}
    for label, weights in (("", "app"), ("weighted_", "edge")):
        prm = T.make_params("node2vec", length=80, p=2.0, q=0.5, rng="philox", seed=3, weights=weights)
        # One walk builds the membership tables (and the weighted records), then the Move's launch
        # shapes are tuned on a 2^21-walker sample before the timed launches.
        T.walk_async(dg, prm, b.io())
        T.synchronize(ctx)
        out[label + "autotune"] = tune_record(dg.autotune(prm, sample_walkers=1 << 21, seed=3))
        b.reset()
        ms = time_launches(ctx, dg, prm, b.io(), 1, reps)
        if b.steps() != per_launch * (reps + 1):
            raise RuntimeError("C3 %s: step count %d != %d launches x %d" % (weights, b.steps(), reps + 1, per_launch))
        out[label + "steps_per_s"] = per_launch / (ms / 1e3 / reps)
        out[label + "ms"] = ms / reps
    b.free()
    dg.close()
    if not args.no_cpu_baseline:
        ref = reference_run(hg, "node2vec", ["-l", "80", "-p", "2", "-q", "0.5"])
        if ref:
            out["reference_steps_per_s"] = ref["value"]
            out["reference_cores"] = ref["cores"]
        # The reference's engine (compute<F> with its ring interleaving) running the weighted functor
        # of tests/dropin_apps/customwalk.cpp (kind 7), built against the reference's sources.
        ref = reference_run(hg, "customwalk", ["-ew", "-l", "80", "-kind", "7", "-em", "2", "-sm", "4", "-p", "2",
                                                "-q", "0.5"])
        if ref:
            out["reference_weighted_steps_per_s"] = ref["value"]
    hg.close()
    log("C3 Node2Vec: %.3e steps/s (weighted %.3e), reference %s (weighted %s)"
        % (out["steps_per_s"], out["weighted_steps_per_s"],
           "%.3e" % out["reference_steps_per_s"] if "reference_steps_per_s" in out else "n/a",
           "%.3e" % out["reference_weighted_steps_per_s"] if "reference_weighted_steps_per_s" in out else "n/a"))
    return out


def headline_walkers_per_lane(tune):
    """Walkers per lane of the timed headline launches (the tuner's choice; 0 is the default, 1)."""
    return int(tune["chosen"].get("walkers_per_thread", 0)) or 1


def tune_record(tune):
    """What the autotuner kept for a walk: its Move, the chosen knobs, and the sample rates."""
    return {"move": tune["move"], "chosen": tune["chosen"],
            "default_steps_per_s": round(tune["default_steps_per_s"], 1),
            "tuned_steps_per_s": round(tune["tuned_steps_per_s"], 1)}


def its_walk(ctx, dg, batch, length, per_launch, args, hg=None):
    """DeepWalk with static ITS and with static rejection sampling on the same graph and walkers
    (paths written).  ITS: prefix sums, search trees and guide records built on the device (a
    fused step is one guide-record read); rejection: maxima and rejection records (a trial is
    one record read).  Launch knobs autotuned.  Beside each, the reference's own deepwalk.out
    with the same sampler (-em 1 -sm 1 / 3) on the host cores, unless --no-cpu-baseline."""
    out = {}
    for sm, key, code in (("its", "its", "1"), ("rejection", "rejection", "3")):
        if hg is not None and not args.no_cpu_baseline:
            ref = reference_run(hg, "deepwalk", ["-ew", "-l", str(length), "-em", "1", "-sm", code])
            if ref:
                out["reference_" + key + "_steps_per_s"] = ref["value"]
                out["reference_" + key + "_cores"] = ref["cores"]
        dg.build_tables(sm, "edge")
        out[key + "_table_build_ms"] = round(dg.last_build_ms(), 3)
        prm = T.make_params("deepwalk", execution="static", sampling=sm, length=length, rng="philox", seed=7,
                            form="fused")
        tune = dg.autotune(prm, sample_walkers=1 << 21, seed=7)
        out[key + "_autotune"] = tune_record(tune)
        batch.reset()
        ms = time_launches(ctx, dg, prm, batch.io(), 1, args.steps)
        out[key + "_steps_per_s"] = per_launch / (ms / 1e3 / args.steps)
        log("fused %s (philox, tuned %s): %.3e steps/s, reference %s" % (
            sm, tune["chosen"], out[key + "_steps_per_s"],
            "%.3e" % out["reference_" + key + "_steps_per_s"] if "reference_" + key + "_steps_per_s" in out else "n/a"))
    return out


def mix_forms(ctx, dg, deg, args, hg=None):
    """Uneven PPR + MetaPath mix: 2^23 walkers from uniform random non-isolated sources
    (generate_random_source), walkers [0, 2^22) PPR with termination 0.15 and length 1
    (g_para.length_ = 1, ppr.cpp:51: the path row holds the source), the rest MetaPath on
    schema 0,1,2,3,4 with length 80 (metapath.cpp:48), paths logged.  Each kind's rows are as
    long as its walks (io.path_stride_b: PPR's 1-entry rows, then MetaPath's 80-entry rows), as
    the two apps' own runs would write them; "*_one_stride" is the same batch in one matrix of
    80-entry rows.  The same batch runs in the persistent work-queue form, the compacted
    hipGraph step form and the plain per-step-launch form; the forms give identical walks."""
    n = 1 << 23
    split = n // 2
    src = T.sources_random(deg, n, 4)
    batch = Batch(ctx, src, 80)
    out = {"walkers": n, "split": split}
    # MetaPath's own launch shapes first, then the mix's queue chunk and blocks per CU on a mixed
    # sample the size of the batch (the chunk is per graph, the shapes per Move; the queue kernel of a
    # mix takes the first kind's shapes).
    pa = T.make_params("ppr", length=1, stop_probability=0.15, rng="philox", seed=4, form="queue")
    pb = T.make_params("metapath", length=80, schema=[0, 1, 2, 3, 4], rng="philox", seed=4, form="queue")
    out["metapath_queue_autotune"] = tune_record(dg.autotune(pb, sample_walkers=1 << 20, seed=4, with_paths=True))
    out["queue_autotune"] = tune_record(dg.autotune_mix(pa, pb, split, sample_walkers=n, seed=4, with_paths=True))
    # "queue_unpadded": the same walks with TRW_PATHS_UNPADDED, each row holding its walker's
    # length entries as the reference's LOG_SEQUENCE output does, with no -1 padding written.
    for form in ("queue", "queue_unpadded", "graph", "step", "queue_one_stride", "queue_unpadded_one_stride"):
        kf = form.split("_")[0]
        a = T.make_params("ppr", length=1, stop_probability=0.15, rng="philox", seed=4, form=kf)
        b = T.make_params("metapath", length=80, schema=[0, 1, 2, 3, 4], rng="philox", seed=4, form=kf)
        io = batch.io(unpadded="unpadded" in form, kind_rows=None if "one_stride" in form else (split, 1, 80))
        T.walk_mix_async(dg, a, b, split, io)
        T.synchronize(ctx)
        batch.reset()
        reps = max(1, min(args.steps, 3))
        timer = T.DeviceTimer(ctx)
        timer.begin()
        for _ in range(reps):
            T.walk_mix_async(dg, a, b, split, io)
        ms = timer.end()
        T.synchronize(ctx)
        steps = batch.steps() // reps
        out[form + "_steps_per_s"] = steps / (ms / 1e3 / reps)
        out[form + "_ms"] = ms / reps
        out["steps_per_batch"] = steps
        log("mix %s form: %.2f ms, %.3e steps/s" % (form, ms / reps, out[form + "_steps_per_s"]))
    out["batched_vs_plain"] = out["queue_steps_per_s"] / out["step_steps_per_s"]
    out["graph_vs_plain"] = out["graph_steps_per_s"] / out["step_steps_per_s"]
    batch.free()
    # The reference's CPU path for each kind on the same graph (it has no mixed batch): its ppr.out at
    # termination 0.15 (|V| walkers from the max-degree vertex, the app's default) and its
    # metapath.out on schema 0,1,2,3,4 with length 80 (one walker per non-isolated vertex).
    if hg is not None and not args.no_cpu_baseline:
        for app, flags in (("ppr", ["-sp", "0.15"]), ("metapath", ["-el", "-s", "0,1,2,3,4", "-l", "80"])):
            ref = reference_run(hg, app, flags)
            if ref:
                out["reference_%s_steps_per_s" % app] = ref["value"]
                out["reference_%s_cores" % app] = ref["cores"]
        log("mix kinds on the reference's CPU path: PPR %s, MetaPath %s steps/s"
            % tuple("%.3e" % out["reference_%s_steps_per_s" % a] if "reference_%s_steps_per_s" % a in out else "n/a"
                    for a in ("ppr", "metapath")))
    return out


def latency_parts(srv, ids):
    """Where the served queries' time went: per query, arrival -> first launch (queue), first
    launch -> its last batch seen done on the device (device), and the host-side completion after
    it; p50 and p99 of each, and their means over the slowest 1% of queries."""
    rows = []
    for q in ids:
        i = srv.info(int(q))
        if i["latency_ms"] >= 0 and i["device_ms"] >= 0:
            rows.append((i["latency_ms"], i["queue_ms"], i["device_ms"],
                         i["latency_ms"] - i["queue_ms"] - i["device_ms"]))
    if not rows:
        return None
    a = np.array(rows)
    slow = a[a[:, 0] >= np.percentile(a[:, 0], 99)]
    out = {}
    for j, name in ((1, "queue"), (2, "device"), (3, "completion")):
        out[name + "_p50_ms"] = float(np.percentile(a[:, j], 50))
        out[name + "_p99_ms"] = float(np.percentile(a[:, j], 99))
        out[name + "_slowest1pct_mean_ms"] = float(slow[:, j].mean())
    return out


def symmetric_sample(deg, adj, w, off, n=4096, seed=7):
    """True when n random edges (u, v, w) each have their reverse (v, u, w): the serving graph is
    symmetrised with mirrored weights by construction, which lets the power iteration use the
    weighted adjacency itself for the transpose."""
    rng = np.random.default_rng(seed)
    e = rng.integers(0, len(adj), n)
    u = np.searchsorted(off, e, side="right") - 1
    for ei, ui in zip(e, u):
        v = int(adj[ei])
        row = adj[off[v]:off[v + 1]]
        j = int(np.searchsorted(row, ui))
        if j >= len(row) or row[j] != ui or w[off[v] + j] != w[ei]:
            return False
    return True


def ppr_power_host(A, strength, starts, alpha, tol=1e-4):
    """FP64 power iteration of the PPR end-vertex law for several start distributions at once (one
    column each): sum_{k>=1} a (1-a)^(k-1) x0 P^k with P(u->v) = w_uv / sum_u w, the reference
    moving before it tests termination (creeper.h:306, ppr.cpp:27-29); a vertex without edges keeps
    its mass.  Iterated until the mass not yet placed is below tol.  A is the weighted adjacency of
    a symmetric graph, so y = A (x / s) is P^T x; its row blocks are multiplied on the host's cores
    in parallel.  Returns (pi, iterations, mass left)."""
    from concurrent.futures import ThreadPoolExecutor
    import scipy.sparse as sp
    n = A.shape[0]
    nt = max(1, min(16, os.cpu_count() or 1))
    cuts = np.searchsorted(A.indptr, np.linspace(0, A.nnz, nt + 1)).clip(0, n)
    cuts[0], cuts[-1] = 0, n
    blocks = []
    for r0, r1 in zip(cuts[:-1], cuts[1:]):
        o0, o1 = int(A.indptr[r0]), int(A.indptr[r1])
        blocks.append(sp.csr_matrix((A.data[o0:o1], A.indices[o0:o1], A.indptr[r0:r1 + 1] - o0),
                                    shape=(int(r1 - r0), n)))
    x = np.zeros((n, len(starts)))
    for j, s in enumerate(starts):
        np.add.at(x[:, j], s, 1.0 / len(s))
    dead = strength == 0
    inv = np.where(dead, 0.0, 1.0 / np.where(dead, 1.0, strength))[:, None]
    pi = np.zeros_like(x)
    coef, left, iters = alpha, 1.0, 0
    with ThreadPoolExecutor(nt) as pool:
        while left > tol:
            xs = x * inv
            y = np.concatenate(list(pool.map(lambda b: b @ xs, blocks)))
            y[dead] += x[dead]
            pi += coef * y
            x = y
            coef *= 1.0 - alpha
            left *= 1.0 - alpha
            iters += 1
    return pi, iters, left


def judge_ppr(pi, checks):
    """Each served estimate (end-vertex walker counts / N) against its law pi (one column per
    check): the L1 distance beside what sampling alone gives (expected_l1), and Pearson's
    chi-square over the vertices where the law expects 5 or more walks.  ok: L1 within 4 / sqrt(N)
    (about four standard deviations of L1) of the sampling figure and the chi-square tail above
    1e-4."""
    out = []
    for j, (src, nw, v, c) in enumerate(checks):
        counts = np.zeros(pi.shape[0])
        counts[v] = c
        p = pi[:, j]
        l1 = float(np.abs(counts / nw - p).sum())
        el1 = expected_l1(p, nw)
        stat, dof, pval = head_chi_square(counts, p, nw)
        out.append({"sources": int(len(src)), "walkers": int(nw), "l1": round(l1, 5),
                    "expected_l1": round(el1, 5), "chi2": round(stat, 2), "chi2_dof": dof, "chi2_p": pval,
                    "ok": bool(l1 <= el1 + 4.0 / np.sqrt(nw) and pval > 1e-4)})
    return out


def ppr_l1(A, strength, checks, alpha, tol=1e-4):
    """Served PPR queries (sources, walkers, vertices, counts) judged against the FP64 power
    iteration from the walkers' own starts (walker i starts at sources[i % |S|])."""
    t0 = time.time()
    pi, iters, left = ppr_power_host(A, strength, [src[np.arange(nw) % len(src)] for src, nw, _, _ in checks],
                                     alpha, tol)
    out = judge_ppr(pi, checks)
    log("served PPR vs FP64 power iteration (%d iterations, %.1fs): %s"
        % (iters, time.time() - t0,
           ", ".join("|S| %d N %d: L1 %.4f (sampling alone %.4f), head chi2 %.1f / %d dof, p %.3g"
                     % (o["sources"], o["walkers"], o["l1"], o["expected_l1"], o["chi2"], o["chi2_dof"], o["chi2_p"])
                     for o in out)))
    return {"iterations": iters, "mass_left": left, "queries": out}


def expected_l1(p, n):
    """E sum_v |X_v / n - p_v| for multinomial counts X of n draws from p: per vertex the binomial's
    mean absolute deviation, 2 (m + 1) C(n, m + 1) p^(m+1) (1 - p)^(n-m) with m = floor(n p)
    (de Moivre), in logs.  A walk batch's counts are a sum of such draws from each source's law,
    whose spread is at most this, so it bounds what sampling alone gives."""
    from scipy.special import gammaln
    p = p[(p > 0) & (p < 1)]
    m = np.floor(n * p)
    lg = (np.log(2.0) + np.log(m + 1) + gammaln(n + 1) - gammaln(m + 2) - gammaln(n - m)
          + (m + 1) * np.log(p) + (n - m) * np.log1p(-p))
    return float(np.exp(lg).sum() / n)


def head_chi_square(counts, p, n, min_expected=5.0):
    """Pearson's statistic over the vertices expected to end at least min_expected walks, the rest
    lumped into one bin; its degrees of freedom and the chi-square upper tail."""
    from scipy.stats import chi2
    e = n * p
    head = e >= min_expected
    obs, exp = list(counts[head]), list(e[head])
    if n - sum(exp) >= min_expected:
        obs.append(n - sum(obs))
        exp.append(n - sum(exp))
    obs, exp = np.array(obs), np.array(exp)
    if len(obs) < 2:
        return 0.0, 0, 1.0
    stat = float(((obs - exp) ** 2 / exp).sum())
    return stat, len(obs) - 1, float(chi2.sf(stat, len(obs) - 1))


def serving(ctx, args):
    """PPR query serving on a 2^24-vertex Chung-Lu graph (exponent 2.1, average degree 16,
    dmin 2, dmax 65536, FP64 U[1,5) weights, weighted alias tables resident).  Poisson
    arrivals at a fixed rate; each query has |S| ~ logU[1, 4096] sources drawn uniformly
    from the non-isolated vertices and N ~ logU[1e3, 1e5] walkers; termination 0.2.  The
    server runs 8 slots (batches in flight) of up to 2^21 walkers with 8 completion threads."""
    t0 = time.time()
    hg = T.HostGraph.chung_lu(1 << 24, 8 << 24, 2.1, 2.0, 65536.0, 5, True, 0)
    deg, adj, w, _ = hg.arrays()
    log("serving graph: %d directed edges (%.1fs)" % (len(adj), time.time() - t0))
    dg = T.DeviceGraph(ctx, deg, adj, w)
    live = np.nonzero(deg > 0)[0].astype(np.int32)
    # The weighted adjacency and vertex strengths for the PPR L1 check (copies: the arrays are views
    # into the host graph, which is freed next).
    adjacency = None
    if args.ppr_l1:
        import scipy.sparse as sp
        off = np.zeros(len(deg) + 1, np.int64)
        np.cumsum(deg, out=off[1:])
        if not symmetric_sample(deg, adj, w, off):
            raise RuntimeError("the serving graph is not symmetric with mirrored weights")
        strength = np.zeros(len(deg))
        nz = deg > 0
        strength[nz] = np.add.reduceat(w, off[:-1][nz])
        adjacency = sp.csr_matrix((w.copy(), adj.copy(), off), shape=(len(deg), len(deg)))
    del deg, adj, w
    hg.close()
    dg.build_tables("alias", "edge")
    rng = np.random.default_rng(5)
    prm = T.make_params("ppr", execution="static", sampling="alias", weights="edge", stop_probability=0.2,
                        rng="philox", seed=5)
    out = {"graph": "chung-lu 2^24 V, avg deg 16, gamma 2.1, dmin 2, dmax 65536, FP64 U[1,5) weights",
           "queries": args.queries, "offered_qps": args.qps, "alias_table_build_ms": round(dg.last_build_ms(), 3)}
    keys = ("achieved_qps", "p50_ms", "p90_ms", "p99_ms", "p999_ms", "mean_ms", "max_ms", "steps_per_s", "steps",
            "batches", "mean_batch_queries", "rejected")
    # The served walk's launch shapes (queue chunk, blocks per CU) tuned on a 2^20-walker sample.
    out["autotune"] = tune_record(dg.autotune(T.make_params("ppr", execution="static", sampling="alias", weights="edge",
                                                            stop_probability=0.2, rng="philox", seed=5, form="queue"),
                                              sample_walkers=1 << 20, seed=5, with_paths=False))
    out["server"] = {"streams": 8, "completion_threads": 8, "max_batch_walkers": 1 << 21}
    with T.Server(dg, **out["server"]) as srv:
        qs = []
        for i in range(args.queries):
            k = int(np.exp(rng.uniform(0.0, np.log(4096.0))))
            nw = int(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
            qs.append(srv.query(prm, rng.choice(live, k), nw, (i + 1) << 32))
        # Warm-up: 4,000 queries at the sweep's highest rate, every slot busy with full batches.  The
        # first time the load reaches a new level the device stalls once for 7 - 12 ms (measured with
        # per-query device times: every batch then in flight is seen done that much late); the
        # server's own warm-up at creation removes most of these, a burst at full load the rest.
        _, ids, _ = srv.run_open_loop(qs[:4000], max(args.sweep + [args.qps]), seed=1)
        for q in ids:
            srv.release(int(q))
        srv.run_open_loop(qs[:200], args.qps, seed=1)
        rep, ids, lat = srv.run_open_loop(qs, args.qps, seed=2)
        st = srv.stats()
        # Served estimates kept for the L1 check: the queries with the fewest and the most walkers
        # and with the fewest and the most sources.
        checks = []
        if adjacency is not None:
            nws = np.array([q.n_walkers for q in qs])
            nss = np.array([q.n_sources for q in qs])
            for i in dict.fromkeys([int(nws.argmin()), int(nws.argmax()), int(nss.argmin()), int(nss.argmax())]):
                v, c = srv.ppr(int(ids[i]))
                checks.append((qs[i]._src, int(qs[i].n_walkers), v, c.astype(np.float64)))
        for q in ids:
            srv.release(int(q))
        # Offered-load sweep: the same queries at rising Poisson rates; where the server saturates the
        # achieved rate falls below the offered one and latency is queueing.
        sweep = []
        for qps in args.sweep:
            r, ids, _ = srv.run_open_loop(qs, qps, seed=3)
            parts = latency_parts(srv, ids)
            for q in ids:
                srv.release(int(q))
            sweep.append(dict({k: r[k] for k in ("offered_qps",) + keys}, parts=parts))
            log("serving sweep: offered %.0f qps -> achieved %.0f qps, p50 %.3f ms, p99 %.3f ms, %.3e steps/s, "
                "%.1f queries per batch; slowest 1%%: queue %.3f, device %.3f, completion %.3f ms"
                % (qps, r["achieved_qps"], r["p50_ms"], r["p99_ms"], r["steps_per_s"], r["mean_batch_queries"],
                   parts["queue_slowest1pct_mean_ms"], parts["device_slowest1pct_mean_ms"],
                   parts["completion_slowest1pct_mean_ms"]))
    dg.close()
    if checks:
        out["ppr_l1"] = ppr_l1(adjacency, strength, checks, 0.2)
    del adjacency
    out.update({k: rep[k] for k in keys})
    out["mean_batch_walkers"] = st["mean_batch_walkers"]
    out["sweep"] = sweep
    if sweep:
        best = max(sweep, key=lambda x: x["steps_per_s"])
        out["saturation"] = {"offered_qps": best["offered_qps"], "achieved_qps": best["achieved_qps"],
                             "steps_per_s": best["steps_per_s"], "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"]}
    log("serving at %.0f qps: p50 %.3f ms, p99 %.3f ms, %.3e steps/s, %.1f queries per batch"
        % (args.qps, rep["p50_ms"], rep["p99_ms"], rep["steps_per_s"], rep["mean_batch_queries"]))
    return out


def kernel_source_hash():
    h = hashlib.sha256()
    for rel in TRAFFIC_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def load_traffic():
    """HBM bytes per headline launch from the counter record, or None when the record is
    missing or was taken on other kernel sources."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        total = float(t["fetch_bytes_per_launch"]) + float(t["write_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None
    if t.get("source_sha256") != kernel_source_hash():
        log("traffic record %s was taken on other kernel sources; roofline.traffic is null" % TRAFFIC_FILE)
        return None
    return total


def read_pmc(pass_dir, counter):
    """{kernel name: [value per dispatch, in dispatch order]} of one rocprofv3 --pmc csv pass."""
    files = glob.glob(os.path.join(pass_dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % pass_dir)
    per = collections.defaultdict(float)
    names = {}
    for fn in files:
        with open(fn) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                key = (fn, int(row["Dispatch_Id"]))
                per[key] += float(row["Counter_Value"])
                names[key] = row["Kernel_Name"]
    out = collections.defaultdict(list)
    for key in sorted(per):
        out[names[key]].append(per[key])
    return out


def record_traffic(d):
    """Headline HBM traffic per launch from two counter passes, each scaled by what the same
    counter reads for the copy kernel over a known size (bytes moved = counter x scale)."""
    with open(os.path.join(d, "bench.json")) as f:
        line = json.loads(f.read().strip().splitlines()[-1])
    steps = int(line["details"]["steps_per_launch"])
    fetch, write = read_pmc(os.path.join(d, "fetch"), "FETCH_SIZE"), read_pmc(os.path.join(d, "write"), "WRITE_SIZE")

    def one(table, name):
        hits = [v for k, v in table.items() if name in k]
        if len(hits) != 1 or not hits[0]:
            raise SystemExit("expected one kernel named %s in the pass, got %d" % (name, len(hits)))
        return hits[0][-1]  # the last dispatch: the timed launch

    nbytes = int(line["details"]["calibrate_copy_bytes"])
    kernel = HEADLINE_KERNEL % headline_walkers_per_lane(line["details"]["autotune"]["alias"])
    f_walk, w_walk = one(fetch, kernel), one(write, kernel)
    f_copy, w_copy = one(fetch, COPY_KERNEL), one(write, COPY_KERNEL)
    fs, ws = nbytes / (f_copy * 1024.0), nbytes / (w_copy * 1024.0)
    rec = {"kernel": line["config"]["kernel"], "workload": line["config"]["workload"],
           "source_sha256": kernel_source_hash(), "sources": list(TRAFFIC_SOURCES),
           "steps_per_launch": steps, "fetch_size_kib_raw": f_walk, "write_size_kib_raw": w_walk,
           "copy_kernel_bytes": nbytes, "copy_fetch_kib_raw": f_copy, "copy_write_kib_raw": w_copy,
           "fetch_scale": fs, "write_scale": ws,
           "fetch_bytes_per_launch": f_walk * 1024.0 * fs, "write_bytes_per_launch": w_walk * 1024.0 * ws}
    rec["bytes_per_step"] = (rec["fetch_bytes_per_launch"] + rec["write_bytes_per_launch"]) / steps
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec))


def main():
    args = parse_args()
    if args.record_traffic:
        record_traffic(args.record_traffic)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU; the library has no multi-GPU path (walkers would shard by "
              "contiguous ranges, the graph and its tables replicated on every GPU)", file=sys.stderr)
        sys.exit(2)
    if args.only in ("serving", "c1", "c3"):
        ctx = T.Context(0)
        fn = {"serving": ("ppr_serving", serving), "c1": ("c1_ppr", c1_ppr), "c3": ("c3_node2vec", c3_node2vec)}
        key, f = fn[args.only]
        print(json.dumps({key: f(ctx, args)}), flush=True)
        ctx.close()
        return
    nv = 1 << args.scale
    length = 80
    t0 = time.time()
    # Labels U{0..4} ride along for the PPR + MetaPath mix; they do not change the structure.
    hg = T.HostGraph.chung_lu(nv, 8 * nv, 2.1, 2.0, 65536.0, args.seed, True, 5)
    deg, adj, w, lab = hg.arrays()
    log("graph: %d vertices, %d directed edges, max degree %d (%.1fs)" % (nv, len(adj), deg.max(), time.time() - t0))

    ctx = T.Context(0)
    dg = T.DeviceGraph(ctx, deg, adj, w, lab)
    dg.build_tables("alias", "edge")
    alias_build_ms = dg.last_build_ms()
    log("alias tables built on the device in %.1f ms" % alias_build_ms)

    if args.only in ("mix", "functor"):
        if args.only == "mix":
            print(json.dumps({"mix_ppr_metapath": mix_forms(ctx, dg, deg, args, hg)}), flush=True)
        else:
            print(json.dumps({"functor_engine": functor_engine(ctx, dg, hg, deg, args)}), flush=True)
        dg.close()
        ctx.close()
        return
    src = T.sources_all(deg)
    n = len(src)
    batch = Batch(ctx, src, length)
    if args.only == "its":
        print(json.dumps(its_walk(ctx, dg, batch, length, (length - 1) * n, args, hg)), flush=True)
        batch.free()
        dg.close()
        ctx.close()
        return
    prm = T.make_params("deepwalk", execution="static", sampling="alias", length=length, rng="philox", seed=2024,
                        form="fused")
    # Launch shapes for this graph and Move (walkers per lane, threads per block, path tile), timed on a
    # 2^21-walker sample; the walks are the same under every choice.
    head_tune = tune_record(dg.autotune(prm, sample_walkers=1 << 21, seed=2024))
    log("autotune (alias): %s" % head_tune)
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
               "b_step_bytes": B_STEP["alias"], "autotune": {"alias": head_tune}}
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

        details.update(its_walk(ctx, dg, batch, length, per_launch, args, hg))

        # The same walk through the host-memory entry point: sources in and 1.3 GB of paths out over
        # PCIe (pageable host buffers), device buffers allocated by the call.  Reported beside the
        # headline, never as it.
        paths_h = np.empty((n, length), dtype=np.int32)
        steps_h = ctypes.c_uint64()
        t2 = time.perf_counter()
        T._abi.check(T.lib().trw_walk_host(dg.h, ctypes.byref(prm), src.ctypes.data, n, None, 0, None, 0,
                                           paths_h.ctypes.data, None, None, None, ctypes.byref(steps_h)))
        secs_h = time.perf_counter() - t2
        del paths_h
        details["host_buffers_steps_per_s"] = steps_h.value / secs_h
        details["host_buffers_ms"] = secs_h * 1e3
        log("same walk through trw_walk_host (paths to pageable host memory): %.1f ms, %.3e steps/s"
            % (secs_h * 1e3, steps_h.value / secs_h))

    if not args.quick:
        # Measured ceiling for this access pattern: dependent random 16-byte gathers at full
        # occupancy over a table the size of the alias walk records (and a 32 MiB cache-resident one).
        for label, nbytes, rec in (("probe_gathers_per_s_4GiB", 4 << 30, 16),
                                   ("probe_gathers_per_s_32MiB", 32 << 20, 16),
                                   ("probe_32B_records_per_s_2GiB", 2 << 30, 32)):
            v = ctypes.c_double()
            T._abi.check(T.lib().trw_probe_random_gather(ctx.h, nbytes, rec, 4, 64, ctypes.byref(v)))
            details[label] = v.value
        log("random-gather ceiling: %.3e 16-byte gathers/s (4 GiB table), %.3e 32-byte records/s (2 GiB)"
            % (details["probe_gathers_per_s_4GiB"], details["probe_32B_records_per_s_2GiB"]))

    if not args.quick:
        details["mix_ppr_metapath"] = mix_forms(ctx, dg, deg, args, hg)
        details["functor_engine"] = functor_engine(ctx, dg, hg, deg, args)
    batch.free()
    dg.close()
    if not args.quick and not args.no_serving:
        details["ppr_serving"] = serving(ctx, args)
    if not args.quick:
        details["c1_ppr"] = c1_ppr(ctx, args)
        details["c3_node2vec"] = c3_node2vec(ctx, args)

    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        a = T.DeviceBuffer(ctx, nbytes)
        b = T.DeviceBuffer(ctx, nbytes)
        a.zero()
        T.lib().trw_device_copy(ctx.h, b.ptr, a.ptr, nbytes, None)
        T.synchronize(ctx)
        details["calibrate_copy_bytes"] = nbytes
        a.free()
        b.free()

    cpu = None if args.no_cpu_baseline else reference_cpu_baseline(hg, length)
    if cpu:
        log("reference CPU baseline: %.3e steps/s on %d cores" % (cpu["value"], cpu["cores"]))

    # The autotuner's choices for every Move the run walked, in one place.
    at = details["autotune"]
    for key in ("its_autotune", "rejection_autotune"):
        if key in details:
            at[details[key]["move"]] = details[key]
    for key in ("metapath_queue_autotune", "queue_autotune"):
        rec = (details.get("mix_ppr_metapath") or {}).get(key)
        if rec:
            at[rec["move"] + "_queue"] = rec
    rec = (details.get("ppr_serving") or {}).get("autotune")
    if rec:
        at[rec["move"] + "_queue_serving"] = rec
    rec = (details.get("c1_ppr") or {}).get("autotune")
    if rec:
        at[rec["move"] + "_queue_c1"] = rec
    for key, name in (("autotune", "node2vec"), ("weighted_autotune", "node2vec_weighted")):
        rec = (details.get("c3_node2vec") or {}).get(key)
        if rec:
            at[name] = rec

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
                   "kernel": "walk_fused_multi_kernel<alias, philox, %d walker(s) per lane>"
                             % headline_walkers_per_lane(head_tune)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "details": details,
    }
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

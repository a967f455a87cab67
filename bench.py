#!/usr/bin/env python3
"""Throughput of whole-read alignment on one MI355X.

Workload: the 5 Mbp synthetic variation graph (SNP and indel bubbles, one
structural-variant bubble of 50-5,000 bp per 20 kbp: insertions, deletions,
inversions) with 20,000 reads of log-normal length (median 12 kbp, clipped to
1-60 kbp) at 12% error, aligned with preset -x vg. One step aligns every read:
device seed lookup, seed ordering, the seed-extension loops (two read groups on
two streams, each round's extensions of a group batched into one gfx950 launch,
later rounds extending several seeds per read ahead and replaying the loop),
selection and the GAF text, handed to Python as a buffer over the library's
GAF bytes (no str decoding of the ~350 MB per step inside the timed region; the
parity check decodes it afterwards). Reads are resident in HBM before the timed
region (ga_reads_upload); their reverse complements and match-mask planes are
computed again by every step.

--workload stream runs the batched form instead: the same graph with a mixed
stream of 100,000 reads (log-normal length, median 10 kbp, sigma 0.9, clipped to
1-100 kbp; error 5-15 % drawn per read; minimizers of 21 bases in windows of 31,
5th-95th percentile 5-193 seed hits per read), each batch of 16,384 reads run by
one captured HIP graph of a persistent kernel whose workers pull whole reads
from a device counter (worker count autotuned once per context).

Prints one JSON line (reads/s, GCUPS, roofline of the extension kernel, and the
CPU baseline: the C++ CPU restatement of the reference run on a sample of the
same reads with the host's cores).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# 256 CUs x 64 int32 lanes per clock x 2.4 GHz
VALU_INT32_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
# algorithmic work of one 64-row column step of the bit-parallel DP:
# 42 int32-equivalent VALU ops, ~6 bytes of HBM traffic (node bases, carries,
# band item state amortised over the node's columns)
OPS_PER_WORD_STEP = 42
BYTES_PER_WORD_STEP = 6
# HBM bytes per extension launch (mean) of the default run of this workload: a
# constant measured offline with rocprofv3 FETCH_SIZE and WRITE_SIZE (each scaled
# by the reading of a 1 GiB copy kernel) on the default command, recorded in
# profiles/extend_vg5m.md; reported as such, not measured by this run
MEASURED_TRAFFIC_PER_LAUNCH = 5.41e9
TRAFFIC_SOURCE = "constant: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of the default run, profiles/extend_vg5m.md"
# The extension kernel runs one extension per single-lane workgroup, so its
# word arithmetic and control flow issue on the scalar unit: one scalar
# instruction per clock per CU, 256 CUs x 2.4 GHz. Scalar instructions per
# forward word step of the default run (SQ_INSTS_SALU over its launches / its
# word steps) are a constant measured offline, recorded in profiles/extend_vg5m.md
SALU_PEAK_GIPS = 256 * 2.4
SALU_PER_WORD_STEP = 158.8
SALU_SOURCE = "constant: rocprofv3 --pmc SQ_INSTS_SALU of the default run, profiles/extend_vg5m.md"


WORKLOADS = {
    "dbg2m": dict(preset="dbg2m", k=19, w=30, params="dbg",
                  text="dbg2m: cyclic de Bruijn graph (2 Mbp genome, repeat families, tandem arrays; k=63 unitigs), "
                       "10,000 reads of 10-30 kbp at 8% error, -x dbg"),
    "vg5m": dict(preset="vg5m", k=15, w=20,
                 text="vg5m: 5 Mbp variation graph with SV bubbles, 20,000 log-normal reads (median 12 kbp) at 12% error, -x vg"),
    "stream": dict(preset="vg5m_stream", k=21, w=31,
                   text="vg5m stream: 5 Mbp variation graph with SV bubbles, 100,000 log-normal reads (median 10 kbp, 1-100 kbp) "
                        "at 5-15% error drawn per read, -x vg with minimizers k=21 w=31"),
}


def cpu_share():
    """Host threads this process may use: the cores it is allowed to run on,
    within the per-GPU CPU share the host grants (OMP_NUM_THREADS when set)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="vg5m")
    ap.add_argument("--reads", type=int, default=0, help="use only the first N reads (0: all of the workload's)")
    ap.add_argument("--warmup-reads", type=int, default=0, help="reads aligned by each warmup step (0: all, as a timed step)")
    ap.add_argument("--cpu-sample", type=int, default=4000, help="reads in the CPU-baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU-baseline threads (0: the host cores this process may use)")
    ap.add_argument("--workdir", default=os.environ.get("GA_BENCH_DIR", ""))
    ap.add_argument("--seed-loop", choices=["auto", "rounds", "device", "batched"], default="auto",
                    help="rounds: host-driven rounds of batched extensions; device: each read's whole seed loop on the GPU; "
                         "batched: the persistent form (auto: batched for the stream, rounds otherwise)")
    ap.add_argument("--seed-order", choices=["host", "device"], default="host",
                    help="seed expansion and OrderSeeds on host threads or on the GPU (same seeds)")
    ap.add_argument("--learned", action="store_true",
                    help="the learned proposal: train the network on outcome records of the plain method (fresh seeds of "
                         "the 5 Mbp variation-graph generator), then time plain and learned on --workload (vg5m or dbg2m)")
    ap.add_argument("--train-records", type=int, default=1_000_000, help="--learned: outcome records to train on (at least)")
    ap.add_argument("--longest-extensions", type=int, default=0, metavar="N",
                    help="only time one launch holding the forward extensions of the N longest reads of --workload from "
                         "their first seeds (the latency of the longest extensions without the rest of a launch) and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes (traffic-counter calibration) and exit")
    return ap.parse_args()


def copy_calibration(nbytes):
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "graphaligner_amd", "_build", "libga_calib.so"))
    lib.ga_calib_copy.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    ms = ctypes.c_double()
    rc = lib.ga_calib_copy(nbytes, 3, ctypes.byref(ms))
    if rc != 0:
        print(f"copy calibration failed ({rc})", file=sys.stderr)
        return 1
    print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": ms.value, "gbs": 2 * nbytes / ms.value / 1e6}))
    return 0


def longest_extensions(args):
    """One launch of the forward extensions of the N longest reads, each from its
    seed nearest the read's start: how long the longest extensions take when no
    other extension shares the GPU."""
    import numpy as np
    import graphaligner_amd as ga
    from graphaligner_amd import synth

    workdir = args.workdir or tempfile.mkdtemp(prefix="ga_bench_")
    os.makedirs(workdir, exist_ok=True)
    wl = WORKLOADS[args.workload]
    prefix = os.path.join(workdir, wl["preset"])
    if not (os.path.exists(prefix + ".gfa") and os.path.exists(prefix + ".fa")):
        synth.generate(wl["preset"], prefix)
    names, seqs = synth.read_fasta(prefix + ".fa")
    longest = sorted(range(len(seqs)), key=lambda i: -len(seqs[i]))[: args.longest_extensions]
    params = ga.params_preset("vg")
    params.minimizerLength, params.minimizerWindowSize = wl["k"], wl["w"]
    graph = ga.Graph(prefix + ".gfa")
    ctx = ga.Context(0)
    index = ga.Index(graph, params, ctx=ctx)
    jobs = []
    for i in longest:
        sd = ga.seeds(ctx, graph, index, [seqs[i]], params.minimizerSeedDensity)[0]
        first = sd[np.argmin(sd[:, 2])]
        jobs.append((seqs[i][int(first[2]) + 1:], int(first[0]) * 2 + int(first[4]), int(first[1])))
    got, ms = ga.extend_batch(ctx, graph, params, [j[0] for j in jobs], [j[1] for j in jobs], [j[2] for j in jobs], summaries=True)
    print(json.dumps({"longest_extensions": len(jobs), "kernel_ms": ms, "lengths": [len(j[0]) for j in jobs],
                      "slices": [len(g[3]) for g in got]}), flush=True)
    return 0


def main():
    args = parse()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.longest_extensions:
        return longest_extensions(args)
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path, and "
              "--gpus %d is not supported (run with --gpus 1)" % args.gpus, file=sys.stderr)
        return 2
    if args.learned:
        return learned_main(args)
    if args.workload == "dbg2m":
        print("--workload dbg2m is timed with --learned (plain and learned side by side)", file=sys.stderr)
        return 2
    import numpy as np
    import graphaligner_amd as ga
    from graphaligner_amd import synth

    workdir = args.workdir or tempfile.mkdtemp(prefix="ga_bench_")
    os.makedirs(workdir, exist_ok=True)
    wl = WORKLOADS[args.workload]
    loop = args.seed_loop if args.seed_loop != "auto" else "batched" if args.workload == "stream" else "rounds"
    prefix = os.path.join(workdir, wl["preset"])
    t = time.time()
    if not (os.path.exists(prefix + ".gfa") and os.path.exists(prefix + ".fa")):
        synth.generate(wl["preset"], prefix)
    names, seqs = synth.read_fasta(prefix + ".fa")
    if args.reads:
        names, seqs = names[: args.reads], seqs[: args.reads]
    gen_s = time.time() - t
    print(f"[bench] workload {args.workload}: {len(seqs)} reads, {sum(map(len, seqs)) / 1e6:.1f} Mbp ({gen_s:.1f} s)", file=sys.stderr, flush=True)

    params = ga.params_preset("vg")
    params.minimizerLength, params.minimizerWindowSize = wl["k"], wl["w"]
    params.deviceSeedLoop = {"rounds": 0, "device": 1, "batched": 2}[loop]
    params.deviceSeedOrder = 1 if args.seed_order == "device" else 0
    t = time.time()
    graph = ga.Graph(prefix + ".gfa")
    load_s = time.time() - t
    t = time.time()
    ctx = ga.Context(0)
    ctx_s = time.time() - t
    # the minimizer index is built on the GPU (ga_index_build_gpu); the host
    # build's time is reported beside it
    t = time.time()
    index = ga.Index(graph, params, ctx=ctx)
    index_s = time.time() - t
    t = time.time()
    ga.Index(graph, params)
    index_host_s = time.time() - t
    t = time.time()
    ga.upload_reads(ctx, seqs)
    ctx.synchronize()
    upload_s = time.time() - t
    print(f"[bench] graph {graph.num_nodes} split nodes ({load_s:.1f} s), index {index.num_kmers} k-mers ({index_s:.1f} s)",
          file=sys.stderr, flush=True)

    for w in range(args.warmup):
        n = min(len(names), args.warmup_reads) if args.warmup_reads else len(names)
        t = time.time()
        ga.align_resident(ctx, graph, index, params, 0, names[:n], gaf_format="buffer")
        print(f"[bench] warmup {w + 1}/{args.warmup}: {n} reads {time.time() - t:.1f} s", file=sys.stderr, flush=True)

    ctx.synchronize()
    t0 = time.perf_counter()
    tot = dict(extendKernelMs=0.0, lookupMs=0.0, hostMs=0.0, extendCallMs=0.0, lookupCallMs=0.0, totalMs=0.0,
               seedingMs=0.0, pickMs=0.0, replayMs=0.0, gafMs=0.0, tuneMs=0.0, unusedExtensions=0, launches=0, extensions=0,
               wordSteps=0, rounds=0, seeds=0, reruns=0)
    gaf = status = None
    for k in range(args.steps):
        ts = time.perf_counter()
        gaf, status, st = ga.align_resident(ctx, graph, index, params, 0, names, gaf_format="buffer")
        for key in tot:
            tot[key] += st[key]
        print(f"[bench] step {k + 1}/{args.steps}: {time.perf_counter() - ts:.2f} s, kernel {st['extendKernelMs'] / 1e3:.2f} s, "
              f"{st['launches']} launches, {st['extensions']} extensions, {st['reruns']} reruns", file=sys.stderr, flush=True)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0

    steps = max(args.steps, 1)
    reads_per_s = len(names) * args.steps / elapsed
    gcups = 64 * tot["wordSteps"] / elapsed / 1e9
    kernel_s = tot["extendKernelMs"] / 1e3
    launches = max(tot["launches"], 1)
    # the dominant kernel: seed extension. Achieved = algorithmic bytes per launch
    # over the launch's mean duration (HIP events on the launching stream).
    achieved_gbs = BYTES_PER_WORD_STEP * tot["wordSteps"] / kernel_s / 1e9 if kernel_s > 0 else 0.0
    valu_tops = OPS_PER_WORD_STEP * tot["wordSteps"] / kernel_s / 1e12 if kernel_s > 0 else 0.0
    salu_gips = SALU_PER_WORD_STEP * tot["wordSteps"] / kernel_s / 1e9 if kernel_s > 0 else 0.0
    default_run = loop == "rounds" and not args.reads and args.workload == "vg5m"
    counts = np.bincount(status, minlength=5).tolist() if status is not None else []

    cpu = None
    parity = None
    if args.cpu_sample > 0:
        cpu, parity = cpu_baseline(prefix, names, seqs, gaf.tobytes().decode(), status, args, wl)

    line = {
        "metric": "reads/s + GCUPS on 1 MI355X (plain GPU path)",
        "value": reads_per_s,
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "gcups": gcups,
        "config": {"workload": wl["text"],
                   "reads": len(names), "read_mbp": sum(map(len, seqs)) / 1e6, "split_nodes": graph.num_nodes,
                   "preset": "vg", "minimizer_k": wl["k"], "minimizer_w": wl["w"], "parallelism": "1 GPU",
                   "seed_loop": loop, "seed_order": args.seed_order, "persistent_workers": st["persistentWorkers"] if loop == "batched" else None},
        "breakdown_s_per_step": {"extend_kernel": kernel_s / steps, "extend_calls": tot["extendCallMs"] / 1e3 / steps,
                                 "seed_lookup_kernel": tot["lookupMs"] / 1e3 / steps,
                                 "seed_lookup_calls": tot["lookupCallMs"] / 1e3 / steps, "host": tot["hostMs"] / 1e3 / steps,
                                 "host_seeding": tot["seedingMs"] / 1e3 / steps, "host_pick": tot["pickMs"] / 1e3 / steps,
                                 "host_replay": tot["replayMs"] / 1e3 / steps, "host_gaf": tot["gafMs"] / 1e3 / steps,
                                 "align_call": tot["totalMs"] / 1e3 / steps, "autotune": tot["tuneMs"] / 1e3 / steps},
        "counters_per_step": {k: tot[k] // steps for k in ("launches", "extensions", "unusedExtensions", "wordSteps", "seeds", "reruns")},
        "read_status_counts": counts,
        # the bit-parallel DP is integer-op bound (42 int32-equivalent ops per
        # 64-cell word step); its HBM view is kept beside it
        "roofline": {"kernel": {"rounds": "extendSingleLaneKernel", "device": "readLoopSingleLaneKernel", "batched": "readLoopPersistentKernel"}[loop],
                     "bound": "valu-int32", "achieved": valu_tops,
                     "peak": VALU_INT32_PEAK_TOPS, "unit": "TOP/s", "frac": valu_tops / VALU_INT32_PEAK_TOPS,
                     "traffic": MEASURED_TRAFFIC_PER_LAUNCH if default_run else None,
                     "traffic_source": TRAFFIC_SOURCE, "mean_launch_ms": tot["extendKernelMs"] / launches,
                     "ops_per_word_step": OPS_PER_WORD_STEP,
                     "hbm": {"achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS,
                             "bytes_per_word_step": BYTES_PER_WORD_STEP},
                     # the ceiling that binds the single-lane kernel: scalar issue
                     "salu": {"achieved": salu_gips, "peak": SALU_PEAK_GIPS, "unit": "G instr/s", "frac": salu_gips / SALU_PEAK_GIPS,
                              "instr_per_word_step": SALU_PER_WORD_STEP, "source": SALU_SOURCE} if default_run else None},
        "cpu_baseline": cpu,
        "gaf_identical_to_cpu_sample": parity,
        "setup_s": {"generate": gen_s, "graph_load": load_s, "context": ctx_s, "index_build_gpu": index_s, "index_build_host": index_host_s,
                    "read_upload": upload_s},
        "pcie_inclusive_reads_per_s": len(names) * args.steps / (elapsed + args.steps * upload_s),
    }
    print(json.dumps(line), flush=True)
    return 0


def gaf_agreement(plain, learned):
    """(identical, plain lines, extra): the plain method's GAF lines the learned
    form reproduces byte for byte (matched per read), and its lines that are not
    among the plain ones."""
    def per_read(text):
        d = {}
        for ln in text.split("\n"):
            if ln:
                d.setdefault(ln.split("\t", 1)[0], []).append(ln)
        return d
    a, b = per_read(plain), per_read(learned)
    same = total = extra = 0
    for name in set(a) | set(b):
        pa, pb = list(a.get(name, [])), list(b.get(name, []))
        total += len(pa)
        for ln in pa:
            if ln in pb:
                pb.remove(ln)
                same += 1
        extra += len(pb)
    return same, total, extra


# the learned form scores a read's first 64 seed clusters and removes the rest
# unscored (in the outcome records of both generators no cluster after the
# first ever made a kept alignment)
MAX_SCORED_CLUSTERS = 64


def learned_main(args):
    """Train the proposal network on the GPU from outcome records of the plain
    method, then time the plain method and the learned form (clusters pruned
    before each read's seed loop) on the same resident reads."""
    import numpy as np
    import graphaligner_amd as ga
    from graphaligner_amd import synth

    workdir = args.workdir or tempfile.mkdtemp(prefix="ga_bench_")
    os.makedirs(workdir, exist_ok=True)
    ctx = ga.Context(0)
    # training data: the 5 Mbp variation-graph generator with fresh seeds
    # (graph 12, reads 2012), reads until the records reach --train-records
    t = time.time()
    tprefix = os.path.join(workdir, "learn_train")
    train_reads = max(2000, args.train_records // 16)
    if not os.path.exists(tprefix + ".fa"):
        synth.generate("vg5m", tprefix, overrides_graph=dict(seed=12), overrides_reads=dict(numReads=train_reads, seed=2012))
    tn, ts = synth.read_fasta(tprefix + ".fa")
    vp = ga.params_preset("vg")
    tg = ga.Graph(tprefix + ".gfa")
    ti = ga.Index(tg, vp)
    ga.upload_reads(ctx, ts)
    F, Y, R = ga.cluster_records(ctx, tg, ti, vp, 0, len(ts))
    rec_s = time.time() - t
    held = R >= int(0.9 * len(ts))  # the last 10 % of the reads: held out
    print(f"[bench] {len(F)} outcome records from {len(ts)} reads ({rec_s:.1f} s), "
          f"{Y[:, 0].mean():.4f} of clusters kept an alignment", file=sys.stderr, flush=True)
    net = ga.Mlp(ctx, seed=1)
    losses, train_ms = net.train(F[~held], Y[~held], epochs=4, batch=4096, lr=2e-3, seed=1)
    P = net.forward(F[held])
    auc = []
    for h in range(2):
        y = Y[held, h] > 0.5
        ranks = np.empty(len(y))
        ranks[np.argsort(P[:, h], kind="stable")] = np.arange(1, len(y) + 1)
        pos, neg = y.sum(), (~y).sum()
        auc.append(float((ranks[y].sum() - pos * (pos + 1) / 2) / (pos * neg)) if pos and neg else None)
    print(f"[bench] trained: loss {losses.tolist()}, {train_ms:.1f} ms, held-out AUC {auc}", file=sys.stderr, flush=True)
    # threshold: the largest of a grid whose held-out reads keep >= 99.9 % of
    # their plain GAF lines
    hn = [tn[i] for i in range(len(ts)) if i >= int(0.9 * len(ts))]
    first = int(0.9 * len(ts))
    hplain, _, _ = ga.align_resident(ctx, tg, ti, vp, first, hn)
    threshold = 0.0
    for thr in (0.01, 0.05, 0.1, 0.2, 0.3, 0.5, 0.7, 0.9):
        ga.set_proposal(ctx, net, threshold=thr, max_clusters=MAX_SCORED_CLUSTERS)
        text, _, _ = ga.align_resident(ctx, tg, ti, vp, first, hn)
        ga.set_proposal(ctx, None)
        same, total, _ = gaf_agreement(hplain, text)
        if same >= 0.999 * total:
            threshold = thr
    print(f"[bench] threshold {threshold}", file=sys.stderr, flush=True)
    del tg, ti

    wl = WORKLOADS[args.workload if args.workload in ("vg5m", "dbg2m") else "vg5m"]
    prefix = os.path.join(workdir, wl["preset"])
    if not (os.path.exists(prefix + ".gfa") and os.path.exists(prefix + ".fa")):
        synth.generate(wl["preset"], prefix)
    names, seqs = synth.read_fasta(prefix + ".fa")
    if args.reads:
        names, seqs = names[: args.reads], seqs[: args.reads]
    params = ga.params_preset(wl.get("params", "vg"))
    graph = ga.Graph(prefix + ".gfa")
    index = ga.Index(graph, params)
    ga.upload_reads(ctx, seqs)
    ctx.synchronize()

    def timed(learned):
        ga.set_proposal(ctx, net if learned else None, threshold=threshold, max_clusters=MAX_SCORED_CLUSTERS)
        for _ in range(args.warmup):
            ga.align_resident(ctx, graph, index, params, 0, names, gaf_format="buffer")
        ctx.synchronize()
        t0 = time.perf_counter()
        tot = dict(extendKernelMs=0.0, wordSteps=0, extensions=0, proposalMs=0.0, prunedClusters=0, proposalClusters=0)
        for _ in range(args.steps):
            gaf, st, s = ga.align_resident(ctx, graph, index, params, 0, names, gaf_format="buffer")
            for k in tot:
                tot[k] += s[k]
        ctx.synchronize()
        el = time.perf_counter() - t0
        ga.set_proposal(ctx, None)
        return gaf.tobytes().decode(), st, el, tot

    plain, pst, pel, ptot = timed(False)
    learned, lst, lel, ltot = timed(True)
    same, total, extra = gaf_agreement(plain, learned)
    steps = max(args.steps, 1)
    plain_rps = len(names) * args.steps / pel
    learned_rps = len(names) * args.steps / lel
    line = {
        "metric": "reads/s + GCUPS on 1 MI355X; learned vs plain at >=99.5% identical GAF",
        "value": learned_rps, "unit": "reads/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": lel * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 (DP), bf16 (network)", "data": "synthetic",
        "gcups": 64 * ltot["wordSteps"] / lel / 1e9,
        "config": {"workload": wl["text"], "reads": len(names), "preset": wl.get("params", "vg"), "parallelism": "1 GPU",
                   "seed_loop": "rounds", "form": "learned proposal (seed clusters scored and pruned before each read's loop)"},
        "plain_reads_per_s": plain_rps, "speedup_vs_plain": learned_rps / plain_rps,
        "gaf_identical_fraction": same / max(total, 1), "gaf_plain_lines": total, "gaf_extra_lines": extra,
        "status_identical": bool((pst == lst).all()),
        "threshold": threshold, "max_scored_clusters": MAX_SCORED_CLUSTERS,
        "per_step": {"plain_word_steps": ptot["wordSteps"] // steps, "learned_word_steps": ltot["wordSteps"] // steps,
                     "plain_extensions": ptot["extensions"] // steps, "learned_extensions": ltot["extensions"] // steps,
                     "clusters": ltot["proposalClusters"] // steps, "pruned_clusters": ltot["prunedClusters"] // steps,
                     "proposal_s": ltot["proposalMs"] / 1e3 / steps},
        "training": {"records": int((~held).sum()), "held_out_records": int(held.sum()), "reads": len(ts),
                     "generator": "vg5m graph seed 12, reads seed 2012", "epochs": 4, "batch": 4096, "lr": 2e-3,
                     "epoch_loss": losses.tolist(), "train_ms": train_ms, "held_out_auc": auc,
                     "kept_fraction": float(Y[:, 0].mean()), "records_s": rec_s},
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline(prefix, names, seqs, gaf, status, args, wl):
    """The oracle (C++ CPU restatement of the reference path, one read per thread
    from a shared queue as src/Aligner.cpp:777-780) on the first reads of the same
    workload, in both builds the reference can be compiled as: with its
    assertions (as it ships) and without (-DNDEBUG); each timed as the median of
    three runs, the faster one reported as the value. The GPU's GAF lines for
    those reads are checked against the asserting build's. The sample keeps the
    CPU leg to about 10-30 s of host work."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as o
    n = min(args.cpu_sample, len(names))
    threads = args.cpu_threads or cpu_share()
    sd = dict(o.PRESET_SEEDING["vg"], k=wl["k"], w=wl["w"])
    params = o.preset_params("vg")
    timed = {}
    for label, ndebug in (("asserting", False), ("ndebug", True)):
        fg = o.Graph(prefix + ".gfa", ndebug=ndebug)
        fi = o.Index(fg, sd["k"], sd["w"], sd["discard"])
        runs = sorted(o.align_reads(fg, fi, params, sd["density"], names[:n], seqs[:n], threads)[2] for _ in range(3))
        timed[label] = (runs[1], runs, fg.L.orc_last_word_steps())
    fast = min(timed, key=lambda k: timed[k][0])
    secs, runs, steps = timed[fast]
    og = o.Graph(prefix + ".gfa")
    oi = o.Index(og, sd["k"], sd["w"], sd["discard"])
    text, ostat, _ = o.align_reads(og, oi, params, sd["density"], names[:n], seqs[:n], threads)
    # GPU lines of the same reads: GAF text is read-ordered, so take the lines up
    # to the first read outside the sample
    gpu_lines = [ln for ln in gaf.split("\n") if ln]
    sample = set(names[:n])
    gpu_sample = [ln for ln in gpu_lines if ln.split("\t", 1)[0] in sample]
    cpu_lines = [ln for ln in text.split("\n") if ln]
    same = sum(1 for a, b in zip(cpu_lines, gpu_sample) if a == b)
    parity = {"reads": n, "lines_cpu": len(cpu_lines), "lines_gpu": len(gpu_sample), "identical_lines": same,
              "status_identical": bool((ostat == status[:n]).all())}
    cpu = {"value": n / secs, "unit": "reads/s", "cores": threads, "kind": "port",
           "gcups": 64 * steps / secs / 1e9,
           "host_nproc": os.cpu_count(),
           "variants": {k: {"reads_per_s": n / v[0], "gcups": 64 * v[2] / v[0] / 1e9, "median_s": v[0],
                            "runs_s": v[1]} for k, v in timed.items()},
           "sample": f"first {n} of the {len(names)} reads of the same workload (the CPU leg is bounded to ~10-30 s), "
                     f"oracle -O3 {'-DNDEBUG' if fast == 'ndebug' else 'with assertions'} (the faster build), "
                     f"{threads} threads (the host cores this process may use), median of 3 runs {secs:.2f} s"}
    return cpu, parity


if __name__ == "__main__":
    sys.exit(main())

#!/usr/bin/env python3
"""Benchmark of the MI355X pixel reconstruction chain on PU50-like events.

Workload: 1000 synthetic PU50 events (50 vertices x Poisson(35) tracks, 1/pT^2
spectrum above 0.5 GeV, ~50k digis and ~13k clusters per event), words
resident in HBM before the timed region.  One step = one pass of the whole
chain (raw words -> digis -> clusters -> rec-hits -> ntuplets with
broken-line fits -> vertices) over all events.  The batched path spreads the
events over S HIP streams and runs each stream's share as equal launches of at
least B events per stage; the plain path runs one launch sequence per event
with the reference's per-event host synchronisation.  Prints one JSON line.

Phases: autotune, warmup, the timed steps (`value`, no instrumentation), the
same steps again with a HIP event pair around every kernel on its own stream
(per-kernel durations under overlap), the first stream's launches alone
(durations without overlap), a parity check of one event per launch against
the plain path, the plain path, the serving loop and the CPU baseline.
Each phase is a ROCTx range, so that a `rocprofv3 --kernel-trace
--marker-trace` run of this script can be summarised per phase with
`python bench.py --summarize-trace DIR` (writes the profiles' markdown).
"""
import argparse
import concurrent.futures as cf
import csv
import ctypes
import glob
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "pixeltrack-standalone_amd"))

METRIC = "events/s at PU50-like load; p50/p99 latency, batched vs plain per-event GPU"
STAGES = {1: "raw->digis", 2: "raw->clusters", 3: "raw->rec-hits", 4: "raw->ntuplets+fits",
          5: "raw->ntuplets+fits->vertices"}
HBM_PEAK_GBS = 8000.0
# host threads of the GPU box's CPU share for one GPU (the rest of the host's cores
# belong to the other GPUs' jobs)
CPU_SHARE = 16

# Canonical algorithmic bytes per event of each stage: the unique input a stage must
# read and the output it must write, with the reference's element widths, from the
# per-event counters W raw words, D digis, H hits, C doublets, L accepted cell links,
# T ntuplets, N hits per ntuplet, T' tracks given to the vertexer, V vertices.  CA
# records count at a minimal 8 bytes per cell, whatever this library's layout is.
# Each row lists the profiler labels of the kernels that implement it.
ROWS = {
    "K1-K3 decode+calib+countModules": (("decode",), lambda c: 4.5 * c["W"] + 16 * c["W"] + 12 * c["D"] + 6 * c["D"]),
    "K6 module scan (ordered starts)": (("moduleStarts",), lambda c: 16384.0),
    "K4+K5 findClus+chargeCut (+K6 clamp scan)": (("clusterize",), lambda c: 10 * c["D"] + 14 * c["D"] + 16384),
    "K7 getHits": (("hits",), lambda c: 12 * c["D"] + 44 * c["H"]),
    "K8 layer start + phi binner": (("phiBinner",), lambda c: 6 * c["H"] + 5120),
    "K9 doublets": (("doubletCount", "doubletScan", "doubletFill", "outerHitScan", "outerHitFill"),
                    lambda c: 14 * c["H"] + 12 * c["C"]),
    "K10 connect": (("connect",), lambda c: 8 * c["C"] + 16 * c["H"] + 4 * c["L"]),
    "K11 fishbone": (("fishbone",), lambda c: 12 * c["C"] + 16 * c["H"]),
    "K12 find_ntuplets": (("startMask", "startCompact", "ntupletCount", "ntupletScan", "ntupletFill"),
                          lambda c: 8 * c["C"] + 4 * c["L"] + 14 * c["T"]),
    "K13 cleaners/multiplicity/classify": (("trackListAlloc", "trackListFill", "trackListSort", "multiplicity",
                                            "classify", "dupPreds", "fastDuplicate", "hitToTupleCount",
                                            "hitToTupleScan", "hitToTupleFill", "tripletCleaner"),
                                           lambda c: 8 * c["C"] + 300 * c["T"]),
    "K14 broken-line fits": (("fit",), lambda c: 22 * c["Nbar"] * c["T"] + 92 * c["T"]),
    "K15 vertexer": (("vertices",), lambda c: 30 * c["T"] + 60 * c["Tp"] + 16 * c["V"]),
}
LABEL_ROW = {lab: row for row, (labs, _) in ROWS.items() for lab in labs}
# device kernel symbol -> profiler label (several kernels make up "clusterize")
KERNEL_LABEL = {
    "decodeKernel": "decode", "moduleStartKernel": "moduleStarts", "clusterSmallKernel": "clusterize",
    "clusterMediumKernel": "clusterize", "clusterBigKernel": "clusterize", "clusStartKernel": "clusterize",
    "hitsKernel": "hits",
    "phiBinnerKernel": "phiBinner", "doubletCountKernel": "doubletCount", "doubletScanKernel": "doubletScan",
    "doubletFillKernel": "doubletFill", "ohScanKernel": "outerHitScan", "ohFillKernel": "outerHitFill",
    "ohBuildKernel": "outerHitFill",
    "connectKernel": "connect", "fishboneKernel": "fishbone", "startMaskKernel": "startMask",
    "startCompactKernel": "startCompact", "ntupletCountKernel": "ntupletCount", "ntupletScanKernel": "ntupletScan",
    "ntupletFillKernel": "ntupletFill", "trackListAllocKernel": "trackListAlloc",
    "trackListFillKernel": "trackListFill", "trackListBuildKernel": "trackListFill",
    "trackListSortKernel": "trackListSort",
    "multiplicityKernel": "multiplicity", "fitKernel": "fit", "classifyKernel": "classify",
    "dupPredsKernel": "dupPreds", "fastDuplicateKernel": "fastDuplicate", "hitToTupleCountKernel": "hitToTupleCount",
    "hitToTupleScanKernel": "hitToTupleScan", "hitToTupleFillKernel": "hitToTupleFill",
    "hitToTupleBuildKernel": "hitToTupleFill",
    "tripletCleanerKernel": "tripletCleaner", "vertexKernel": "vertices",
}


def row_bytes(row, c):
    return ROWS[row][1](c)


def kernel_symbol(name):
    """Short name of a profiled kernel: 'void pxg::(anonymous namespace)::fitKernel<3>(...)'
    gives 'fitKernel<3>'."""
    name = name.replace("(anonymous namespace)", "anon").replace("void ", "")
    return name.split("(")[0].split("::")[-1].strip()


def kernel_label(sym):
    return KERNEL_LABEL.get(sym.split("<")[0])


class Roctx:
    """ROCTx ranges around the phases (no-ops when the library is absent)."""

    def __init__(self):
        self.lib = None
        # rocprofv3 --marker-trace intercepts the rocprofiler-sdk ROCTx library (the older
        # roctracer one is not seen by it)
        for name in ("librocprofiler-sdk-roctx.so", "/opt/rocm/lib/librocprofiler-sdk-roctx.so", "libroctx64.so",
                     "/opt/rocm/lib/libroctx64.so"):
            try:
                self.lib = ctypes.CDLL(name)
                self.lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
                break
            except OSError:
                continue

    def push(self, name):
        if self.lib:
            self.lib.roctxRangePushA(name.encode())

    def pop(self):
        if self.lib:
            self.lib.roctxRangePop()


def gen_events(detector, cfg, n, threads, first=0):
    with cf.ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda i: detector.event(cfg, i)[0], range(first, first + n)))


def serve_at(ctx, inputs, rate, server_kw, seed=1, classes=None, pinned=None):
    """Submits `inputs` with Poisson arrivals at `rate` events/s (all at once if rate is None)
    to a serving loop and returns throughput and latency percentiles (ms).  With `pinned`
    (the same events in mapped pinned buffers) the events are handed over without a host
    copy (pxg_server_submit_pinned), else copied at submission (pxg_server_submit)."""
    from pxtrack import gpu
    sv = gpu.Server(ctx, **server_kw)
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(inputs)) if rate else np.zeros(len(inputs))
    ids = {}
    t0 = time.perf_counter()
    t = t0
    for i, (w, f) in enumerate(inputs):
        t += gaps[i]
        while True:  # spin until the arrival time (sleep when far)
            d = t - time.perf_counter()
            if d <= 0:
                break
            if d > 2e-4:
                time.sleep(d - 1e-4)
        ids[sv.submit_pinned(pinned[i]) if pinned is not None else sv.submit(w, f)] = i
    res = []
    deadline = time.time() + 120
    while len(res) < len(ids) and time.time() < deadline:
        res += sv.poll(wait_ms=50)
    info = sv.info()
    sv.close()
    if len(res) < len(ids):
        return {"rate": rate, "error": f"{len(res)} of {len(ids)} completed"}
    lat = np.array([(r["t_done_ns"] - r["t_submit_ns"]) * 1e-6 for r in res])
    span = (max(r["t_done_ns"] for r in res) - min(r["t_submit_ns"] for r in res)) * 1e-9
    out = {"rate": rate, "submit": "pinned" if pinned is not None else "copy", "events": len(res),
           "events_per_s": round(len(res) / span, 1),
           "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "p999_ms": round(float(np.percentile(lat, 99.9)), 3),
           "mean_batch": round(float(np.mean([r["batch_size"] for r in res])), 1),
           "mean_queue_ms": round(float(np.mean([(r["t_launch_ns"] - r["t_submit_ns"]) * 1e-6 for r in res])), 3),
           "mean_batch_ms": round(float(np.mean([(r["t_done_ns"] - r["t_launch_ns"]) * 1e-6 for r in res])), 3),
           "batches": info["batches"], "status_or": int(np.bitwise_or.reduce([r["status"] for r in res]))}
    if classes is not None:
        for name in sorted(set(classes)):
            lc = np.array([(r["t_done_ns"] - r["t_submit_ns"]) * 1e-6 for r in res if classes[ids[r["id"]]] == name])
            out[f"{name}_p50_ms"] = round(float(np.percentile(lc, 50)), 3)
            out[f"{name}_p99_ms"] = round(float(np.percentile(lc, 99)), 3)
    return out


def serving_sweep(ctx, det, cond, args, tuned):
    """Continuous-batching serving loop vs the plain per-event path at the same arrival rates."""
    from pxtrack import gen, gpu
    threads = min(CPU_SHARE, os.cpu_count() or 1)
    raws = gen_events(det, gen.preset("pu_poisson60"), args.serve_events, threads)
    ev = [gpu.unpack_raw(r, cond["fed_ids"]) for r in raws]
    budget = int(args.serve_budget_gb * 1e9)
    shape = dict(doublet_lanes=tuned["doublet_lanes"], grid_target=tuned["grid_target"]) if tuned else {}
    # products stay on the device (no --transfer, the reference's default); counters come back
    batched_kw = dict(max_batch=args.batch, n_streams=3, hbm_budget=budget, upto=args.upto, transfer=False, **shape)
    plain_kw = dict(max_batch=1, n_streams=1, upto=args.upto, transfer=False, **shape)
    rates = [1000, 3000, 10000, 30000, None]
    # the events as a front end holding them in mapped pinned buffers hands them over: the
    # device reads them while it loads each batch (PCIe-inclusive, no host copy)
    pev = [gpu.PinnedEvent(w, f) for w, f in ev]
    out = {"workload": f"{len(ev)} events, pileup Poisson(60) clipped to [10, 200], Poisson arrivals",
           "hbm_budget_gb": args.serve_budget_gb, "batched": [], "plain": []}
    for r in rates:
        out["batched"].append(serve_at(ctx, ev, r, batched_kw, pinned=pev))
        print(f"[bench] serving batched rate {r}: {out['batched'][-1]}", file=sys.stderr, flush=True)
    # the same at saturation with a host copy of every event at submission
    out["batched_copy_submit"] = serve_at(ctx, ev, None, batched_kw)
    print(f"[bench] serving batched, copied at submission: {out['batched_copy_submit']}", file=sys.stderr,
          flush=True)
    for pe in pev:
        pe.close()
    for r in rates[:2] + [None]:
        sub = ev[: min(len(ev), 300)]
        out["plain"].append(serve_at(ctx, sub, r, plain_kw))
        print(f"[bench] serving plain rate {r}: {out['plain'][-1]}", file=sys.stderr, flush=True)
    # 80/20 mix of PU50 and PU200 with the high-occupancy capacities, latency per class
    n50 = int(0.8 * args.mix_events)
    r50 = gen_events(det, gen.preset("pu50"), n50, threads)
    r200 = gen_events(det, gen.preset("pu200"), args.mix_events - n50, threads)
    mix = [(gpu.unpack_raw(r, cond["fed_ids"]), "pu50") for r in r50] + \
          [(gpu.unpack_raw(r, cond["fed_ids"]), "pu200") for r in r200]
    order = np.random.default_rng(7).permutation(len(mix))
    mix = [mix[i] for i in order]
    pmix = [gpu.PinnedEvent(*m[0]) for m in mix]
    mix_kw = dict(max_batch=args.batch // 2, n_streams=3, hbm_budget=budget, upto=args.upto,
                  profile=gpu.PROFILE_HIGH_OCCUPANCY, transfer=False, **shape)
    out["mix_pu50_pu200"] = [serve_at(ctx, [m[0] for m in mix], r, mix_kw, classes=[m[1] for m in mix],
                                      pinned=pmix) for r in (2000, None)]
    for pe in pmix:
        pe.close()
    print(f"[bench] serving mix: {out['mix_pu50_pu200']}", file=sys.stderr, flush=True)
    return out


def roofline_of(label, prof, counters, steps_events):
    """Roofline entry of one profiled kernel label: its row's algorithmic bytes per event x
    events per launch / average launch duration."""
    calls, total_ms = prof[label]
    row = LABEL_ROW[label]
    per_launch = steps_events / calls
    alg = row_bytes(row, counters) * per_launch
    avg_s = total_ms / 1e3 / calls
    achieved = alg / avg_s / 1e9
    return {"bound": "hbm", "kernel": label, "row": row, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "algorithmic_bytes_per_event": round(row_bytes(row, counters)),
            "algorithmic_bytes_per_launch": round(alg), "avg_launch_ms": round(avg_s * 1e3, 4), "launches": calls,
            "events_per_launch": round(per_launch, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--events", type=int, default=1000)
    ap.add_argument("--batch", type=int, default=128,
                    help="at least this many events per launch (each stream's share is cut into equal launches)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tuning", metavar="LANES,GRID[,PER_EVENT,STAGE,WINDOW]",
                    help="launch shape (doublet lanes per slot, workgroups per launch, at most per event, connect's "
                         "staged candidates, fishbone window) instead of the autotuner's")
    ap.add_argument("--streams", type=int, default=3, help="HIP streams the batches are spread over")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop sweep")
    ap.add_argument("--serve-events", type=int, default=2000)
    ap.add_argument("--mix-events", type=int, default=200)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--upto", type=int, default=5, help="last stage: 3 rec-hits, 4 ntuplets + fits, 5 vertices")
    ap.add_argument("--graphs", action="store_true",
                    help="replay each batch's stage sequence as a captured HIP graph (batched path)")
    ap.add_argument("--riemann-fit", action="store_true",
                    help="fit with the Riemann fit instead of the broken line (batched path)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB three times (the reference kernel for scaling "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--isolated-passes", type=int, default=2,
                    help="passes over the first stream's batches with the other streams idle, for the "
                         "per-kernel durations without overlap")
    ap.add_argument("--only-timed", action="store_true",
                    help="stop after the timed steps (for counter runs: every chain dispatch is a timed one)")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="summarise a rocprofv3 --kernel-trace --marker-trace csv output directory of this "
                         "script per phase (markdown on stdout) and exit")
    ap.add_argument("--summarize-sq", metavar="DIR",
                    help="summarise a rocprofv3 --pmc csv run of this script with SQ_WAVE_CYCLES, "
                         "SQ_LDS_BANK_CONFLICT, SQ_ACTIVE_INST_VALU, SQ_ACTIVE_INST_LDS and GRBM_GUI_ACTIVE into "
                         "occupancy, LDS bank-conflict and VALU-busy percentages per kernel (markdown) and exit")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc csv output directories (FETCH_SIZE, WRITE_SIZE) of an "
                         "--only-timed --calibrate-copy run of this script into HBM bytes per event and kernel "
                         "(json on stdout, markdown on stderr) and exit; needs --events and --steps of that run")
    args = ap.parse_args()
    if args.summarize_trace:
        summarize_trace(args.summarize_trace)
        return
    if args.summarize_sq:
        summarize_sq(args.summarize_sq)
        return
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1], args.events * args.steps, args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: each box here has one MI355X and events shard across GPUs "
              f"without any exchange, so --gpus {args.gpus} is not run.", file=sys.stderr)
        sys.exit(2)

    from pxtrack import gen, gpu
    import torch

    rtx = Roctx()
    if args.calibrate_copy > 0:
        nbytes = args.calibrate_copy << 20
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: 3 copies of {nbytes} bytes", file=sys.stderr, flush=True)
        del a, b

    t0 = time.time()
    det = gen.Detector()
    cond = det.conditions()
    cfg = gen.preset("pu50")
    threads = min(CPU_SHARE, os.cpu_count() or 1)
    raws = gen_events(det, cfg, args.events, threads)
    inputs = [gpu.unpack_raw(r, cond["fed_ids"]) for r in raws]
    W = np.array([len(w) for w, _ in inputs], dtype=np.int64)
    print(f"[bench] generated {len(raws)} PU50 events in {time.time() - t0:.1f}s, mean words {W.mean():.0f}",
          file=sys.stderr, flush=True)

    ctx = gpu.Context(cond)
    n = len(inputs)
    B = max(1, min(args.batch, n))
    # the events are split over S HIP streams, each with its own resident pool;
    # a pass enqueues every stream's batches, which then overlap on the device
    S = max(1, min(args.streams, (n + B - 1) // B))
    bounds = [round(k * n / S) for k in range(S + 1)]
    # each stream's share is cut into equal launches of at least --batch events: a
    # short last launch would cost a whole launch's latency for a few events
    share = max(bounds[k + 1] - bounds[k] for k in range(S))
    per = max(1, share // B)
    B = (share + per - 1) // per
    streams = []
    for k in range(S):
        part = inputs[bounds[k]:bounds[k + 1]]
        st = ctx.stream(max_events=len(part), max_words=int(sum(len(w) for w, _ in part)))
        if args.riemann_fit:
            st.set_riemann_fit(True)
        if args.graphs:
            st.set_graphs(True)
        st.load(part)
        streams.append(st)
    stream = streams[0]
    launches = [(k, first, min(B, bounds[k + 1] - bounds[k] - first)) for k in range(S)
                for first in range(0, bounds[k + 1] - bounds[k], B)]

    def batched_pass():
        for k, first, count in launches:
            streams[k].run(first, count, args.upto)
        status = 0
        for k, st in enumerate(streams):
            status |= st.fetch_counts(0, bounds[k + 1] - bounds[k])
        return status

    # ---- launch shape chosen by the autotuner on the first batch
    tuned = None
    if args.tuning:
        v = [int(x) for x in args.tuning.split(",")]
        keys = ("doublet_lanes", "grid_target", "max_blocks_per_event", "connect_stage", "fishbone_window")
        shape = dict(zip(keys, v))
        for st in streams:
            st.set_launch_shape(shape)
        tuned = dict(stream.launch_shape(), ms_per_batch=None)
    elif not args.no_tune:
        rtx.push("autotune")
        shape, tms = stream.autotune_shape(0, min(B, bounds[1]), args.upto)
        rtx.pop()
        for st in streams[1:]:
            st.set_launch_shape(shape)
        tuned = dict(shape, ms_per_batch=round(tms, 4))
        print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)

    # ---- batched path: warmup, then K timed steps (one step = all events), nothing instrumented
    rtx.push("warmup")
    for _ in range(args.warmup):
        batched_pass()
    torch.cuda.synchronize()
    rtx.pop()
    rtx.push("timed")
    t_start = time.perf_counter()
    for _ in range(args.steps):
        status = batched_pass()
    for st in streams:
        st.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    rtx.pop()
    ms_per_step = 1e3 * elapsed / args.steps
    evps = n * args.steps / elapsed
    print(f"[bench] timed: {evps:.1f} events/s, {ms_per_step:.3f} ms per step, status {status}", file=sys.stderr,
          flush=True)
    if args.only_timed:
        print(json.dumps({"metric": METRIC, "value": round(evps, 1), "unit": "events/s", "only_timed": True,
                          "events": n, "steps": args.steps, "status": status}))
        return

    # ---- the same steps with every kernel bracketed by HIP events on its stream
    for st in streams:
        st.profile(True)
        st.profile_read()
    rtx.push("profiled")
    t_prof = time.perf_counter()
    for _ in range(args.steps):
        batched_pass()
    for st in streams:
        st.sync()
    elapsed_prof = time.perf_counter() - t_prof
    rtx.pop()
    prof = {}
    for st in streams:
        for k, (c, ms) in st.profile_read().items():
            a = prof.get(k, (0, 0.0))
            prof[k] = (a[0] + c, a[1] + ms)
        st.profile(False)

    # ---- the first stream's batches alone (no overlap with the other streams)
    iso = {}
    iso_events = 0
    if args.isolated_passes > 0:
        st = streams[0]
        st.profile(True)
        st.profile_read()
        rtx.push("isolated")
        for _ in range(args.isolated_passes):
            for k, first, count in launches:
                if k == 0:
                    st.run(first, count, args.upto)
        st.sync()
        rtx.pop()
        iso = st.profile_read()
        st.profile(False)
        iso_events = bounds[1] * args.isolated_passes

    # ---- parity: one event of every launch of the last pass against the plain path
    parity = None
    if args.upto >= 5:
        rtx.push("parity")
        plain_stream = ctx.stream(max_events=1, max_words=int(W.max()))
        checked, bad = 0, []
        for k, first, count in launches:
            e = first + count // 2
            got = streams[k].event_products(e, upto=5)
            alone = plain_stream.process(*inputs[bounds[k] + e], upto=5)
            diff = gpu.differing_products(alone, got)
            checked += 1
            if diff:
                bad.append({"stream": k, "event": e, "differ": diff})
        plain_stream.close()
        rtx.pop()
        parity = {"events_checked": checked, "identical_to_plain": not bad, "mismatches": bad}
        if bad:
            print(f"[bench] PARITY FAILURE: batched products differ from the plain path: {bad}", file=sys.stderr,
                  flush=True)
            print(json.dumps({"metric": METRIC, "error": "batched products differ from the plain path",
                              "parity": parity}))
            sys.exit(3)

    # per-event counters for the algorithmic bytes
    where = [(st, e) for k, st in enumerate(streams) for e in range(bounds[k + 1] - bounds[k])]
    cnts = np.array([st.event_counts(e) for st, e in where], dtype=np.float64)
    mean = {"W": float(W.mean()), "D": float(W.mean()), "M": float(cnts[:, 0].mean()), "H": float(cnts[:, 1].mean())}
    for key in ("C", "L", "T", "TH", "F", "Tp", "V", "Nbar"):
        mean[key] = 0.0
    if args.upto >= 4:
        cac = [st.ca_counts(e) for st, e in where]
        for key, name in (("C", "nCells"), ("L", "nNeighbors"), ("T", "nTuples"), ("TH", "apcN"), ("F", "nMult"),
                          ("Tp", "vtxTracks"), ("V", "nVertices")):
            mean[key] = float(np.mean([c[name] for c in cac]))
        mean["Nbar"] = mean["TH"] / mean["T"] if mean["T"] else 0.0

    # roofline: of the kernels that make up a scope row on their own (so that the row's
    # algorithmic bytes are that kernel's), the one with the largest share of the profiled
    # timed steps; the track-list build shares its row with the cleaners and is not a candidate
    steps_events = n * args.steps
    single = [lab for lab in prof if lab in LABEL_ROW and len(ROWS[LABEL_ROW[lab]][0]) == 1]
    total_prof = sum(v[1] for v in prof.values())
    share = {k: round(v[1] / total_prof, 4) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])}
    kern = max(single, key=lambda k: prof[k][1])
    tfile = os.path.join(REPO, "profiles", "traffic_chain_pu50_batched167.json")
    traffic_per_event = json.load(open(tfile))["labels"] if os.path.exists(tfile) else {}

    def with_extras(label):
        r = roofline_of(label, prof, mean, steps_events)
        r["traffic"] = (round(traffic_per_event[label] * r["events_per_launch"]) if label in traffic_per_event
                        else None)
        if label in iso:
            ic, ims = iso[label]
            per_launch = iso_events / ic
            alg = row_bytes(LABEL_ROW[label], mean) * per_launch
            r["isolated"] = {"avg_launch_ms": round(ims / ic, 4), "launches": ic,
                             "events_per_launch": round(per_launch, 1),
                             "achieved": round(alg / (ims / ic / 1e3) / 1e9, 2),
                             "frac": round(alg / (ims / ic / 1e3) / 1e9 / HBM_PEAK_GBS, 5)}
        return r

    roof = with_extras(kern)
    others = {lab: with_extras(lab) for lab in ("connect", "fit") if lab in prof and lab != kern}
    # every row of the canonical table: summed kernel time per launch of the row's kernels
    rows = {}
    for row, (labs, _) in ROWS.items():
        have = [lab for lab in labs if lab in prof]
        if not have:
            continue
        per_launch_ms = sum(prof[lab][1] / prof[lab][0] for lab in have)
        per_launch_events = steps_events / prof[have[0]][0]
        alg = row_bytes(row, mean) * per_launch_events
        rows[row] = {"ms_per_launch": round(per_launch_ms, 4), "bytes_per_event": round(row_bytes(row, mean)),
                     "frac": round(alg / (per_launch_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 5)}
    chain_bytes = sum(row_bytes(row, mean) for row in ROWS) * n
    chain = {"algorithmic_bytes_per_step": round(chain_bytes), "frac_of_hbm_peak_per_step":
             round(chain_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 5)}

    # ---- plain per-event path: one launch sequence + host sync per event
    plain = None
    if not args.no_plain:
        rtx.push("plain")
        lat = []
        for st, e in where[:50]:  # warmup
            st.run(e, 1, args.upto)
            st.fetch_counts(e, 1)
        t1 = time.perf_counter()
        for st, e in where:
            a = time.perf_counter()
            st.run(e, 1, args.upto)
            st.fetch_counts(e, 1)
            lat.append(time.perf_counter() - a)
        tp = time.perf_counter() - t1
        rtx.pop()
        lat = np.array(lat) * 1e3
        plain = {"events_per_s": round(n / tp, 1), "latency_ms_p50": round(float(np.percentile(lat, 50)), 4),
                 "latency_ms_p99": round(float(np.percentile(lat, 99)), 4),
                 "latency_ms_p999": round(float(np.percentile(lat, 99.9)), 4)}

    # ---- serving loop: continuous batching under Poisson arrivals
    serving = None
    if not args.no_serve:
        rtx.push("serving")
        serving = serving_sweep(ctx, det, cond, args, tuned)
        rtx.pop()

    # ---- CPU baseline: the oracle (serial restatement) on the host cores, bounded sample
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import pyoracle
        orc = pyoracle.Oracle(cond)
        sample = raws[:100]
        ncores = min(CPU_SHARE, os.cpu_count() or 1)
        a = time.perf_counter()
        done = orc.run_many(sample, upto=args.upto, threads=ncores, repeat=1)
        one = time.perf_counter() - a
        rep = max(1, int(args.cpu_seconds / max(one, 1e-3)))
        a = time.perf_counter()
        done = orc.run_many(sample, upto=args.upto, threads=ncores, repeat=rep)
        ct = time.perf_counter() - a
        # one thread, one event at a time: the serial program's per-event latency
        a = time.perf_counter()
        one_done = orc.run_many(sample[:20], upto=args.upto, threads=1, repeat=1)
        one_ms = 1e3 * (time.perf_counter() - a) / max(1, one_done)
        model = ""
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": round(done / ct, 1), "unit": "events/s", "cores": ncores, "kind": "port",
               "sample": f"{len(sample)} of the PU50 events x {rep} passes, {STAGES[args.upto]}, one event per thread "
                         f"at a time ({ncores} threads), {ct:.1f} s",
               "host_cpu_count": os.cpu_count(),
               "cores_note": f"{CPU_SHARE} threads: the host CPU share of one GPU on the box",
               "single_thread_ms_per_event": round(one_ms, 3), "cpu_model": model}

    batch_lat_ms = ms_per_step / len(launches)
    kernel_ms = {k: round(v[1] / v[0], 4) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])}
    kernel_ms_iso = {k: round(v[1] / v[0], 4) for k, v in sorted(iso.items(), key=lambda kv: -kv[1][1])}
    out = {
        "metric": METRIC,
        "value": round(evps, 1),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32/f32/f64",
        "data": "synthetic",
        "config": {"workload": f"{n} PU50 events ({STAGES[args.upto]}), batched path, {B} events per launch, "
                               f"{S} streams" + (", Riemann fit" if args.riemann_fit else ""),
                   "fit": "riemann" if args.riemann_fit else "broken line",
                   "graphs": bool(args.graphs),
                   "events": n, "batch": B, "streams": S, "mean_words": round(float(W.mean()), 1),
                   "mean_per_event": {k: round(v, 2) for k, v in mean.items()},
                   "stages": STAGES[args.upto]},
        "batched": {"events_per_s": round(evps, 1), "batch_latency_ms": round(batch_lat_ms, 4),
                    "status": status, "tuning": tuned,
                    "profiled_pass_events_per_s": round(steps_events / elapsed_prof, 1)},
        "parity": parity,
        "plain": plain,
        "speedup_batched_vs_plain": round(evps / plain["events_per_s"], 2) if plain else None,
        "serving": serving,
        "kernel_ms_per_launch": kernel_ms,
        "kernel_ms_per_launch_isolated": kernel_ms_iso,
        "kernel_time_share": share,
        "roofline": roof,
        "roofline_other_kernels": others,
        "roofline_rows": rows,
        "chain": chain,
        "cpu_baseline": cpu,
        "device": gpu.device_name(),
    }
    print(json.dumps(out))


# ---- rocprofv3 trace summary ------------------------------------------------------------

def _read_csv(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def summarize_trace(d):
    """Markdown summary of a `rocprofv3 --kernel-trace --marker-trace -f csv` run of this
    script: per kernel symbol, the launches and average duration in each phase (the ROCTx
    ranges this script opens; every phase ends with a device synchronisation, so each
    dispatch belongs to the phase it started in)."""
    kfile = sorted(glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True))
    mfile = sorted(glob.glob(os.path.join(d, "**", "*marker_api_trace.csv"), recursive=True))
    if not kfile:
        sys.exit(f"no kernel_trace.csv under {d}")
    ks = _read_csv(kfile[0])
    phases = []
    if mfile:
        for m in _read_csv(mfile[0]):
            name = m.get("Function") or m.get("Marker_Name") or m.get("Name") or ""
            if name in ("autotune", "warmup", "timed", "profiled", "isolated", "parity", "plain", "serving"):
                phases.append((int(m["Start_Timestamp"]), int(m["End_Timestamp"]), name))
    order = ["timed", "profiled", "isolated", "warmup", "autotune", "parity", "plain", "serving", "other"]
    stats = {}
    for k in ks:
        sym = kernel_symbol(k["Kernel_Name"])
        t0, t1 = int(k["Start_Timestamp"]), int(k["End_Timestamp"])
        ph = next((p for a, b, p in phases if a <= t0 <= b), "other")
        s = stats.setdefault(sym, {})
        c, tot = s.get(ph, (0, 0.0))
        s[ph] = (c + 1, tot + (t1 - t0) / 1e3)
    tot_timed = sum(s.get("timed", (0, 0.0))[1] for s in stats.values())
    print("| kernel | label | timed launches | timed avg us | timed share | profiled avg us | isolated avg us |")
    print("|---|---|---|---|---|---|---|")
    iso_sum = 0.0
    for sym, s in sorted(stats.items(), key=lambda kv: -kv[1].get("timed", (0, 0.0))[1]):
        if not any(ph in s for ph in ("timed", "profiled", "isolated")):
            continue  # ran only while the launch shape was tuned
        ci, ti = s.get("isolated", (0, 0.0))
        iso_sum += ti / ci if ci else 0.0
        def avg(ph):
            c, t = s.get(ph, (0, 0.0))
            return f"{t / c:.2f}" if c else ""
        c, t = s.get("timed", (0, 0.0))
        share = f"{100 * t / tot_timed:.2f} %" if tot_timed and c else ""
        print(f"| {sym} | {kernel_label(sym) or ''} | {c} | {avg('timed')} | {share} | {avg('profiled')} | "
              f"{avg('isolated')} |")
    print(f"\nSum of the isolated averages (one launch of each kernel, nothing overlapping): "
          f"{iso_sum:.1f} us")
    if phases:
        print()
        for a, b, p in phases:
            print(f"- phase `{p}`: {(b - a) / 1e6:.2f} ms")
    else:
        print("\n(no ROCTx phases in the trace: every dispatch counted as `other`)")


def _pmc_per_kernel(d, counter):
    """Sum of one counter over the dispatches of each kernel symbol in a --pmc csv run."""
    f = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))
    if not f:
        sys.exit(f"no counter_collection.csv under {d}")
    out = {}
    for r in _read_csv(f[0]):
        if r.get("Counter_Name") != counter:
            continue
        sym = kernel_symbol(r["Kernel_Name"])
        c, v, m = out.get(sym, (0, 0.0, 0.0))
        x = float(r["Counter_Value"])
        out[sym] = (c + 1, v + x, max(m, x))
    return out


def summarize_sq(d, cu_num=256, waves_per_cu=32, xcds=8):
    """Per kernel, from the dispatches of a --pmc run (rocprofv3's derived-metric formulas):
    occupancy % = 400 SQ_WAVE_CYCLES / GRBM_GUI_ACTIVE / CUs / max waves per CU (SQ_WAVE_CYCLES
    counts in quad cycles), LDS bank-conflict % = 100 SQ_LDS_BANK_CONFLICT / GRBM_GUI_ACTIVE /
    CUs, VALU busy % = 100 SQ_ACTIVE_INST_VALU / CUs / GRBM_GUI_ACTIVE, LDS busy % likewise with
    SQ_ACTIVE_INST_LDS; each summed over the kernel's dispatches.  On this MI355X the SQ counters
    of a dispatch cover one of the 8 XCDs against the whole GPU's GRBM_GUI_ACTIVE: a copy kernel
    of 2048 waves (at most 8 waves per CU, 25 % occupancy) reads 2.8 % by the plain formula,
    so every ratio is multiplied by the XCD count."""
    f = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))
    if not f:
        sys.exit(f"no counter_collection.csv under {d}")
    acc = {}
    for r in _read_csv(f[0]):
        sym = kernel_symbol(r["Kernel_Name"])
        a = acc.setdefault(sym, {"dispatches": set()})
        a["dispatches"].add(r.get("Dispatch_Id") or r.get("Correlation_Id"))
        name = r["Counter_Name"]
        a[name] = a.get(name, 0.0) + float(r["Counter_Value"])
    print("| kernel | label | dispatches | occupancy % | LDS bank conflict % | VALU busy % | LDS busy % |")
    print("|---|---|---|---|---|---|---|")
    rows = []
    for sym, a in acc.items():
        gui = a.get("GRBM_GUI_ACTIVE", 0.0)
        if not gui:
            continue
        occ = 400 * xcds * a.get("SQ_WAVE_CYCLES", 0.0) / gui / cu_num / waves_per_cu
        bank = 100 * xcds * a.get("SQ_LDS_BANK_CONFLICT", 0.0) / gui / cu_num
        valu = 100 * xcds * a.get("SQ_ACTIVE_INST_VALU", 0.0) / cu_num / gui
        lds = 100 * xcds * a.get("SQ_ACTIVE_INST_LDS", 0.0) / cu_num / gui
        rows.append((gui, sym, len(a["dispatches"]), occ, bank, valu, lds))
    for gui, sym, n, occ, bank, valu, lds in sorted(rows, reverse=True):
        print(f"| {sym} | {kernel_label(sym) or ''} | {n} | {occ:.1f} | {bank:.2f} | {valu:.1f} | {lds:.1f} |")


def summarize_pmc(fetch_dir, write_dir, events, copy_mib):
    """HBM bytes per event and profiler label from FETCH_SIZE and WRITE_SIZE runs.  The
    counters are scaled by what they read for a plain device copy of copy_mib MiB run three
    times at the start of the same program (the largest dispatch of the copy kernel)."""
    fetch, write = _pmc_per_kernel(fetch_dir, "FETCH_SIZE"), _pmc_per_kernel(write_dir, "WRITE_SIZE")
    nbytes = copy_mib << 20
    copy_sym = [k for k in fetch if "copy" in k.lower() or "elementwise" in k.lower()]
    if not copy_mib or not copy_sym:
        sys.exit("the runs hold no copy-calibration dispatch (--calibrate-copy)")
    cf_ = max(fetch[k][2] for k in copy_sym)
    cw_ = max(write[k][2] for k in copy_sym if k in write)
    fpu, wpu = nbytes / cf_, nbytes / cw_
    labels = {}
    for sym, (c, v, _) in fetch.items():
        lab = kernel_label(sym)
        if not lab:
            continue
        a = labels.setdefault(lab, [0.0, 0.0, 0])
        a[0] += v * fpu / events
        a[2] += c
    for sym, (c, v, _) in write.items():
        lab = kernel_label(sym)
        if lab:
            labels.setdefault(lab, [0.0, 0.0, 0])[1] += v * wpu / events
    out = {"what": "HBM bytes per event of each profiler label: FETCH_SIZE and WRITE_SIZE summed over its "
                   "dispatches, scaled by the bytes per counter unit of a plain device copy of known size",
           "events": events,
           "copy_calibration": {"bytes": nbytes, "fetch_units": cf_, "write_units": cw_,
                                "fetch_bytes_per_unit": round(fpu, 2), "write_bytes_per_unit": round(wpu, 2)},
           "labels": {k: round(v[0] + v[1]) for k, v in sorted(labels.items(), key=lambda kv: -(kv[1][0] + kv[1][1]))},
           "detail": {k: {"fetch_bytes_per_event": round(v[0]), "write_bytes_per_event": round(v[1]),
                          "dispatches": v[2]} for k, v in labels.items()}}
    print(json.dumps(out, indent=1))
    print(f"Copy calibration: a {nbytes} B device copy reads {cf_:.0f} FETCH_SIZE units and writes {cw_:.0f} "
          f"WRITE_SIZE units: {fpu:.1f} and {wpu:.1f} bytes per unit.  Events: {events}.\n", file=sys.stderr)
    print("| label | fetch B/event | write B/event | total B/event | dispatches |", file=sys.stderr)
    print("|---|---|---|---|---|", file=sys.stderr)
    for k, v in sorted(labels.items(), key=lambda kv: -(kv[1][0] + kv[1][1])):
        print(f"| {k} | {v[0]:.0f} | {v[1]:.0f} | {v[0] + v[1]:.0f} | {v[2]} |", file=sys.stderr)


if __name__ == "__main__":
    main()

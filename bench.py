#!/usr/bin/env python3
"""Benchmark of the MI355X pixel reconstruction chain on PU50-like events.

Workload: 1000 synthetic PU50 events (50 vertices x Poisson(35) tracks, 1/pT^2
spectrum above 0.5 GeV, ~50k digis and ~13k clusters per event), words
resident in HBM before the timed region.  One step = one pass of the whole
chain (raw words -> digis -> clusters -> rec-hits -> ntuplets with
broken-line fits -> vertices) over all events.  The batched path spreads the events over S
HIP streams and runs each stream's share as equal launches of at least B events per stage;
the plain path runs one launch sequence per event with the reference's
per-event host synchronisation.  Prints one JSON line.
"""
import argparse
import concurrent.futures as cf
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

# Algorithmic bytes per event of each kernel: the unique input it must read and
# the output it must write, with the reference's element widths, from the mean
# per-event counters (c): W words, D digis, H hits, M module runs, S doublet
# slots (layer pair x inner hit), C cells, N neighbour entries, T tuples, TH
# tuple hits, F fitted tuples, V vertex-finder tracks.  Per hit the CA kernels
# read x, y, z, r (16 B), iphi, ysize, detInd (6 B); a cell is inner + outer hit
# (4 B), pair and flags (2 B), list slots and counts (13 B).
def algorithmic_bytes(kernel, c):
    W, D, H, M = c["W"], c["D"], c["H"], c["M"]
    S, C, N, T, TH, F, V = c["S"], c["C"], c["N"], c["T"], c["TH"], c["F"], c["V"]
    table = {
        # stage A/B: decode reads words (4 B) + fedIds (0.5 B) per word, writes seven digi
        # arrays (20 B) and reads ~2 B of gains per digi; clusterize reads x, y, adc,
        # moduleInd (8 B) and writes clus + moduleInd (6 B); hits reads 12 B per digi and
        # writes 44 B per hit; moduleStarts reads 16 B of masks per 64 words and writes
        # 4 B per module run; phiBinner reads iphi and writes the bin contents + offsets,
        # and copies each hit's packed position and attributes (24 B) into histogram order
        "decode": 24.5 * W + 2 * D,
        "moduleStarts": W / 64 * 16 + 4 * M,
        "clusterize": 14 * D,
        "hits": 12 * D + 44 * H,
        "phiBinner": 4 * H + 48 * H + 5124,
        # doublets: hits (z, r, iphi, ysize, detInd: 14 B) + phi histogram (2 B per hit
        # + 5 kB) in, one count (4 B) per slot out; the fill pass writes the cells
        "doubletCount": 16 * H + 5124 + 4 * S,
        "doubletScan": 8 * S,
        "doubletFill": 16 * H + 5124 + 4 * S + 19 * C + 4 * H,
        "outerHitScan": 8 * H,
        "outerHitFill": 2 * C + 4 * H + 4 * C,
        "outerHitSort": 8 * C + C,
        # connect: every cell (6 B), the hits' x, y, z, r, detInd (18 B), the slot offsets,
        # the neighbour lists written (4 B per entry) and the used bits
        "connect": 6 * C + 18 * H + 4 * S + 4 * N + 5 * C + C / 8,
        "fishbone": 4 * C + 2 * C + 16 * H + C / 8 + C,
        "ntupletCount": 7 * C + 4 * N + 8 * C,
        "ntupletScan": 16 * C,
        "ntupletFill": 7 * C + 4 * N + 8 * C + 4 * T + 2 * TH + 16 * T + T,
        "trackListAlloc": 12 * C,
        "trackListFill": 16 * T + 8 * C + 2 * TH,
        "trackListSort": 8 * C + 2 * TH + 4 * T + T,
        "multiplicity": 5 * T + 2 * T + 4 * TH + 92 * T,
        # fit: per fitted tuple the offsets and hit indices, per hit x, y, z, xerr, yerr,
        # detInd (22 B) and its 36 B local frame, 92 B of results out
        "fit": F * (4 + 2 * 4 + 4 * (22 + 36) + 92),
        "classify": T * (4 + 1 + 4 * 5 + 12 + 1),
        # dupPreds: per cell with two or more tuples (MU) its list (2 B per tuple, ~3), the
        # tuples' cells (16 B each) and those cells' counts (4 B, ~3 per tuple), 33 B out
        "dupPreds": c.get("MU", 0.0) * (4 + 4 + 3 * (2 + 16 + 12) + 33),
        "fastDuplicate": 8 * C + 2 * TH + 4 * T + T,
        "hitToTupleCount": T * 5 + 2 * TH + 4 * H,
        "hitToTupleScan": 12 * H,
        "hitToTupleFill": T * 5 + 2 * TH + 8 * H + 2 * TH,
        "tripletCleaner": 8 * H + 2 * TH + 4 * T + T,
        "vertices": T * (4 + 1 + 4 + 4 + 4) + V * 24 + 4096 * 5,
    }
    return table.get(kernel)


def gen_events(detector, cfg, n, threads):
    from pxtrack import gen  # noqa: F401
    with cf.ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda i: detector.event(cfg, i)[0], range(n)))


def serve_at(ctx, inputs, rate, server_kw, seed=1, classes=None):
    """Submits `inputs` with Poisson arrivals at `rate` events/s (all at once if rate is None)
    to a serving loop and returns throughput and latency percentiles (ms)."""
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
        ids[sv.submit(w, f)] = i
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
    out = {"rate": rate, "events": len(res), "events_per_s": round(len(res) / span, 1),
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
    threads = min(16, os.cpu_count() or 1)
    raws = gen_events(det, gen.preset("pu_poisson60"), args.serve_events, threads)
    ev = [gpu.unpack_raw(r, cond["fed_ids"]) for r in raws]
    budget = int(args.serve_budget_gb * 1e9)
    shape = dict(doublet_lanes=tuned["doublet_lanes"], grid_target=tuned["grid_target"]) if tuned else {}
    # products stay on the device (no --transfer, the reference's default); counters come back
    batched_kw = dict(max_batch=args.batch, n_streams=3, hbm_budget=budget, upto=args.upto, transfer=False, **shape)
    plain_kw = dict(max_batch=1, n_streams=1, upto=args.upto, transfer=False, **shape)
    rates = [1000, 3000, 10000, 30000, None]
    out = {"workload": f"{len(ev)} events, pileup Poisson(60) clipped to [10, 200], Poisson arrivals",
           "hbm_budget_gb": args.serve_budget_gb, "batched": [], "plain": []}
    for r in rates:
        out["batched"].append(serve_at(ctx, ev, r, batched_kw))
        print(f"[bench] serving batched rate {r}: {out['batched'][-1]}", file=sys.stderr, flush=True)
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
    mix_kw = dict(max_batch=args.batch // 2, n_streams=3, hbm_budget=budget, upto=args.upto,
                  profile=gpu.PROFILE_HIGH_OCCUPANCY, transfer=False, **shape)
    out["mix_pu50_pu200"] = [serve_at(ctx, [m[0] for m in mix], r, mix_kw, classes=[m[1] for m in mix])
                             for r in (2000, None)]
    print(f"[bench] serving mix: {out['mix_pu50_pu200']}", file=sys.stderr, flush=True)
    return out


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
    ap.add_argument("--streams", type=int, default=3, help="HIP streams the batches are spread over")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop sweep")
    ap.add_argument("--serve-events", type=int, default=600)
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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: each box here has one MI355X and events shard across GPUs "
              f"without any exchange, so --gpus {args.gpus} is not run.", file=sys.stderr)
        sys.exit(2)

    from pxtrack import gen, gpu
    import torch

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
    threads = min(16, os.cpu_count() or 1)
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

    def batched_pass():
        for k, st in enumerate(streams):
            m = bounds[k + 1] - bounds[k]
            for first in range(0, m, B):
                st.run(first, min(B, m - first), args.upto)
        status = 0
        for k, st in enumerate(streams):
            status |= st.fetch_counts(0, bounds[k + 1] - bounds[k])
        return status

    # ---- launch shape chosen by the autotuner on the first batch
    tuned = None
    if not args.no_tune:
        lanes, grid, tms = stream.autotune(0, min(B, bounds[1]), args.upto)
        for st in streams[1:]:
            st.set_tuning(lanes, grid)
        tuned = {"doublet_lanes": lanes, "grid_target": grid, "ms_per_batch": round(tms, 4)}
        print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)

    # ---- batched path: warmup, then K timed steps (one step = all events)
    for _ in range(args.warmup):
        batched_pass()
    torch.cuda.synchronize()
    for st in streams:
        st.profile(True)
        st.profile_read()
        st.sync()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        status = batched_pass()
    for st in streams:
        st.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    prof = {}
    for st in streams:
        for k, (c, ms) in st.profile_read().items():
            a = prof.get(k, (0, 0.0))
            prof[k] = (a[0] + c, a[1] + ms)
        st.profile(False)
    ms_per_step = 1e3 * elapsed / args.steps
    evps = n * args.steps / elapsed

    # ---- the first stream's batches alone (no overlap with the other streams), for the
    # kernels' own durations
    iso = {}
    if args.isolated_passes > 0:
        st = streams[0]
        st.profile(True)
        st.profile_read()
        for _ in range(args.isolated_passes):
            for first in range(0, bounds[1], B):
                st.run(first, min(B, bounds[1] - first), args.upto)
        st.sync()
        iso = st.profile_read()
        st.profile(False)
        iso_events = bounds[1] * args.isolated_passes

    # per-event counters for the algorithmic bytes
    where = [(st, e) for k, st in enumerate(streams) for e in range(bounds[k + 1] - bounds[k])]
    cnts = np.array([st.event_counts(e) for st, e in where], dtype=np.float64)
    mean = {"W": float(W.mean()), "D": float(W.mean()), "M": float(cnts[:, 0].mean()), "H": float(cnts[:, 1].mean())}
    if args.upto >= 4:
        cac = [st.ca_counts(e) for st, e in where]
        for key, name in (("S", "nSlots"), ("C", "nCells"), ("N", "nNeighbors"), ("T", "nTuples"), ("TH", "apcN"),
                          ("F", "nMult"), ("V", "vtxTracks"), ("MU", "nMulti")):
            mean[key] = float(np.mean([c[name] for c in cac]))
        mean["vertices"] = float(np.mean([c["nVertices"] for c in cac]))
    else:
        for key in ("S", "C", "N", "T", "TH", "F", "V", "MU"):
            mean[key] = 0.0
    launches_per_kernel = {k: v[0] for k, v in prof.items()}
    # the dominant kernel: the most time with no other stream running (durations in the
    # timed region depend on what the other streams overlap them with)
    rank = iso if iso else prof
    kern = max((k for k in rank if k in prof), key=lambda k: rank[k][1])
    calls, total_ms = prof[kern]
    avg_s = total_ms / 1e3 / calls
    events_per_launch = n * args.steps / calls
    alg = algorithmic_bytes(kern, mean) * events_per_launch
    achieved = alg / avg_s / 1e9
    kernel_ms = {k: round(v[1] / v[0], 4) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])}
    kernel_ms_iso = {k: round(v[1] / v[0], 4) for k, v in sorted(iso.items(), key=lambda kv: -kv[1][1])}
    traffic, iso_roof = None, None
    tfile = os.path.join(REPO, "profiles", "traffic_chain_pu50_batched128.json")
    per_event = json.load(open(tfile))["kernels"] if os.path.exists(tfile) else {}
    if kern in per_event:
        traffic = round(per_event[kern]["hbm_bytes_per_event"] * events_per_launch)
    if kern in iso:
        ic, ims = iso[kern]
        iso_per_launch = iso_events / ic
        ialg = algorithmic_bytes(kern, mean) * iso_per_launch
        iso_roof = {"avg_launch_ms": round(ims / ic, 4), "launches": ic,
                    "events_per_launch": round(iso_per_launch, 1), "algorithmic_bytes_per_launch": round(ialg),
                    "achieved": round(ialg / (ims / ic / 1e3) / 1e9, 2),
                    "frac": round(ialg / (ims / ic / 1e3) / 1e9 / HBM_PEAK_GBS, 5)}
        if kern in per_event:
            iso_roof["traffic"] = round(per_event[kern]["hbm_bytes_per_event"] * iso_per_launch)
    share = {k: round(v[1] / sum(x[1] for x in prof.values()), 4) for k, v in prof.items()}

    # ---- plain per-event path: one launch sequence + host sync per event
    plain = None
    if not args.no_plain:
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
        lat = np.array(lat) * 1e3
        plain = {"events_per_s": round(n / tp, 1), "latency_ms_p50": round(float(np.percentile(lat, 50)), 4),
                 "latency_ms_p99": round(float(np.percentile(lat, 99)), 4),
                 "latency_ms_p999": round(float(np.percentile(lat, 99.9)), 4)}

    # ---- serving loop: continuous batching under Poisson arrivals
    serving = None
    if not args.no_serve:
        serving = serving_sweep(ctx, det, cond, args, tuned)

    # ---- CPU baseline: the oracle (serial restatement) on the host cores, bounded sample
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import pyoracle
        orc = pyoracle.Oracle(cond)
        sample = raws[:100]
        ncores = min(16, os.cpu_count() or 1)
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
               "single_thread_ms_per_event": round(one_ms, 3), "cpu_model": model}

    batch_lat_ms = ms_per_step / ((n + B - 1) // B)
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
                   "mean_per_event": {k: (round(v, 1) if v is not None else None) for k, v in mean.items()},
                   "stages": STAGES[args.upto]},
        "batched": {"events_per_s": round(evps, 1), "batch_latency_ms": round(batch_lat_ms, 4),
                    "status": status, "tuning": tuned},
        "plain": plain,
        "speedup_batched_vs_plain": round(evps / plain["events_per_s"], 2) if plain else None,
        "serving": serving,
        "kernel_ms_per_launch": kernel_ms,
        "kernel_ms_per_launch_isolated": kernel_ms_iso,
        "kernel_time_share": share,
        "launches": launches_per_kernel,
        "roofline": {"bound": "hbm", "kernel": kern, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "algorithmic_bytes_per_launch": round(alg), "avg_launch_ms": round(avg_s * 1e3, 4),
                     "launches": calls, "events_per_launch": round(events_per_launch, 1),
                     "isolated": iso_roof},
        "cpu_baseline": cpu,
        "device": gpu.device_name(),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

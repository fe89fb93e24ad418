"""GPU ntuplet and vertex stages against the oracle (through the C ABI).

Bit-exact: cells, outer-hit lists, neighbour lists, fishbone kills, tuples,
per-cell track lists, multiplicity, detector indices, quality flags, hit ->
tuple association, vertex assignment, ndof and ordering.  Fitted track
parameters and vertex z / weight / chi2 are floats computed in double (fits)
or float (vertices) with the serial operation order; the device and host
math libraries may differ in the last bit of atan2/log/sin/cos, so they are
compared with a relative tolerance of 1e-5 (float outputs) and an absolute
floor of 1e-10 for values that cancel to zero.
"""
import numpy as np
import pytest

from conftest import assert_same, event

pytestmark = pytest.mark.gpu

RTOL = 1e-5
ATOL = 1e-10  # values that cancel to ~0 (e.g. cot(theta) of a track at eta 0) differ by double-precision noise

CASES = [
    ("serial100", 0, {}),
    ("serial100", 7, {}),
    ("pu50", 0, dict(pu_lo=5, pu_hi=5)),
    ("pu50", 0, {}),
    ("pu50", 3, {}),
    ("pu50", 2, {}),  # 3 vertices split
    ("pu50", 4, {}),  # 2 vertices split
    ("pu20_80", 2, {}),
    ("pu50", 5, {"error_rate": 0.1}),
]

EXACT_CA = ["cellInner", "cellOuter", "cellLayerPair", "cellDoubletId", "cellUsed", "outerHitOff", "outerHitCells",
            "cellNeighborOff", "cellNeighbors", "cellTrackOff", "cellTracks", "multOff"]


def _close(a, b, name, ctx):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    if a.shape != b.shape:
        return f"{ctx} {name}: shape {a.shape} != {b.shape}"
    same = (a == b) | (np.isnan(a) & np.isnan(b)) | \
        (np.abs(a - b) <= RTOL * np.maximum(np.abs(a), np.abs(b)) + ATOL)
    if not same.all():
        bad = np.nonzero(~same)[0]
        return (f"{ctx} {name}: {bad.size} of {a.size} differ beyond {RTOL}, first at {bad[:5]}: "
                f"{a[bad[:5]]} vs {b[bad[:5]]}")
    return None


def _exact(a, b, name, ctx):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        return f"{ctx} {name}: shape {a.shape} != {b.shape}"
    if not np.array_equal(a, b):
        bad = np.nonzero(a != b)[0]
        return f"{ctx} {name}: {bad.size} of {a.size} differ, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"
    return None


def compare_ca(o, g, ctx, internals=True):
    """Collects every mismatching product of one event, then fails with all of them."""
    errs = []
    n_t = o["nTuples"]
    if g["nTuples"] != n_t:
        errs.append(f"{ctx}: tuples {g['nTuples']} != {n_t}")
    if internals:
        for k in ("nCells", "caStatus"):
            if g[k] != o[k]:
                errs.append(f"{ctx}: {k} {g[k]} != {o[k]}")
        for k in EXACT_CA:
            errs.append(_exact(o[k], g[k], k, ctx))
        nm = int(o["multOff"][8])
        errs.append(_exact(o["multBins"][:nm], g["multBins"], "multBins", ctx))
        nh = o["nHits"]
        errs.append(_exact(o["hitToTupleOff"][:nh + 1], g["hitToTupleOff"], "hitToTupleOff", ctx))
        ne = int(o["hitToTupleOff"][nh])
        errs.append(_exact(o["hitToTupleBins"][:ne], g["hitToTupleBins"], "hitToTupleBins", ctx))
    for k in ("tupleOff", "tupleBins", "detIndices"):
        errs.append(_exact(o[k], g[k], k, ctx))
    errs.append(_exact(o["quality"][:n_t], g["quality"][:n_t], "quality", ctx))
    for k in ("pt", "eta", "chi2"):
        errs.append(_close(o[k][:n_t], g[k][:n_t], k, ctx))
    errs.append(_close(o["state"][:5 * n_t], g["state"][:5 * n_t], "state", ctx))
    errs.append(_close(o["covariance"][:15 * n_t], g["covariance"][:15 * n_t], "covariance", ctx))
    if "nVertices" in g:
        nv = o["nVertices"]
        if g["nVertices"] != nv:
            errs.append(f"{ctx}: vertices {g['nVertices']} != {nv}")
        else:
            errs.append(_exact(o["idv"][:n_t], g["idv"][:n_t], "idv", ctx))
            for k in ("ndof", "sortInd"):
                errs.append(_exact(o[k][:nv], g[k][:nv], k, ctx))
            for k in ("zv", "wv", "chi2v", "ptv2"):
                errs.append(_close(o[k][:nv], g[k][:nv], k, ctx))
    errs = [e for e in errs if e]
    assert not errs, "\n".join(errs)


def _gpu_event(ctx_gpu, stream, raw, upto=5, internals=True):
    from pxtrack import gpu
    words, fed = gpu.unpack_raw(raw, ctx_gpu.fed_ids)
    return stream.process(words, fed, upto=upto, internals=internals)


@pytest.mark.parametrize("preset,index,kw", CASES)
def test_plain_ntuplets_and_vertices_match_oracle(detector, oracle, gpu_ctx, gpu_stream, preset, index, kw):
    raw, _ = event(detector, preset, index, **kw)
    o = oracle.process(raw, upto=5)
    print(f"{preset}[{index}] hits={o['nHits']} cells={o['nCells']} tuples={o['nTuples']} "
          f"vertices={o['nVertices']}", flush=True)
    g = _gpu_event(gpu_ctx, gpu_stream, raw)
    compare_ca(o, g, f"{preset}[{index}]")


def test_empty_event(detector, oracle, gpu_ctx, gpu_stream):
    g = gpu_stream.process(np.zeros(0, np.uint32), np.zeros(0, np.uint8), upto=5, internals=True)
    assert g["nTuples"] == 0 and g["nCells"] == 0 and g["nVertices"] == 0


def test_batched_matches_plain(detector, oracle, gpu_ctx):
    from pxtrack import gpu
    evs = [event(detector, "pu20_80", i)[0] for i in range(6)]
    unpacked = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in evs]
    s = gpu_ctx.stream(max_events=len(evs), max_words=sum(w.size for w, _ in unpacked) + 256 * len(evs))
    s.load(unpacked)
    s.run(0, len(evs), upto=5)
    s.fetch_counts(0, len(evs))
    for e, raw in enumerate(evs):
        print(f"batched event {e}", flush=True)
        o = oracle.process(raw, upto=5)
        g = s.event_products(e, upto=5)
        g.update(s.ca_internals(e, int(g["nHits"])))
        compare_ca(o, g, f"batch[{e}]")
    # a sub-range of the resident events gives the same products
    s.run(2, 3, upto=5)
    s.fetch_counts(2, 3)
    for e in range(2, 5):
        o = oracle.process(evs[e], upto=5)
        compare_ca(o, s.event_products(e, upto=5), f"sub-batch[{e}]", internals=False)


def test_high_occupancy_profile(detector, conditions, gpu_ctx):
    """PU200 events under the high-occupancy capacities: nothing overflows (status 0),
    more than 65535 tuples (indices past the reference's 16-bit tindex_type) and outer
    hits with more than the reference's 128 cells."""
    import pyoracle
    from pxtrack import gpu
    orc = pyoracle.Oracle(conditions, profile=1)
    s = gpu_ctx.stream(max_events=1, max_words=300000, profile=gpu.PROFILE_HIGH_OCCUPANCY)
    for idx in (1, 4):
        raw, _ = event(detector, "pu200", idx)
        o = orc.process(raw, upto=5)
        print(f"pu200[{idx}] hits={o['nHits']} cells={o['nCells']} tuples={o['nTuples']} "
              f"status={o['caStatus']:#x}", flush=True)
        assert o["caStatus"] == 0 and o["nTuples"] > 65535
        assert np.diff(o["outerHitOff"]).max() > 128
        compare_ca(o, _gpu_event(gpu_ctx, s, raw), f"pu200[{idx}]/high-occupancy")


def test_high_occupancy_batch(detector, conditions, gpu_ctx):
    """A batch of 8 events (two PU200, six PU50) under the high-occupancy capacities, as one
    launch per stage: the per-event LDS builds of the outer-hit lists (hits past 32768, so
    both 16-bit halves of the counters), the track lists (cells in many slices) and the
    hit -> tuple map (hits in several slices), and module runs of more than 256 digis in the
    clustering and rec-hit kernels, each event against the oracle."""
    import pyoracle
    from pxtrack import gpu
    orc = pyoracle.Oracle(conditions, profile=1)
    evs = [event(detector, "pu200", i)[0] for i in (1, 4)] + [event(detector, "pu50", i)[0] for i in range(6)]
    unpacked = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in evs]
    s = gpu_ctx.stream(max_events=len(evs), max_words=sum(w.size for w, _ in unpacked) + 256 * len(evs),
                       profile=gpu.PROFILE_HIGH_OCCUPANCY)
    s.load(unpacked)
    s.run(0, len(evs), upto=5)
    s.fetch_counts(0, len(evs))
    for e, raw in enumerate(evs):
        print(f"high-occupancy batch event {e}", flush=True)
        o = orc.process(raw, upto=5)
        assert o["caStatus"] == 0
        g = s.event_products(e, upto=5)
        g.update(s.ca_internals(e, int(g["nHits"])))
        compare_ca(o, g, f"high-occupancy batch[{e}]")
    assert max(int(orc.process(r, upto=3)["nHits"]) for r in evs[:2]) > 32768


def test_reference_capacities_overflow(detector, oracle, gpu_ctx):
    """PU200 under the reference's capacities (CAConstants.h): more doublets than
    maxNumberOfDoublets, more tuples than maxNumberOfTuples, a full multiplicity
    container.  The products keep what the serial program keeps (its first cells and
    tuples in serial order) and the status bits are the same."""
    s = gpu_ctx.stream(max_events=1, max_words=400000)
    for idx in (1, 2):
        raw, _ = event(detector, "pu200", idx)
        o = oracle.process(raw, upto=5)
        print(f"pu200[{idx}] reference caps: cells={o['nCells']} tuples={o['nTuples']} status={o['caStatus']:#x}",
              flush=True)
        assert o["caStatus"] & 1  # the cell container overflowed
        compare_ca(o, _gpu_event(gpu_ctx, s, raw), f"pu200[{idx}]/reference")


@pytest.mark.parametrize("algo", [1, 2])
def test_alternative_vertex_clustering(detector, conditions, gpu_ctx, algo):
    """DBSCAN (1) and iterative (2) vertex clustering, the reference's off-by-default
    switches (gpuVertexFinder.h:46-60), on the whole chain against the oracle."""
    import pyoracle
    orc = pyoracle.Oracle(conditions, vertex_algo=algo)
    s = gpu_ctx.stream(max_events=1, max_words=300000)
    s.set_vertex_algorithm(algo)
    for idx in (0, 3):
        raw, _ = event(detector, "pu50", idx)
        o = orc.process(raw, upto=5)
        print(f"algo {algo} pu50[{idx}] tracks={o['vtxTracks']} vertices={o['nVertices']}", flush=True)
        compare_ca(o, _gpu_event(gpu_ctx, s, raw, internals=False), f"algo {algo} pu50[{idx}]", internals=False)


def test_riemann_fit(detector, conditions, gpu_ctx):
    """The reference's useRiemannFit switch (CAHitNtupletGeneratorOnGPU.cc:168-172,
    RiemannFitOnGPU.cc:3-113) on the whole chain against the oracle: fitted
    parameters and covariances within the fits' tolerance, everything else exact."""
    import pyoracle
    orc = pyoracle.Oracle(conditions, riemann_fit=True)
    s = gpu_ctx.stream(max_events=1, max_words=300000)
    s.set_riemann_fit(True)
    for idx in (0, 3):
        raw, _ = event(detector, "pu50", idx)
        o = orc.process(raw, upto=5)
        print(f"riemann pu50[{idx}] tuples={o['nTuples']} vertices={o['nVertices']}", flush=True)
        compare_ca(o, _gpu_event(gpu_ctx, s, raw), f"riemann pu50[{idx}]")


def test_bfield_and_vertex_ptmin(detector, conditions, gpu_ctx):
    """The field makeTuples takes (CAHitNtupletCUDA.cc:31) and the vertex finder's pt cut
    (PixelVertexProducerCUDA.cc:37) as parameters: a weaker field and a higher cut on both
    sides change the products the same way, on the plain path and on a batched launch
    (the stream settings), for the broken-line and the Riemann fit."""
    import pyoracle
    from pxtrack import gpu
    bf, ptmin = 0.0095, 0.8
    raws = [event(detector, "pu50", idx)[0] for idx in (1, 4)]
    default = pyoracle.Oracle(conditions).process(raws[0], upto=5)
    for riemann in (False, True):
        orc = pyoracle.Oracle(conditions, bfield=bf, vertex_ptmin=ptmin, riemann_fit=riemann)
        plain = gpu_ctx.stream(max_events=1, max_words=300000)
        plain.set_riemann_fit(riemann)
        plain.bfield, plain.vertex_ptmin = bf, ptmin  # passed by pxg_ntuplets / pxg_find_vertices
        batch = gpu_ctx.stream(max_events=2, max_words=600000)
        batch.set_riemann_fit(riemann)
        batch.set_bfield(bf)
        batch.set_vertex_ptmin(ptmin)
        batch.load([gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in raws])
        batch.run(0, 2, upto=5)
        batch.fetch_counts(0, 2)
        for i, raw in enumerate(raws):
            o = orc.process(raw, upto=5)
            if i == 0 and not riemann:
                n = o["nTuples"]
                assert not np.array_equal(o["pt"][:n], default["pt"][:n])
                assert o["vtxTracks"] < default["vtxTracks"]
            print(f"bfield {bf}, ptMin {ptmin}, riemann {riemann}: event {i} tracks to the vertexer "
                  f"{o['vtxTracks']}, vertices {o['nVertices']}", flush=True)
            compare_ca(o, _gpu_event(gpu_ctx, plain, raw), f"bf/ptmin plain[{i}] riemann={riemann}")
            compare_ca(o, batch.event_products(i, upto=5), f"bf/ptmin batch[{i}] riemann={riemann}",
                       internals=False)
        plain.close()
        batch.close()


@pytest.mark.parametrize("early,late", [(False, True), (True, True), (False, False)])
def test_fishbone_switches(detector, conditions, gpu_ctx, early, late):
    """The reference's earlyFishbone / lateFishbone switches
    (CAHitNtupletGeneratorKernels.cc:109-112, :133-136, :161-164) on the whole chain,
    kills and per-cell track lists included, against the oracle."""
    import pyoracle
    orc = pyoracle.Oracle(conditions, early_fishbone=early, late_fishbone=late)
    s = gpu_ctx.stream(max_events=1, max_words=300000)
    s.set_fishbone(early, late)
    for idx in (0, 3):
        raw, _ = event(detector, "pu50", idx)
        o = orc.process(raw, upto=5)
        kills = int((o["cellDoubletId"] < 0).sum())
        print(f"fishbone early={early} late={late} pu50[{idx}] kills={kills} tuples={o['nTuples']}", flush=True)
        compare_ca(o, _gpu_event(gpu_ctx, s, raw), f"fishbone {early}/{late} pu50[{idx}]")


def test_alternative_switches_together_at_pu200(detector, conditions, gpu_ctx):
    """The Riemann fit, the late fishbone without the early one and DBSCAN vertexing at
    once, on a PU200 event under the high-occupancy capacities, against the oracle."""
    import pyoracle
    from pxtrack import gpu
    orc = pyoracle.Oracle(conditions, profile=1, vertex_algo=1, early_fishbone=False, late_fishbone=True,
                          riemann_fit=True)
    s = gpu_ctx.stream(max_events=1, max_words=400000, profile=gpu.PROFILE_HIGH_OCCUPANCY)
    s.set_riemann_fit(True)
    s.set_fishbone(False, True)
    s.set_vertex_algorithm(1)
    raw, _ = event(detector, "pu200", 3)
    o = orc.process(raw, upto=5)
    print(f"pu200 switches: cells={o['nCells']} tuples={o['nTuples']} vertices={o['nVertices']} "
          f"status={o['caStatus']:#x}", flush=True)
    compare_ca(o, _gpu_event(gpu_ctx, s, raw), "pu200/switches")


def test_stats_counters(detector, oracle, gpu_ctx):
    """doStats (CAHitNtupletGeneratorKernels.h:14-26): the counters the GPU accumulates
    over a batch equal the oracle's per-event counts summed; off by default, a second
    pass adds the same again, and a reset clears them."""
    from pxtrack import gpu
    evs = [event(detector, "pu20_80", 30 + i)[0] for i in range(4)]
    unpacked = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in evs]
    s = gpu_ctx.stream(max_events=4, max_words=sum(w.size for w, _ in unpacked) + 1024)
    s.load(unpacked)
    s.run(0, 4, upto=5)
    s.fetch_counts(0, 4)
    assert all(v == 0 for v in s.counters().values())
    s.set_stats(True)
    s.run(0, 4, upto=5)
    s.fetch_counts(0, 4)
    exp = np.sum([oracle.process(r, upto=5)["stats"].astype(np.int64) for r in evs], axis=0)
    got = s.counters()
    print("counters", got, flush=True)
    assert [got[k] for k in s.COUNTERS] == exp.tolist()
    s.run(0, 4, upto=5)
    s.fetch_counts(0, 4)
    assert [s.counters()[k] for k in s.COUNTERS] == (2 * exp).tolist()
    s.print_counters()
    s.reset_counters()
    assert all(v == 0 for v in s.counters().values())


def test_graph_replays_equal_eager_runs(detector, oracle, gpu_ctx, gpu_stream):
    """pxg_set_graphs: the captured stage sequences give the eager run's products on
    every replay, with other eager work and synchronisations between the replays, and
    for two batch shapes; a setting change drops them (the next run captures again)."""
    from pxtrack import gpu
    evs = [event(detector, "pu20_80", 50 + i)[0] for i in range(3)]
    unpacked = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in evs]
    s = gpu_ctx.stream(max_events=3, max_words=sum(w.size for w, _ in unpacked) + 1024)
    s.load(unpacked)

    def products(first, count):
        s.fetch_counts(first, count)
        got = [s.event_products(e, upto=5) for e in range(first, first + count)]
        for k, e in enumerate(range(first, first + count)):
            got[k].update(s.ca_internals(e, int(got[k]["nHits"])))
        return got

    def same(ref, got, ctx):
        for e in range(len(ref)):
            nt, nv = int(ref[e]["nTuples"]), int(ref[e]["nVertices"])
            assert (int(got[e]["nTuples"]), int(got[e]["nVertices"])) == (nt, nv), ctx
            used = {"quality": nt, "state": 5 * nt, "idv": nt, "zv": nv}
            for k in ("tupleOff", "tupleBins", "quality", "state", "cellInner", "cellOuter", "cellNeighbors",
                      "cellTracks", "idv", "zv", "xx", "adc", "clus", "xg", "zg"):
                n = used.get(k, len(ref[e][k]))
                assert np.array_equal(ref[e][k][:n], got[e][k][:n]), f"{ctx} event {e}: {k}"

    s.run(0, 3, upto=5)
    ref = products(0, 3)
    compare_ca(oracle.process(evs[0], upto=5), ref[0], "eager event 0")
    s.run(1, 2, upto=5)
    ref12 = products(1, 2)
    s.set_graphs(True)
    other = gpu.unpack_raw(event(detector, "pu50", 9)[0], gpu_ctx.fed_ids)
    for rep in range(4):
        print(f"graph replay {rep}", flush=True)
        s.run(0, 3, upto=5)
        same(ref, products(0, 3), f"replay {rep}")
        gpu_stream.process(*other, upto=5)  # eager work on another stream in between
        s.run(1, 2, upto=5)
        same(ref12, products(1, 2), f"replay {rep} of events 1-2")
    s.set_tuning(8, 1024)  # drops the graphs: the next run captures with the new shape
    s.run(0, 3, upto=5)
    same(ref, products(0, 3), "after a setting change")
    s.set_graphs(False)


def test_products_do_not_depend_on_launch_shape(detector, oracle, gpu_ctx):
    """Every launch shape the autotuner may pick gives the same products."""
    from pxtrack import gpu
    evs = [event(detector, "pu20_80", 10 + i)[0] for i in range(3)]
    unpacked = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in evs]
    s = gpu_ctx.stream(max_events=3, max_words=sum(w.size for w, _ in unpacked) + 1024)
    s.load(unpacked)
    ref = None
    shapes = [dict(doublet_lanes=1, grid_target=64), dict(doublet_lanes=8, grid_target=2048),
              dict(doublet_lanes=64, grid_target=8192), dict(doublet_lanes=4, grid_target=333),
              dict(connect_stage=128, fishbone_window=128, max_blocks_per_event=4),
              dict(connect_stage=512, fishbone_window=384, max_blocks_per_event=4096, grid_target=65536),
              dict(connect_stage=256, fishbone_window=256, max_blocks_per_event=512, grid_target=16384)]
    for shape in shapes:
        s.set_launch_shape(shape)
        lanes, grid = s.launch_shape()["doublet_lanes"], s.launch_shape()["grid_target"]
        s.run(0, 3, upto=5)
        s.fetch_counts(0, 3)
        got = [s.event_products(e, upto=5) for e in range(3)]
        for e in range(3):
            got[e].update(s.ca_internals(e, int(got[e]["nHits"])))
        if ref is None:
            ref = got
            for e in range(3):
                compare_ca(oracle.process(evs[e], upto=5), got[e], f"shape {lanes}/{grid} event {e}")
        else:
            for e in range(3):
                # the stored products (the containers' tails past them hold stale memory)
                nt, nv = int(ref[e]["nTuples"]), int(ref[e]["nVertices"])
                assert (int(got[e]["nTuples"]), int(got[e]["nVertices"])) == (nt, nv), f"shape {lanes}/{grid}"
                used = {"quality": nt, "state": 5 * nt, "idv": nt, "zv": nv}
                for k in ("tupleOff", "tupleBins", "quality", "state", "cellInner", "cellOuter", "cellNeighbors",
                          "cellTracks", "idv", "zv"):
                    n = used.get(k, len(ref[e][k]))
                    assert np.array_equal(ref[e][k][:n], got[e][k][:n]), f"shape {lanes}/{grid} event {e}: {k}"
    lanes, grid, ms = s.autotune(0, 3, upto=5)
    assert lanes in (2, 4, 8, 16, 32, 64) and grid >= 1024 and ms > 0
    shape, ms = s.autotune_shape(0, 3, upto=5)
    assert shape == s.launch_shape() and ms > 0
    assert shape["connect_stage"] in (128, 256, 512) and shape["fishbone_window"] in (128, 256, 384)
    with pytest.raises(RuntimeError):
        s.set_launch_shape(dict(connect_stage=200))

"""Parity of the HIP stage A/B path (raw -> digis -> clusters -> rec-hits) with
the oracle, bit-exact on every product field, on the plain per-event path and
on the batched path."""
import numpy as np
import pytest

from conftest import assert_same, assert_same_errors, event

pytestmark = pytest.mark.gpu

DIGI_KEYS = ["xx", "yy", "adc", "moduleInd", "clus", "pdigi", "rawIdArr"]
CLUSTER_KEYS = ["clusInModule", "moduleId", "clusModuleStart"]
HIT_KEYS = ["xl", "yl", "xerr", "yerr", "xg", "yg", "zg", "rg", "iphi", "charge", "xsize", "ysize", "detInd",
            "hitsLayerStart", "phiOff", "phiBins", "avgCorrected"]

CASES = [("serial100", 0, {}), ("serial100", 1, {}), ("serial100", 7, {}), ("pu50", 0, {}), ("pu50", 3, {}),
         ("serial100", 2, {"error_rate": 0.5}), ("pu50", 1, {"error_rate": 0.2}), ("pu20_80", 5, {}),
         ("pu200", 0, {})]


def _compare(g, o, ctx):
    assert g["nDigis"] == o["nDigis"], ctx
    assert g["nModules"] == o["nModules"], ctx
    assert_same(g, o, DIGI_KEYS, ctx)
    assert_same_errors(g["errors"], o["errors"], ctx)
    n = o["nModules"]
    assert np.array_equal(g["moduleStart"][: n + 1], o["moduleStart"][: n + 1]), f"{ctx}: moduleStart"
    assert_same(g, o, CLUSTER_KEYS, ctx)
    assert g["nClusters"] == o["nClusters"], ctx
    if "xl" in g:
        assert g["nHits"] == o["nHits"], ctx
        assert_same(g, o, HIT_KEYS, ctx)


@pytest.mark.parametrize("preset,index,kw", CASES)
def test_plain_path_matches_oracle(detector, conditions, oracle, gpu_stream, preset, index, kw):
    from pxtrack import gpu
    raw, _ = event(detector, preset, index, **kw)
    words, feds = gpu.unpack_raw(raw, conditions["fed_ids"])
    o = oracle.process(raw)
    g = gpu_stream.process(words, feds)
    print(f"{preset}[{index}] W={o['nWords']} modules={o['nModules']} clusters={o['nClusters']}", flush=True)
    assert g["status"] == 0
    _compare(g, o, f"{preset}[{index}]")


def test_raw_to_digi_only(detector, conditions, oracle, gpu_stream):
    """The decode-only entry point leaves clus = digi index, as countModules does."""
    from pxtrack import gpu
    raw, _ = event(detector, "pu50", 2)
    words, feds = gpu.unpack_raw(raw, conditions["fed_ids"])
    o = oracle.process(raw, upto=1)
    g = gpu_stream.process(words, feds, upto=1)
    assert_same(g, o, ["xx", "yy", "adc", "moduleInd", "pdigi", "rawIdArr"], "digi-only")
    assert np.array_equal(g["clus"], np.arange(len(words), dtype=np.int32))
    assert_same_errors(g["errors"], o["errors"], "digi-only")


def test_empty_event(gpu_stream):
    g = gpu_stream.process(np.zeros(0, np.uint32), np.zeros(0, np.uint8))
    assert g["nDigis"] == 0 and g["nModules"] == 0 and g["nClusters"] == 0 and g["nHits"] == 0
    assert g["errors"].size == 0


def test_batched_equals_plain_and_oracle(detector, conditions, oracle, gpu_ctx):
    """One launch per stage over events of different sizes gives the per-event products."""
    from pxtrack import gpu
    cases = [("pu20_80", i, {}) for i in range(6)] + [("serial100", 3, {}), ("pu50", 4, {"error_rate": 0.3})]
    raws = [event(detector, p, i, **kw)[0] for p, i, kw in cases]
    inputs = [gpu.unpack_raw(r, conditions["fed_ids"]) for r in raws]
    s = gpu_ctx.stream(max_events=len(inputs), max_words=sum(len(w) for w, _ in inputs))
    s.load(inputs)
    s.run(0, len(inputs), 3)
    assert s.fetch_counts(0, len(inputs)) == 0
    for e, raw in enumerate(raws):
        g = s.event_products(e)
        o = oracle.process(raw)
        _compare(g, o, f"batch event {e}")
    # a sub-range launch (events 2..4) reproduces the same products
    s.run(2, 3, 3)
    assert s.fetch_counts(2, 3) == 0
    for e in range(2, 5):
        g = s.event_products(e)
        _compare(g, oracle.process(raws[e]), f"sub-batch event {e}")
    s.close()


def test_status_reports_only_the_stages_run(detector, conditions, gpu_ctx):
    """A raw-to-cluster call reports the status of its own stages, not the overflow
    bits an earlier event left in the stream's ntuplet counters."""
    from pxtrack import gpu
    s = gpu_ctx.stream(max_events=1, max_words=400000)
    big = gpu.unpack_raw(event(detector, "pu200", 1)[0], conditions["fed_ids"])
    assert s.process(*big, upto=5)["status"] >> 8 & 1, "the PU200 event overflows the reference cell capacity"
    small = gpu.unpack_raw(event(detector, "serial100", 4)[0], conditions["fed_ids"])
    assert s.process(*small, upto=2)["status"] == 0
    assert s.process(*small, upto=3)["status"] == 0
    s.close()

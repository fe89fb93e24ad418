"""Parity of the launch shapes the benchmark runs: hundreds of events per launch
on three HIP streams at once, launches that start past event 0, the launch
shape the autotuner picks, and the 128- and 32-event batches of the batched
configuration (pileup uniform in 20-80).

Every event is compared with the plain per-event path bit for bit (digis,
errors, clusters, hits, tuples, quality, fitted parameters and covariances,
vertex assignment, vertices), and a spread sample of them with the oracle
(integer products exact, fitted values within the fits' tolerance).
"""
import concurrent.futures as cf

import pytest

from conftest import event
from test_gpu_ca import compare_ca
from test_gpu_stageAB import _compare as compare_ab

pytestmark = pytest.mark.gpu


def _events(detector, preset, first, n):
    from pxtrack import gen
    cfg = gen.preset(preset)
    with cf.ThreadPoolExecutor(8) as ex:
        return list(ex.map(lambda i: detector.event(cfg, i)[0], range(first, first + n)))


def assert_identical(a, b, ctx):
    """Bit-exact equality of two events' stored products (floats as bit patterns)."""
    from pxtrack import gpu
    bad = gpu.differing_products(a, b)
    assert not bad, f"{ctx}: {bad} differ from the plain path"


def _load(gpu_ctx, inputs):
    s = gpu_ctx.stream(max_events=len(inputs), max_words=int(sum(len(w) for w, _ in inputs)))
    s.load(inputs)
    return s


def test_bench_launch_shapes_match_plain_and_oracle(detector, oracle, gpu_ctx, gpu_stream):
    from pxtrack import gpu
    pu50 = _events(detector, "pu50", 0, 334)
    mixed = _events(detector, "pu20_80", 200, 128)
    in50 = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in pu50]
    inmix = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in mixed]
    # stream 0: the benchmark's share of one stream (two launches of 167 PU50 events);
    # stream 1: the mixed events as one launch of 128; stream 2: the same as four of 32
    plan = [(_load(gpu_ctx, in50), [(0, 167), (167, 167)]),
            (_load(gpu_ctx, inmix), [(0, 128)]),
            (_load(gpu_ctx, inmix), [(32 * k, 32) for k in range(4)])]
    shape, ms = plan[0][0].autotune_shape(0, 167, upto=5)
    print(f"autotuned shape: {shape} ({ms:.2f} ms)", flush=True)
    for s, _ in plan[1:]:
        s.set_launch_shape(shape)
    # every stream's launches enqueued before any result is read, as the benchmark does
    for s, launches in plan:
        for first, count in launches:
            s.run(first, count, upto=5)
    status = [s.fetch_counts(0, sum(c for _, c in launches)) for s, launches in plan]
    print(f"status per stream {status}", flush=True)

    for (s, launches), inputs, raws, name in ((plan[0], in50, pu50, "pu50"), (plan[1], inmix, mixed, "pu20_80x128"),
                                              (plan[2], inmix, mixed, "pu20_80x32")):
        for e in range(len(inputs)):
            g = s.event_products(e, upto=5)
            ctx = f"{name} event {e}"
            alone = gpu_stream.process(*inputs[e], upto=5)
            assert_identical(alone, g, ctx)
            if e % 23 == 0 or e == len(inputs) - 1:
                o = oracle.process(raws[e], upto=5)
                compare_ab(g, o, ctx)
                compare_ca(o, g, ctx, internals=False)
        print(f"{name}: {len(inputs)} events identical to the plain path", flush=True)
    for s, _ in plan:
        s.close()


def test_sub_range_launches_on_a_fresh_pool(detector, oracle, gpu_ctx, gpu_stream):
    """Two launches of 9 and 12 events on a freshly loaded pool, the second starting at
    event 9, with the grid spread the benchmark uses (about 49 workgroups per event),
    against the plain path and the oracle for every event."""
    from pxtrack import gpu
    raws = _events(detector, "pu50", 400, 21)
    inputs = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in raws]
    s = _load(gpu_ctx, inputs)
    s.set_tuning(4, 8192)
    s.run(0, 9, upto=5)
    s.run(9, 12, upto=5)
    s.fetch_counts(0, 21)
    for e in range(21):
        g = s.event_products(e, upto=5)
        assert_identical(gpu_stream.process(*inputs[e], upto=5), g, f"sub-range event {e}")
        o = oracle.process(raws[e], upto=5)
        compare_ab(g, o, f"sub-range event {e}")
        compare_ca(o, g, f"sub-range event {e}", internals=False)
    s.close()

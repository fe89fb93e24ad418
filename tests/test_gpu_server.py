"""The serving loop: events submitted one by one are batched continuously and
each event's products are the same as when it runs alone on the plain path."""
import time

import numpy as np
import pytest

from conftest import event

pytestmark = pytest.mark.gpu


def _plain(gpu_stream, words, fed):
    return gpu_stream.process(words, fed, upto=5)


def _same(alone, mixed, ctx):
    n = alone["nTuples"]
    assert mixed["nTuples"] == n, ctx
    assert np.array_equal(alone["tupleOff"][:n + 1], mixed["tupleOff"]), f"{ctx}: tupleOff"
    nb = int(alone["tupleOff"][n])
    assert np.array_equal(alone["tupleBins"][:nb], mixed["tupleBins"]), f"{ctx}: tupleBins"
    assert np.array_equal(alone["quality"][:n], mixed["quality"]), f"{ctx}: quality"
    assert np.array_equal(alone["state"][:5 * n].view(np.uint32), mixed["state"].view(np.uint32)), f"{ctx}: state"
    assert np.array_equal(alone["idv"][:n], mixed["idv"]), f"{ctx}: idv"
    nv = alone["nVertices"]
    assert mixed["nVertices"] == nv, ctx
    assert np.array_equal(alone["zv"][:nv].view(np.uint32), mixed["zv"].view(np.uint32)), f"{ctx}: zv"
    assert np.array_equal(alone["sortInd"][:nv], mixed["sortInd"]), f"{ctx}: sortInd"


def test_server_mix_equals_alone(detector, gpu_ctx, gpu_stream):
    from pxtrack import gpu
    rng = np.random.default_rng(11)
    raws = [event(detector, "pu20_80", 40 + i)[0] for i in range(24)]
    inputs = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in raws]
    sv = gpu.Server(gpu_ctx, max_batch=8, n_streams=2)
    ids = {}
    for i in rng.permutation(len(inputs)):
        ids[sv.submit(*inputs[i])] = i
        time.sleep(float(rng.exponential(0.0005)))
    results = []
    deadline = time.time() + 60
    while len(results) < len(ids) and time.time() < deadline:
        results += sv.poll(wait_ms=100)
    assert len(results) == len(ids), f"{len(results)} of {len(ids)} completed"
    assert max(r["batch_size"] for r in results) > 1, "no batching happened"
    for r in results:
        i = ids[r["id"]]
        assert r["t_submit_ns"] <= r["t_launch_ns"] <= r["t_done_ns"]
        mixed = sv.products(r["id"])
        _same(_plain(gpu_stream, *inputs[i]), mixed, f"event {i} (batch of {r['batch_size']})")
    info = sv.info()
    assert info["events"] == len(ids) and info["batches"] < len(ids)
    sv.close()


def test_server_pinned_submission(detector, gpu_ctx, gpu_stream):
    """Events handed over in the caller's mapped pinned buffers (pxg_server_submit_pinned:
    the device reads them, no host copy) mixed with copied ones in the same batches: every
    result equals the plain path's; memory that is not mapped pinned is refused."""
    from pxtrack import gpu
    raws = [event(detector, "pu20_80", 90 + i)[0] for i in range(16)]
    inputs = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in raws]
    pinned = [gpu.PinnedEvent(*x) for x in inputs]
    sv = gpu.Server(gpu_ctx, max_batch=8, n_streams=2)
    ids = {}
    for i, x in enumerate(inputs):
        rid = sv.submit_pinned(pinned[i]) if i % 3 else sv.submit(*x)
        ids[rid] = i
    results = []
    deadline = time.time() + 60
    while len(results) < len(ids) and time.time() < deadline:
        results += sv.poll(wait_ms=100)
    assert len(results) == len(ids)
    assert max(r["batch_size"] for r in results) > 1
    for r in results:
        i = ids[r["id"]]
        _same(_plain(gpu_stream, *inputs[i]), sv.products(r["id"]), f"pinned event {i}")
    class Fake:  # pageable numpy memory
        words, fed_ids = inputs[0][0].copy(), inputs[0][1].copy()
    with pytest.raises(RuntimeError, match="mapped pinned"):
        sv.submit_pinned(Fake())
    sv.close()
    for pe in pinned:
        pe.close()


def test_server_budget_limits_batch_words(detector, gpu_ctx, gpu_stream):
    """A small HBM budget shrinks the slots' word capacity; events still complete, alone or batched."""
    from pxtrack import gpu
    raws = [event(detector, "pu50", 60 + i)[0] for i in range(6)]
    inputs = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in raws]
    # room for 4 events but the words of only 2 full-size ones
    sv = gpu.Server(gpu_ctx, max_batch=4, n_streams=1, hbm_budget=gpu.pool_bytes(4, 2 * 300000))
    cap = sv.info()["max_batch_words"]
    assert cap < 4 * 300000
    ids = {sv.submit(*x): i for i, x in enumerate(inputs)}
    results = []
    deadline = time.time() + 60
    while len(results) < len(ids) and time.time() < deadline:
        results += sv.poll(wait_ms=100)
    assert len(results) == len(ids)
    for r in results:
        assert r["n_words"] == inputs[ids[r["id"]]][0].size
        _same(_plain(gpu_stream, *inputs[ids[r["id"]]]), sv.products(r["id"]), f"budget event {ids[r['id']]}")
    sv.close()


def test_server_keeps_each_requests_beamspot(detector, gpu_ctx, gpu_stream):
    """Requests batched together keep their own beam spot: the same event submitted with
    two beam spots gives, for each, the plain path's products under that beam spot, and
    the two results differ."""
    from pxtrack import gpu
    words, fed = gpu.unpack_raw(event(detector, "pu50", 21, pu_lo=20, pu_hi=20)[0], gpu_ctx.fed_ids)
    bs0 = gpu_ctx.beamspot
    bs1 = gpu.BeamSpot(*[getattr(bs0, n) for n, _ in gpu.BeamSpot._fields_])
    bs1.x, bs1.y, bs1.z = bs0.x + 0.05, bs0.y - 0.03, bs0.z + 0.4
    sv = gpu.Server(gpu_ctx, max_batch=8, n_streams=1)
    which = {}
    for k in range(9):  # the first request occupies the slot, the next ones wait and share a batch
        which[sv.submit(words, fed, beamspot=bs1 if k % 2 else bs0)] = k % 2
    results = []
    deadline = time.time() + 60
    while len(results) < len(which) and time.time() < deadline:
        results += sv.poll(wait_ms=100)
    assert len(results) == len(which)
    launches = {}
    for r in results:
        launches.setdefault(r["t_launch_ns"], set()).add(which[r["id"]])
    assert any(v == {0, 1} for v in launches.values()), "no batch held both beam spots"
    alone = [_plain(gpu_stream, words, fed), gpu_stream.process(words, fed, upto=5, beamspot=bs1)]
    assert not np.array_equal(alone[0]["state"].view(np.uint32), alone[1]["state"].view(np.uint32))
    for r in results:
        _same(alone[which[r["id"]]], sv.products(r["id"]), f"beam spot {which[r['id']]} (batch of {r['batch_size']})")
    sv.close()


def _serve(sv, inputs, rate, seed):
    """Submits inputs with Poisson arrivals at `rate` events/s; returns the results by input."""
    rng = np.random.default_rng(seed)
    ids, t = {}, time.perf_counter()
    for i in rng.permutation(len(inputs)):
        t += rng.exponential(1.0 / rate) if rate else 0.0
        while time.perf_counter() < t:
            pass
        ids[sv.submit(*inputs[i])] = i
    results = []
    deadline = time.time() + 120
    while len(results) < len(ids) and time.time() < deadline:
        results += sv.poll(wait_ms=100)
    assert len(results) == len(ids), f"{len(results)} of {len(ids)} completed"
    return {ids[r["id"]]: r for r in results}


@pytest.mark.parametrize("budget_gb,n_events,rate", [(64, 256, 10000), (256, 96, None)])
def test_server_poisson60_under_budget(detector, gpu_ctx, gpu_stream, budget_gb, n_events, rate):
    """The serving configuration of the benchmark: pileup Poisson(60) clipped to [10, 200],
    batches of up to 128 events on 3 slots under an HBM budget of 64 GB (Poisson arrivals
    at 10k events/s) and 256 GB (all at once); every event's products equal the plain path's."""
    from pxtrack import gen, gpu
    cfg = gen.preset("pu_poisson60")
    raws = [detector.event(cfg, 1000 + i)[0] for i in range(n_events)]
    inputs = [gpu.unpack_raw(r, gpu_ctx.fed_ids) for r in raws]
    sv = gpu.Server(gpu_ctx, max_batch=128, n_streams=3, hbm_budget=int(budget_gb * 1e9))
    res = _serve(sv, inputs, rate, seed=budget_gb)
    sizes = [r["batch_size"] for r in res.values()]
    print(f"budget {budget_gb} GB: {len(res)} events in {sv.info()['batches']} batches, largest {max(sizes)}",
          flush=True)
    assert max(sizes) > 1
    for i, r in res.items():
        assert r["n_words"] == inputs[i][0].size
        alone = _plain(gpu_stream, *inputs[i])
        assert r["status"] == alone["status"], f"event {i}: status {r['status']} != {alone['status']}"
        _same(alone, sv.products(r["id"]), f"poisson60 event {i} (batch of {r['batch_size']}, {budget_gb} GB)")
    sv.close()

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

"""pxg_raw_to_cluster_noalloc: the products of pxg_raw_to_cluster for one event from device
inputs into the caller's device arrays, with the kernels' scratch in a static device array
(no device allocation by the library).  Two calls in a row, the second on a smaller event,
are compared bit for bit with pxg_raw_to_cluster on the same events; a size past the
scratch array's capacity is rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest

from conftest import event

pytestmark = pytest.mark.gpu

MAX_WORDS = 300000
DIGI_TYPES = {"xx": np.uint16, "yy": np.uint16, "adc": np.uint16, "moduleInd": np.uint16, "clus": np.int32,
              "pdigi": np.uint32, "rawIdArr": np.uint32}
TORCH_OF = {np.uint16: "int16", np.int32: "int32", np.uint32: "int32"}


def _run_noalloc(gpu_ctx, words, feds):
    """One call through the C ABI; outputs start as a 0xA5 byte pattern so that anything the
    call leaves unwritten shows."""
    import torch
    from pxtrack import gpu
    n = len(words)
    dev = torch.device("cuda")
    d_words = torch.from_numpy(np.ascontiguousarray(words, np.uint32).view(np.int32)).to(dev)
    d_feds = torch.from_numpy(np.ascontiguousarray(feds, np.uint8)).to(dev)
    fill = lambda count, dt: torch.full((count,), -0x5a5a5a5b if dt == "int32" else -0x5a5b,  # noqa: E731
                                        dtype=getattr(torch, dt), device=dev)
    digi = {k: fill(max(n, 1), TORCH_OF[t]) for k, t in DIGI_TYPES.items()}
    clus = {k: fill(2001, "int32") for k in ("moduleStart", "clusInModule", "moduleId", "clusModuleStart")}
    errs = torch.full((12 * max(n, 1),), 0xA5, dtype=torch.uint8, device=dev)
    d = gpu.Digis(*[digi[k].data_ptr() for k in DIGI_TYPES], 0, 0)
    c = gpu.Clusters(*[clus[k].data_ptr() for k in ("moduleStart", "clusInModule", "moduleId", "clusModuleStart")],
                     0)
    e = gpu.Errors(errs.data_ptr(), 0)
    stream = torch.cuda.current_stream().cuda_stream
    st = gpu_ctx.lib.pxg_raw_to_cluster_noalloc(gpu_ctx._h, stream, d_words.data_ptr(), d_feds.data_ptr(), n,
                                                C.byref(d), C.byref(c), C.byref(e))
    assert st >= 0, gpu.last_error()
    out = {k: digi[k].cpu().numpy().view(t)[:d.nDigis] for k, t in DIGI_TYPES.items()}
    for k, size in gpu.CLUSTER_FIELDS.items():
        out[k] = clus[k].cpu().numpy().view(np.uint32)[:size]
    out["errors"] = errs.cpu().numpy()[:12 * e.size]
    out.update(nModules=d.nModules, nDigis=d.nDigis, nClusters=c.nClusters, status=st)
    return out


def test_noalloc_matches_raw_to_cluster_twice(detector, gpu_ctx, gpu_stream):
    from pxtrack import gen, gpu
    big = gpu.unpack_raw(detector.event(gen.preset("pu50"), 7)[0], gpu_ctx.fed_ids)
    small = gpu.unpack_raw(event(detector, "serial100", 3)[0], gpu_ctx.fed_ids)
    assert 0 < len(small[0]) < len(big[0]) // 2
    for name, (words, feds) in (("pu50", big), ("serial100", small)):
        print(f"noalloc {name}: {len(words)} words", flush=True)
        got = _run_noalloc(gpu_ctx, words, feds)
        want = gpu_stream.process(words, feds, upto=2)
        # moduleStart holds nModules + 1 entries (the rest is not written by either path)
        nm = int(want["nModules"])
        got["moduleStart"], want["moduleStart"] = got["moduleStart"][:nm + 1], want["moduleStart"][:nm + 1]
        for k in list(DIGI_TYPES) + list(gpu.CLUSTER_FIELDS) + ["errors", "nModules", "nDigis", "nClusters",
                                                                 "status"]:
            assert np.array_equal(np.asarray(got[k]), np.asarray(want[k])), (name, k)
        assert got["nClusters"] > 0


def test_noalloc_rejects_sizes_past_its_scratch(gpu_ctx):
    import torch
    from pxtrack import gpu
    buf = torch.zeros(16, dtype=torch.int32, device="cuda")
    d = gpu.Digis(*([buf.data_ptr()] * 7), 0, 0)
    c = gpu.Clusters(*([buf.data_ptr()] * 4), 0)
    e = gpu.Errors(buf.data_ptr(), 0)
    st = gpu_ctx.lib.pxg_raw_to_cluster_noalloc(gpu_ctx._h, None, buf.data_ptr(), buf.data_ptr(), MAX_WORDS + 2,
                                                C.byref(d), C.byref(c), C.byref(e))
    assert st == -3 and "capacity" in gpu.last_error()

"""The shared device primitives (pixeltrack-standalone_amd/csrc/primitives.h) on
their own, through the self-test kernels of tests/native: the wavefront and
workgroup prefix scans against sequential scans for sizes around every wavefront
and workgroup boundary, in 32 and 64 bits (the reference's
src/cuda/test/prefixScan_t.cu:22-60 checks its warp and block scans the same
way), the one-workgroup array scan up to 300k entries, the per-wavefront counter
allocation and sums, and the count -> scan -> fill one-to-many container against
a bucket sort of the same entries, within and past its capacity (the checks of
test/HistoContainer_t_cpu.cc:11-145 and test/OneToManyAssoc_t.h:89-212)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu
LIB = os.path.join(REPO, "tests", "native", "libpxprim_selftest.so")


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available()
    lib = C.CDLL(LIB)
    vp, u32 = C.c_void_p, C.c_uint32
    lib.pxs_wave.argtypes = [vp, vp, vp, vp, vp, u32]
    lib.pxs_wg_scan.argtypes = [C.c_int, C.c_int, vp, vp, vp, u32]
    lib.pxs_scan_array.argtypes = [C.c_int, vp, u32]
    lib.pxs_wave_alloc.argtypes = [vp, vp, vp, u32]
    lib.pxs_wave_add.argtypes = [vp, vp, u32]
    lib.pxs_one_to_many.argtypes = [vp, u32, u32, u32, vp, vp, vp, vp, vp]
    return lib


def dev(a):
    import torch
    a = np.ascontiguousarray(a)
    view = {np.uint32: np.int32, np.uint64: np.int64, np.uint8: np.uint8}[a.dtype.type]
    return torch.from_numpy(a.view(view)).cuda()


def host(t, dtype):
    return t.cpu().numpy().view(dtype)


def test_wave_scans(lib):
    import torch
    rng = np.random.default_rng(1)
    for n in (1, 2, 63, 64, 65, 127, 128, 129, 1000, 4096 + 17):
        x = rng.integers(0, 1 << 20, n, dtype=np.uint32)
        nw = (n + 255) // 256 * 4
        incl, excl = torch.zeros(n, dtype=torch.int32, device="cuda"), torch.zeros(n, dtype=torch.int32, device="cuda")
        wmax, wsum = torch.zeros(nw, dtype=torch.int32, device="cuda"), torch.zeros(nw, dtype=torch.int32,
                                                                                    device="cuda")
        d = dev(x)
        assert lib.pxs_wave(d.data_ptr(), incl.data_ptr(), excl.data_ptr(), wmax.data_ptr(), wsum.data_ptr(), n) == 0
        pad = np.zeros(nw * 64, np.uint64)
        pad[:n] = x
        chunks = pad.reshape(-1, 64)
        want_incl = np.cumsum(chunks, axis=1).ravel()[:n]
        assert np.array_equal(host(incl, np.uint32), want_incl.astype(np.uint32)), n
        assert np.array_equal(host(excl, np.uint32), (want_incl - x).astype(np.uint32)), n
        used = (n + 63) // 64
        assert np.array_equal(host(wmax, np.uint32)[:used], chunks.max(axis=1)[:used].astype(np.uint32)), n
        assert np.array_equal(host(wsum, np.uint32)[:used], chunks.sum(axis=1)[:used].astype(np.uint32)), n


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("nt", [64, 256, 640, 1024])
def test_workgroup_scans(lib, nt, bits):
    import torch
    rng = np.random.default_rng(nt + bits)
    dt = np.uint32 if bits == 32 else np.uint64
    top = (1 << 20) if bits == 32 else (1 << 40)
    for n in sorted({1, 63, 64, 65, nt - 1, nt, nt + 1, 3 * nt + 5, 20000}):
        x = rng.integers(0, top, n, dtype=dt)
        nb = (n + nt - 1) // nt
        tdt = torch.int32 if bits == 32 else torch.int64
        out, tot = torch.zeros(n, dtype=tdt, device="cuda"), torch.zeros(nb, dtype=tdt, device="cuda")
        d = dev(x)
        assert lib.pxs_wg_scan(nt, bits, d.data_ptr(), out.data_ptr(), tot.data_ptr(), n) == 0
        pad = np.zeros(nb * nt, np.uint64)
        pad[:n] = x
        chunks = pad.reshape(nb, nt)
        excl = (np.cumsum(chunks, axis=1) - chunks).ravel()[:n]
        assert np.array_equal(host(out, dt), excl.astype(dt)), (nt, bits, n)
        assert np.array_equal(host(tot, dt), chunks.sum(axis=1).astype(dt)), (nt, bits, n)


@pytest.mark.parametrize("nt", [256, 1024])
def test_array_scan(lib, nt):
    rng = np.random.default_rng(nt)
    for n in (0, 1, nt - 1, nt, nt + 1, 49152, 300000):
        x = rng.integers(0, 200, n + 1, dtype=np.uint32)
        d = dev(x)
        assert lib.pxs_scan_array(nt, d.data_ptr(), n) == 0
        got = host(d, np.uint32)
        want = np.concatenate([[0], np.cumsum(x[:n], dtype=np.uint64)]).astype(np.uint32)
        assert np.array_equal(got, want), n


def test_wave_alloc_and_add(lib):
    import torch
    rng = np.random.default_rng(3)
    for n in (1, 64, 100, 5000):
        need = (rng.random(n) < 0.4).astype(np.uint8)
        counter = torch.zeros(1, dtype=torch.int32, device="cuda")
        slot = torch.zeros(n, dtype=torch.int32, device="cuda")
        assert lib.pxs_wave_alloc(counter.data_ptr(), dev(need).data_ptr(), slot.data_ptr(), n) == 0
        s = host(slot, np.uint32)
        k = int(need.sum())
        assert int(counter.item()) == k
        taken = s[need == 1]
        assert np.array_equal(np.sort(taken), np.arange(k, dtype=np.uint32)), n
        assert (s[need == 0] == 0xffffffff).all()
        # within a wavefront the slots follow lane order
        for w in range(0, n, 64):
            ts = s[w:w + 64][need[w:w + 64] == 1]
            assert np.array_equal(ts, np.sort(ts))
        v = rng.integers(0, 1000, n, dtype=np.uint32)
        v[rng.random(n) < 0.3] = 0
        counter.zero_()
        assert lib.pxs_wave_add(counter.data_ptr(), dev(v).data_ptr(), n) == 0
        assert int(counter.item()) == int(v.sum())


@pytest.mark.parametrize("n,nkeys,capacity", [(0, 10, 16), (1000, 7, 4096), (200000, 49152, 1 << 20),
                                              (50000, 300, 20000), (100000, 1, 100000)])
def test_one_to_many(lib, n, nkeys, capacity):
    import torch
    rng = np.random.default_rng(n + nkeys)
    keys = (rng.zipf(1.3, n) % nkeys).astype(np.uint32) if n else np.zeros(0, np.uint32)
    count = torch.zeros(nkeys, dtype=torch.int32, device="cuda")
    off = torch.zeros(nkeys + 1, dtype=torch.int32, device="cuda")
    content = torch.full((capacity,), -1, dtype=torch.int32, device="cuda")
    total = torch.zeros(1, dtype=torch.int32, device="cuda")
    dropped = torch.zeros(1, dtype=torch.int32, device="cuda")
    dk = dev(keys) if n else torch.zeros(1, dtype=torch.int32, device="cuda")
    assert lib.pxs_one_to_many(dk.data_ptr(), n, nkeys, capacity, count.data_ptr(), off.data_ptr(),
                               content.data_ptr(), total.data_ptr(), dropped.data_ptr()) == 0
    counts = np.bincount(keys, minlength=nkeys).astype(np.int64)
    want_off = np.minimum(np.concatenate([[0], np.cumsum(counts)]), capacity)
    o = host(off, np.uint32).astype(np.int64)
    assert np.array_equal(o, want_off)
    assert int(total.item()) == n
    assert int(dropped.item()) == max(0, n - capacity)
    c = host(content, np.uint32)
    # every stored entry sits under its own key, once
    stored = c[:o[-1]]
    assert len(np.unique(stored)) == len(stored)
    for k in np.nonzero(counts)[0][:2000]:
        vals = c[o[k]:o[k + 1]]
        assert (keys[vals] == k).all(), k
        if o[k + 1] - o[k] == counts[k]:
            assert np.array_equal(np.sort(vals), np.nonzero(keys == k)[0].astype(np.uint32)), k

"""ctypes binding of libpxgpu.so (include/pxgpu.h) and of the host FED loop.

No fallback: if the HIP library is not built, or no GPU is present, these
calls raise.
"""
import ctypes as C

import numpy as np

from ._lib import load_library

vp = C.c_void_p
u32 = C.c_uint32


class Conditions(C.Structure):
    _fields_ = [("cabling", vp), ("mod_to_unp", vp), ("gain", vp), ("gain_data", vp), ("gain_data_bytes", u32),
                ("cpe_common", vp), ("cpe_det", vp), ("n_det", u32), ("cpe_layer", vp), ("cpe_average", vp)]


class BeamSpot(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("x", "y", "z", "sigmaZ", "beamWidthX", "beamWidthY", "dxdz", "dydz",
                                         "emittanceX", "emittanceY", "betaStar")]


class Digis(C.Structure):
    _fields_ = [("xx", vp), ("yy", vp), ("adc", vp), ("moduleInd", vp), ("clus", vp), ("pdigi", vp),
                ("rawIdArr", vp), ("nModules", u32), ("nDigis", u32)]


class Clusters(C.Structure):
    _fields_ = [("moduleStart", vp), ("clusInModule", vp), ("moduleId", vp), ("clusModuleStart", vp),
                ("nClusters", u32)]


class Errors(C.Structure):
    _fields_ = [("data", vp), ("size", u32)]


class Hits(C.Structure):
    _fields_ = [(n, vp) for n in ("xl", "yl", "xerr", "yerr", "xg", "yg", "zg", "rg", "iphi", "charge", "xsize",
                                  "ysize", "detInd", "hitsModuleStart", "hitsLayerStart", "phiOff", "phiBins",
                                  "averageGeometry")] + [("nHits", u32)]


class RawView(C.Structure):
    _fields_ = [("words", vp), ("fedIds", vp), ("nWords", u32)]


class Tracks(C.Structure):
    _fields_ = [(n, vp) for n in ("hitOff", "hitIndices", "detIndices", "quality", "chi2", "eta", "pt", "state",
                                  "covariance")] + [("stride", u32), ("nTuples", u32)]


class Vertices(C.Structure):
    _fields_ = [(n, vp) for n in ("idv", "zv", "wv", "chi2", "ptv2", "ndof", "sortInd")] + [("nvFinal", u32)]


class CAInternals(C.Structure):
    _fields_ = [(n, vp) for n in ("cellInner", "cellOuter", "cellPair", "cellFlags", "cellUsed", "outerHitOff",
                                  "outerHitCells", "neighborSlot", "neighborCount", "neighborPool", "trackSlot",
                                  "trackCount", "trackPool", "multOff", "multBins", "hitToTupleOff",
                                  "hitToTupleBins")] + [("counts", u32 * 20)]


class LaunchShape(C.Structure):
    _fields_ = [("doublet_lanes", u32), ("grid_target", u32), ("max_blocks_per_event", u32), ("connect_stage", u32),
                ("fishbone_window", u32)]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch", u32), ("n_streams", u32), ("hbm_budget", C.c_uint64), ("profile", C.c_int),
                ("upto", C.c_int), ("doublet_lanes", u32), ("grid_target", u32), ("transfer", C.c_int)]


class ServerResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("t_submit_ns", C.c_uint64), ("t_launch_ns", C.c_uint64),
                ("t_done_ns", C.c_uint64), ("batch_size", u32), ("n_words", u32), ("n_hits", u32),
                ("n_tuples", u32), ("n_vertices", u32)]


CA_COUNTS = ["nCells", "nDoublets", "apcM", "apcN", "nTuples", "caStatus", "nbLists", "trLists", "nMulti", "nMult",
             "nVertices", "vtxTracks", "vstatus", "nSlots", "nNeighbors", "nStart", "nPaths"]
PROFILE_REFERENCE, PROFILE_HIGH_OCCUPANCY = 0, 1
BFIELD = 0.0114256972711507  # 1/GeV, CAHitNtupletCUDA.cc:31
VERTEX_PTMIN = 0.5           # GeV, PixelVertexProducerCUDA.cc:37
VERTEX_DENSITY, VERTEX_DBSCAN, VERTEX_ITERATIVE = 0, 1, 2  # pxg_set_vertex_algorithm


DIGI_FIELDS = {"xx": np.uint16, "yy": np.uint16, "adc": np.uint16, "moduleInd": np.uint16, "clus": np.int32,
               "pdigi": np.uint32, "rawIdArr": np.uint32}
CLUSTER_FIELDS = {"moduleStart": 2001, "clusInModule": 2000, "moduleId": 2000, "clusModuleStart": 2001}
HIT_FIELDS = {"xl": np.float32, "yl": np.float32, "xerr": np.float32, "yerr": np.float32, "xg": np.float32,
              "yg": np.float32, "zg": np.float32, "rg": np.float32, "iphi": np.int16, "charge": np.int32,
              "xsize": np.int16, "ysize": np.int16, "detInd": np.uint16}


def _gpu():
    lib = load_library("libpxgpu.so")
    if not getattr(lib, "_typed", False):
        lib.pxg_last_error.restype = C.c_char_p
        lib.pxg_init.argtypes = [C.POINTER(Conditions), C.POINTER(vp)]
        lib.pxg_finalize.argtypes = [vp]
        lib.pxg_device_name.argtypes = [C.c_char_p, C.c_size_t]
        lib.pxg_stream_create.argtypes = [vp, u32, u32, C.POINTER(vp)]
        lib.pxg_stream_create_profile.argtypes = [vp, u32, u32, C.c_int, C.POINTER(vp)]
        lib.pxg_ntuplets.argtypes = [vp, C.POINTER(Hits), C.c_float, C.POINTER(Tracks)]
        lib.pxg_find_vertices.argtypes = [vp, C.POINTER(Tracks), C.c_float, C.POINTER(Vertices)]
        lib.pxg_event_tracks.argtypes = [vp, u32, C.POINTER(Tracks), C.POINTER(Vertices)]
        lib.pxg_event_ca_internals.argtypes = [vp, u32, C.POINTER(CAInternals)]
        lib.pxg_pool_bytes.restype = C.c_size_t
        lib.pxg_pool_bytes.argtypes = [u32, u32, C.c_int]
        lib.pxg_profile_capacities.argtypes = [C.c_int, C.POINTER(u32 * 6)]
        lib.pxg_raw_to_cluster_noalloc.argtypes = [vp, vp, vp, vp, u32, C.POINTER(Digis), C.POINTER(Clusters),
                                                   C.POINTER(Errors)]
        lib.pxg_server_create.argtypes = [vp, C.POINTER(ServerConfig), C.POINTER(vp)]
        lib.pxg_server_destroy.argtypes = [vp]
        lib.pxg_server_submit.argtypes = [vp, vp, vp, u32, C.POINTER(BeamSpot), C.POINTER(C.c_uint64)]
        lib.pxg_server_submit_pinned.argtypes = [vp, vp, vp, u32, C.POINTER(BeamSpot), C.POINTER(C.c_uint64)]
        lib.pxg_host_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
        lib.pxg_host_free.argtypes = [vp]
        lib.pxg_server_poll.argtypes = [vp, C.POINTER(ServerResult), C.c_int, C.c_int]
        lib.pxg_server_products.argtypes = [vp, C.c_uint64, C.POINTER(vp), C.POINTER(u32)]
        lib.pxg_server_release.argtypes = [vp, C.c_uint64]
        lib.pxg_server_info.argtypes = [vp, C.POINTER(u32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        lib.pxg_set_tuning.argtypes = [vp, u32, u32]
        lib.pxg_set_launch_shape.argtypes = [vp, C.POINTER(LaunchShape)]
        lib.pxg_get_launch_shape.argtypes = [vp, C.POINTER(LaunchShape)]
        lib.pxg_autotune_shape.argtypes = [vp, u32, u32, C.c_int, C.POINTER(LaunchShape), C.POINTER(C.c_float)]
        lib.pxg_set_vertex_algorithm.argtypes = [vp, C.c_int]
        lib.pxg_set_fishbone.argtypes = [vp, C.c_int, C.c_int]
        lib.pxg_set_riemann_fit.argtypes = [vp, C.c_int]
        lib.pxg_set_bfield.argtypes = [vp, C.c_float]
        lib.pxg_set_vertex_ptmin.argtypes = [vp, C.c_float]
        lib.pxg_set_stats.argtypes = [vp, C.c_int]
        lib.pxg_set_graphs.argtypes = [vp, C.c_int]
        lib.pxg_get_counters.argtypes = [vp, C.POINTER(C.c_uint64)]
        lib.pxg_reset_counters.argtypes = [vp]
        lib.pxg_print_counters.argtypes = [vp]
        lib.pxg_get_tuning.argtypes = [vp, C.POINTER(u32), C.POINTER(u32)]
        lib.pxg_autotune.argtypes = [vp, u32, u32, C.c_int, C.POINTER(u32), C.POINTER(u32),
                                     C.POINTER(C.c_float)]
        lib.pxg_stream_destroy.argtypes = [vp]
        lib.pxg_sync.argtypes = [vp]
        lib.pxg_raw_to_digi.argtypes = [vp, vp, vp, u32, C.POINTER(Digis), C.POINTER(Errors)]
        lib.pxg_raw_to_cluster.argtypes = [vp, vp, vp, u32, C.POINTER(Digis), C.POINTER(Clusters),
                                           C.POINTER(Errors)]
        lib.pxg_rechits.argtypes = [vp, C.POINTER(Digis), C.POINTER(Clusters), C.POINTER(BeamSpot), C.POINTER(Hits)]
        lib.pxg_load_events.argtypes = [vp, u32, C.POINTER(RawView), C.POINTER(BeamSpot)]
        lib.pxg_run.argtypes = [vp, u32, u32, C.c_int]
        lib.pxg_fetch_counts.argtypes = [vp, u32, u32]
        lib.pxg_event_products.argtypes = [vp, u32, C.POINTER(Digis), C.POINTER(Clusters), C.POINTER(Errors),
                                           C.POINTER(Hits)]
        lib.pxg_copy_to_host.argtypes = [vp, vp, vp, C.c_size_t]
        lib.pxg_event_counts.argtypes = [vp, u32, C.POINTER(u32)]
        lib.pxg_profile_enable.argtypes = [vp, C.c_int]
        lib.pxg_profile_read.argtypes = [vp, C.POINTER(C.c_char_p), C.POINTER(u32), C.POINTER(C.c_double), C.c_int]
        lib._typed = True
    return lib


def _host():
    lib = load_library("libpxhost.so")
    if not getattr(lib, "_typed", False):
        lib.pxh_unpack_raw_record.argtypes = [vp, C.c_size_t, vp, u32, vp, vp, u32, C.POINTER(u32)]
        lib._typed = True
    return lib


def unpack_raw(raw, fed_ids, capacity=300000):
    """Host FED loop of the raw-to-cluster producer: raw record -> (words, fedIds)."""
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    fed_ids = np.ascontiguousarray(fed_ids, dtype=np.uint32)
    words = np.zeros(capacity, dtype=np.uint32)
    fed = np.zeros(capacity // 2, dtype=np.uint8)
    n = u32()
    rc = _host().pxh_unpack_raw_record(raw.ctypes.data, raw.size, fed_ids.ctypes.data, fed_ids.size, words.ctypes.data,
                                       fed.ctypes.data, capacity, C.byref(n))
    if rc != 0:
        raise RuntimeError(f"FED unpacking failed ({rc})")
    return words[: n.value].copy(), fed[: n.value // 2].copy()


def _check(rc, what):
    if rc < 0:
        raise RuntimeError(f"{what}: {_gpu().pxg_last_error().decode()}")
    return rc


def last_error():
    """The message of the library's last failed call (pxg_last_error)."""
    return _gpu().pxg_last_error().decode()


def device_name():
    buf = C.create_string_buffer(256)
    _check(_gpu().pxg_device_name(buf, 256), "pxg_device_name")
    return buf.value.decode()


class Context:
    """Device context with the conditions uploaded once."""

    def __init__(self, cond):
        self.lib = _gpu()
        self._keep = cond
        c = Conditions()
        c.cabling = cond["cabling"].ctypes.data
        c.mod_to_unp = cond["mod_to_unp"].ctypes.data
        c.gain = cond["gain"].ctypes.data
        c.gain_data = cond["gain_data"].ctypes.data
        c.gain_data_bytes = cond["gain_data"].size
        c.cpe_common = cond["cpe_common"].ctypes.data
        c.cpe_det = cond["cpe_det"].ctypes.data
        c.n_det = cond["n_det"]
        c.cpe_layer = cond["cpe_layer"].ctypes.data
        c.cpe_average = cond["cpe_average"].ctypes.data
        self._h = vp()
        _check(self.lib.pxg_init(C.byref(c), C.byref(self._h)), "pxg_init")
        self.fed_ids = cond["fed_ids"]
        bs = np.frombuffer(cond["beamspot"].tobytes(), dtype=np.float32)
        self.beamspot = BeamSpot(*[float(v) for v in bs])

    def close(self):
        if getattr(self, "_h", None):
            self.lib.pxg_finalize(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def stream(self, max_events=1, max_words=300000, profile=PROFILE_REFERENCE):
        return Stream(self, max_events, max_words, profile)


class Stream:
    """One HIP stream with its device-resident product pool."""

    def __init__(self, ctx, max_events, max_words, profile=PROFILE_REFERENCE):
        self.ctx, self.lib = ctx, ctx.lib
        self._h = vp()
        _check(self.lib.pxg_stream_create_profile(ctx._h, max_events, max_words, profile, C.byref(self._h)),
               "pxg_stream_create_profile")
        self.caps = capacities(profile)
        self.bfield, self.vertex_ptmin = BFIELD, VERTEX_PTMIN
        self._inputs = []

    def close(self):
        if getattr(self, "_h", None):
            self.lib.pxg_stream_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # -- device -> host copies of products
    def _copy(self, ptr, dtype, n):
        out = np.zeros(n, dtype=dtype)
        if n:
            _check(self.lib.pxg_copy_to_host(self._h, out.ctypes.data, ptr, out.nbytes), "pxg_copy_to_host")
        return out

    def digis_to_host(self, d):
        out = {k: self._copy(getattr(d, k), t, d.nDigis) for k, t in DIGI_FIELDS.items()}
        out["nModules"], out["nDigis"] = d.nModules, d.nDigis
        return out

    def clusters_to_host(self, c):
        out = {k: self._copy(getattr(c, k), np.uint32, n) for k, n in CLUSTER_FIELDS.items()}
        out["nClusters"] = c.nClusters
        return out

    def errors_to_host(self, e):
        return self._copy(e.data, np.uint8, 12 * e.size)

    def hits_to_host(self, h):
        out = {k: self._copy(getattr(h, k), t, h.nHits) for k, t in HIT_FIELDS.items()}
        out["hitsLayerStart"] = self._copy(h.hitsLayerStart, np.uint32, 11)
        out["phiOff"] = self._copy(h.phiOff, np.uint32, 1281)
        out["phiBins"] = self._copy(h.phiBins, np.uint16, h.nHits)
        out["avgCorrected"] = self._copy(h.averageGeometry, np.float32, 6 * 148 + 2)
        out["nHits"] = h.nHits
        return out

    def tracks_to_host(self, t):
        """Track products in the oracle's layout (per-tuple rows for state and covariance)."""
        S, n = t.stride, t.nTuples
        out = {"tupleOff": self._copy(t.hitOff, np.uint32, S + 1)}
        nb = int(out["tupleOff"][S])
        out["tupleBins"] = self._copy(t.hitIndices, np.uint16, nb)
        out["detIndices"] = self._copy(t.detIndices, np.uint16, nb)
        out["quality"] = self._copy(t.quality, np.uint8, S)
        for k in ("chi2", "eta", "pt"):
            out[k] = self._copy(getattr(t, k), np.float32, S)
        out["state"] = self._copy(t.state, np.float32, 5 * S).reshape(5, S).T.copy().ravel()
        out["covariance"] = self._copy(t.covariance, np.float32, 15 * S).reshape(15, S).T.copy().ravel()
        out["nTuples"] = n
        return out

    def vertices_to_host(self, v, stride):
        out = {"idv": self._copy(v.idv, np.int16, stride), "ndof": self._copy(v.ndof, np.int32, stride)}
        for k, name in (("zv", "zv"), ("wv", "wv"), ("chi2", "chi2v"), ("ptv2", "ptv2")):
            out[name] = self._copy(getattr(v, k), np.float32, 1024)
        out["sortInd"] = self._copy(v.sortInd, np.uint16, 1024)
        out["nVertices"] = v.nvFinal
        return out

    def ca_counts(self, e):
        """Counters of the ntuplet and vertex stages of resident event e (after fetch_counts)."""
        ci = CAInternals()
        _check(self.lib.pxg_event_ca_internals(self._h, e, C.byref(ci)), "pxg_event_ca_internals")
        return dict(zip(CA_COUNTS, list(ci.counts)))

    def ca_internals(self, e, n_hits):
        """The ntuplet stage's containers of resident event e, in the oracle's layout."""
        ci = CAInternals()
        _check(self.lib.pxg_event_ca_internals(self._h, e, C.byref(ci)), "pxg_event_ca_internals")
        cnt = dict(zip(CA_COUNTS, list(ci.counts)))
        nc = cnt["nCells"]
        out = {"cellInner": self._copy(ci.cellInner, np.uint16, nc), "cellOuter": self._copy(ci.cellOuter, np.uint16, nc)}
        out["cellLayerPair"] = self._copy(ci.cellPair, np.uint8, nc).astype(np.int16)
        flags = self._copy(ci.cellFlags, np.uint8, nc)
        out["cellDoubletId"] = np.where(flags & 2, -1, np.arange(nc)).astype(np.int32)
        out["cellUsed"] = self._copy(ci.cellUsed, np.uint8, nc).astype(np.uint16)
        ohoff = self._copy(ci.outerHitOff, np.uint32, n_hits + 1)
        ohc = self._copy(ci.outerHitCells, np.uint32, 2 * int(ohoff[-1]) if n_hits else 0)[0::2]
        # the device keeps each hit's cells in fill order; the serial container holds the
        # first maxCellsPerHit in ascending order
        cap = self.caps["max_cells_per_hit"]
        lists = [np.sort(ohc[ohoff[h]:ohoff[h + 1]])[:cap] for h in range(n_hits)]
        out["outerHitOff"] = np.concatenate([[0], np.cumsum([len(x) for x in lists])]).astype(np.uint32)
        out["outerHitCells"] = np.concatenate(lists).astype(np.uint32) if lists else np.zeros(0, np.uint32)

        def csr(slot_p, count_p, pool_p, width, dtype, count_dtype):
            slot = self._copy(slot_p, np.int32, nc)
            count = self._copy(count_p, count_dtype, nc).astype(np.int64)
            count[slot < 0] = 0
            npool = int(slot.max()) + 1 if nc and slot.max() >= 0 else 0
            pool = self._copy(pool_p, dtype, npool * width).reshape(npool, width) if npool else None
            vals = [pool[slot[c], :count[c]] for c in range(nc) if count[c]]
            off = np.concatenate([[0], np.cumsum(count)]).astype(np.uint32)
            return off, (np.concatenate(vals).astype(dtype) if vals else np.zeros(0, dtype))

        out["cellNeighborOff"], out["cellNeighbors"] = csr(ci.neighborSlot, ci.neighborCount, ci.neighborPool, 36,
                                                           np.uint32, np.uint8)
        out["cellTrackOff"], out["cellTracks"] = csr(ci.trackSlot, ci.trackCount, ci.trackPool, 48, np.uint32,
                                                     np.uint16)
        out["multOff"] = self._copy(ci.multOff, np.uint32, 9)
        out["multBins"] = self._copy(ci.multBins, np.uint32, int(out["multOff"][8]))
        out["hitToTupleOff"] = self._copy(ci.hitToTupleOff, np.uint32, n_hits + 1)
        out["hitToTupleBins"] = self._copy(ci.hitToTupleBins, np.uint32, int(out["hitToTupleOff"][-1]) if n_hits else 0)
        out.update(cnt)
        return out

    # -- plain per-event path
    def raw_to_cluster(self, words, fed_ids):
        words = np.ascontiguousarray(words, dtype=np.uint32)
        fed_ids = np.ascontiguousarray(fed_ids, dtype=np.uint8)
        d, c, e = Digis(), Clusters(), Errors()
        st = _check(self.lib.pxg_raw_to_cluster(self._h, words.ctypes.data, fed_ids.ctypes.data, words.size,
                                                C.byref(d), C.byref(c), C.byref(e)), "pxg_raw_to_cluster")
        return st, d, c, e

    def raw_to_digi(self, words, fed_ids):
        words = np.ascontiguousarray(words, dtype=np.uint32)
        fed_ids = np.ascontiguousarray(fed_ids, dtype=np.uint8)
        d, e = Digis(), Errors()
        st = _check(self.lib.pxg_raw_to_digi(self._h, words.ctypes.data, fed_ids.ctypes.data, words.size,
                                             C.byref(d), C.byref(e)), "pxg_raw_to_digi")
        return st, d, e

    def rechits(self, d, c, beamspot=None):
        h = Hits()
        bs = beamspot if beamspot is not None else self.ctx.beamspot
        st = _check(self.lib.pxg_rechits(self._h, C.byref(d), C.byref(c), C.byref(bs), C.byref(h)), "pxg_rechits")
        return st, h

    def ntuplets(self, h):
        t = Tracks()
        st = _check(self.lib.pxg_ntuplets(self._h, C.byref(h), self.bfield, C.byref(t)), "pxg_ntuplets")
        return st, t

    def find_vertices(self, t):
        v = Vertices()
        st = _check(self.lib.pxg_find_vertices(self._h, C.byref(t), self.vertex_ptmin, C.byref(v)),
                    "pxg_find_vertices")
        return st, v

    def process(self, words, fed_ids, upto=3, internals=False, beamspot=None):
        """Plain path on one event; returns a dict of host arrays named like the oracle's fields.
        beamspot: a BeamSpot (default: the conditions' beam spot)."""
        if upto == 1:
            st, d, e = self.raw_to_digi(words, fed_ids)
            out = self.digis_to_host(d)
        else:
            st, d, c, e = self.raw_to_cluster(words, fed_ids)
            out = self.digis_to_host(d)
            out.update(self.clusters_to_host(c))
            if upto >= 3:
                st2, h = self.rechits(d, c, beamspot)
                st |= st2
                out.update(self.hits_to_host(h))
            if upto >= 4:
                st3, t = self.ntuplets(h)
                st |= st3
                out.update(self.tracks_to_host(t))
                if internals:
                    out.update(self.ca_internals(0, h.nHits))
            if upto >= 5:
                st4, v = self.find_vertices(t)
                st |= st4
                out.update(self.vertices_to_host(v, t.stride))
        out["errors"] = self.errors_to_host(e)
        out["status"] = st
        return out

    # -- resident / batched path
    def load(self, events, beamspot=None):
        """events: list of (words, fedIds) host arrays; copies them into the device pool."""
        self._inputs = [(np.ascontiguousarray(w, dtype=np.uint32), np.ascontiguousarray(f, dtype=np.uint8))
                        for w, f in events]
        views = (RawView * len(self._inputs))()
        for i, (w, f) in enumerate(self._inputs):
            views[i].words, views[i].fedIds, views[i].nWords = w.ctypes.data, f.ctypes.data, w.size
        bs = beamspot if beamspot is not None else self.ctx.beamspot
        _check(self.lib.pxg_load_events(self._h, len(self._inputs), views, C.byref(bs)), "pxg_load_events")

    def run(self, first, count, upto=3):
        _check(self.lib.pxg_run(self._h, first, count, upto), "pxg_run")

    def fetch_counts(self, first, count):
        return _check(self.lib.pxg_fetch_counts(self._h, first, count), "pxg_fetch_counts")

    def event_products(self, e, upto=3):
        d, c, er, h = Digis(), Clusters(), Errors(), Hits()
        _check(self.lib.pxg_event_products(self._h, e, C.byref(d), C.byref(c), C.byref(er), C.byref(h)),
               "pxg_event_products")
        out = self.digis_to_host(d)
        if upto >= 2:
            out.update(self.clusters_to_host(c))
        if upto >= 3:
            out.update(self.hits_to_host(h))
        out["errors"] = self.errors_to_host(er)
        if upto >= 4:
            t, v = Tracks(), Vertices()
            _check(self.lib.pxg_event_tracks(self._h, e, C.byref(t), C.byref(v)), "pxg_event_tracks")
            out.update(self.tracks_to_host(t))
            if upto >= 5:
                out.update(self.vertices_to_host(v, t.stride))
        return out

    def event_counts(self, e):
        """(module runs, clusters, error records, status) of resident event e."""
        out = (u32 * 4)()
        _check(self.lib.pxg_event_counts(self._h, e, out), "pxg_event_counts")
        return tuple(out)

    def set_tuning(self, doublet_lanes, grid_target):
        _check(self.lib.pxg_set_tuning(self._h, doublet_lanes, grid_target), "pxg_set_tuning")

    def set_launch_shape(self, shape):
        """shape: dict with the fields of LaunchShape (missing ones keep their value)."""
        cur = self.launch_shape()
        cur.update(shape)
        _check(self.lib.pxg_set_launch_shape(self._h, C.byref(LaunchShape(**cur))), "pxg_set_launch_shape")

    def launch_shape(self):
        h = LaunchShape()
        _check(self.lib.pxg_get_launch_shape(self._h, C.byref(h)), "pxg_get_launch_shape")
        return {f: getattr(h, f) for f, _ in LaunchShape._fields_}

    def autotune_shape(self, first, count, upto=5):
        """Chooses the fastest launch shape on resident events; returns (shape dict, ms per pass)."""
        h, ms = LaunchShape(), C.c_float()
        _check(self.lib.pxg_autotune_shape(self._h, first, count, upto, C.byref(h), C.byref(ms)),
               "pxg_autotune_shape")
        return {f: getattr(h, f) for f, _ in LaunchShape._fields_}, ms.value

    def set_vertex_algorithm(self, algo):
        """VERTEX_DENSITY (default), VERTEX_DBSCAN or VERTEX_ITERATIVE."""
        _check(self.lib.pxg_set_vertex_algorithm(self._h, algo), "pxg_set_vertex_algorithm")

    def set_fishbone(self, early=True, late=False):
        """The fishbone passes: early (before the ntuplet search, default on) and late
        (on the cells in tuples, with the cleaner after classification, default off)."""
        _check(self.lib.pxg_set_fishbone(self._h, int(early), int(late)), "pxg_set_fishbone")

    COUNTERS = ("nEvents", "nHits", "nCells", "nTuples", "nFitTracks", "nGoodTracks", "nUsedHits", "nDupHits",
                "nKilledCells", "nEmptyCells", "nZeroTrackCells")

    def set_graphs(self, on=True):
        """Capture run()'s stage sequences as HIP graphs and replay them."""
        _check(self.lib.pxg_set_graphs(self._h, int(on)), "pxg_set_graphs")

    def set_stats(self, on=True):
        """Accumulate the ntuplet generator's counters (the reference's doStats)."""
        _check(self.lib.pxg_set_stats(self._h, int(on)), "pxg_set_stats")

    def counters(self):
        out = (C.c_uint64 * len(self.COUNTERS))()
        _check(self.lib.pxg_get_counters(self._h, out), "pxg_get_counters")
        return dict(zip(self.COUNTERS, (int(x) for x in out)))

    def reset_counters(self):
        _check(self.lib.pxg_reset_counters(self._h), "pxg_reset_counters")

    def print_counters(self):
        _check(self.lib.pxg_print_counters(self._h), "pxg_print_counters")

    def set_riemann_fit(self, on=True):
        """Riemann fit in place of the broken-line fit (default off)."""
        _check(self.lib.pxg_set_riemann_fit(self._h, int(on)), "pxg_set_riemann_fit")

    def set_bfield(self, bfield=None):
        """Field of the helix fits in 1/GeV (default: CAHitNtupletCUDA's 0.0114256972711507)."""
        b = BFIELD if bfield is None else bfield
        _check(self.lib.pxg_set_bfield(self._h, b), "pxg_set_bfield")
        self.bfield = b

    def set_vertex_ptmin(self, pt=None):
        """Track pt cut of the vertex finder in GeV (default: PixelVertexProducerCUDA's 0.5)."""
        p = VERTEX_PTMIN if pt is None else pt
        _check(self.lib.pxg_set_vertex_ptmin(self._h, p), "pxg_set_vertex_ptmin")
        self.vertex_ptmin = p

    def tuning(self):
        a, b = u32(), u32()
        _check(self.lib.pxg_get_tuning(self._h, C.byref(a), C.byref(b)), "pxg_get_tuning")
        return a.value, b.value

    def autotune(self, first, count, upto=5):
        """Chooses the fastest launch shape on resident events; returns (lanes, grid, ms per pass)."""
        a, b, ms = u32(), u32(), C.c_float()
        _check(self.lib.pxg_autotune(self._h, first, count, upto, C.byref(a), C.byref(b), C.byref(ms)),
               "pxg_autotune")
        return a.value, b.value, ms.value

    def sync(self):
        _check(self.lib.pxg_sync(self._h), "pxg_sync")

    def profile(self, on=True):
        _check(self.lib.pxg_profile_enable(self._h, 1 if on else 0), "pxg_profile_enable")

    def profile_read(self):
        n = 32
        names, calls, ms = (C.c_char_p * n)(), (u32 * n)(), (C.c_double * n)()
        k = _check(self.lib.pxg_profile_read(self._h, names, calls, ms, n), "pxg_profile_read")
        return {names[i].decode(): (calls[i], ms[i]) for i in range(min(k, n))}


def unpack_products(buf):
    """Splits one event's packed products (pxg_server_products layout) into arrays."""
    b = np.frombuffer(buf, dtype=np.uint8)
    h = b[:64].view(np.uint32)
    nT, nV = int(h[5]), int(min(h[6], 1024))

    def a4(v):
        return (v + 3) & ~3

    o = 64
    out = {"nWords": int(h[0]), "nModules": int(h[1]), "nHits": int(h[2]), "status": int(h[3]),
           "caStatus": int(h[4]), "nTuples": nT, "nVertices": int(h[6]), "nCells": int(h[8])}
    out["tupleOff"] = b[o:o + 4 * (nT + 1)].view(np.uint32).copy()
    o += a4(4 * (nT + 1))
    nb = int(h[7])
    out["tupleBins"] = b[o:o + 2 * nb].view(np.uint16).copy()
    o += a4(2 * 5 * nT)
    out["quality"] = b[o:o + nT].copy()
    o += a4(nT)
    out["state"] = b[o:o + 20 * nT].view(np.float32).reshape(5, nT).T.copy().ravel()
    o += 20 * nT
    out["idv"] = b[o:o + 2 * nT].view(np.int16).copy()
    o += a4(2 * nT)
    out["zv"] = b[o:o + 4 * nV].view(np.float32).copy()
    o += 4 * nV
    out["sortInd"] = b[o:o + 2 * nV].view(np.uint16).copy()
    return out


class Server:
    """Continuous-batching serving loop of the library (pxg_server_*)."""

    def __init__(self, ctx, max_batch=64, n_streams=3, hbm_budget=0, profile=PROFILE_REFERENCE, upto=5,
                 doublet_lanes=0, grid_target=0, transfer=True):
        self.ctx, self.lib = ctx, ctx.lib
        cfg = ServerConfig(max_batch, n_streams, hbm_budget, profile, upto, doublet_lanes, grid_target,
                           1 if transfer else 0)
        self._h = vp()
        _check(self.lib.pxg_server_create(ctx._h, C.byref(cfg), C.byref(self._h)), "pxg_server_create")

    def close(self):
        if getattr(self, "_h", None):
            self.lib.pxg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, words, fed_ids, beamspot=None):
        words = np.ascontiguousarray(words, dtype=np.uint32)
        fed_ids = np.ascontiguousarray(fed_ids, dtype=np.uint8)
        rid = C.c_uint64()
        bs = beamspot if beamspot is not None else self.ctx.beamspot
        _check(self.lib.pxg_server_submit(self._h, words.ctypes.data, fed_ids.ctypes.data, words.size, C.byref(bs),
                                          C.byref(rid)), "pxg_server_submit")
        return rid.value

    def submit_pinned(self, event, beamspot=None):
        """Queues a PinnedEvent without a host copy; it must stay unchanged until polled."""
        rid = C.c_uint64()
        bs = beamspot if beamspot is not None else self.ctx.beamspot
        _check(self.lib.pxg_server_submit_pinned(self._h, event.words.ctypes.data, event.fed_ids.ctypes.data,
                                                 event.words.size, C.byref(bs), C.byref(rid)),
               "pxg_server_submit_pinned")
        return rid.value

    def poll(self, max_results=1024, wait_ms=0):
        out = (ServerResult * max_results)()
        n = _check(self.lib.pxg_server_poll(self._h, out, max_results, wait_ms), "pxg_server_poll")
        return [{f: getattr(out[i], f) for f, _ in ServerResult._fields_} for i in range(n)]

    def products(self, rid, release=True):
        ptr, n = vp(), u32()
        _check(self.lib.pxg_server_products(self._h, rid, C.byref(ptr), C.byref(n)), "pxg_server_products")
        buf = bytes((C.c_uint8 * n.value).from_address(ptr.value))
        if release:
            self.lib.pxg_server_release(self._h, rid)
        return unpack_products(buf)

    def release(self, rid):
        self.lib.pxg_server_release(self._h, rid)

    def info(self):
        w, b, e = u32(), C.c_uint64(), C.c_uint64()
        _check(self.lib.pxg_server_info(self._h, C.byref(w), C.byref(b), C.byref(e)), "pxg_server_info")
        return {"max_batch_words": w.value, "batches": b.value, "events": e.value}


class PinnedEvent:
    """One event's words and FED ids in mapped pinned host memory (pxg_host_alloc), for
    Server.submit_pinned: the device reads them directly while it loads the batch."""

    def __init__(self, words, fed_ids):
        words = np.ascontiguousarray(words, dtype=np.uint32)
        fed_ids = np.ascontiguousarray(fed_ids, dtype=np.uint8)
        self._lib = _gpu()
        self._ptr = vp()
        nb = max(1, 4 * words.size + fed_ids.size + 16)
        _check(self._lib.pxg_host_alloc(nb, C.byref(self._ptr)), "pxg_host_alloc")
        raw = (C.c_uint8 * nb).from_address(self._ptr.value)
        buf = np.frombuffer(raw, dtype=np.uint8)
        self.words = buf[:4 * words.size].view(np.uint32)
        off = (4 * words.size + 15) // 16 * 16
        self.fed_ids = buf[off:off + fed_ids.size]
        self.words[:] = words
        self.fed_ids[:] = fed_ids

    def close(self):
        if getattr(self, "_ptr", None) and self._ptr.value:
            self.words = self.fed_ids = None
            self._lib.pxg_host_free(self._ptr)
            self._ptr = vp()

    def __del__(self):
        self.close()


CAPACITY_FIELDS = ("max_doublets", "max_tuples", "max_tuples_mult", "max_hit_to_tuple", "max_cell_lists",
                   "max_cells_per_hit")


def capacities(profile=PROFILE_REFERENCE):
    """The ntuplet container capacities of a profile (pxg_profile_capacities)."""
    out = (u32 * 6)()
    _check(_gpu().pxg_profile_capacities(profile, C.byref(out)), "pxg_profile_capacities")
    return dict(zip(CAPACITY_FIELDS, list(out)))


def pool_bytes(max_events, max_words, profile=PROFILE_REFERENCE):
    """Device memory of a stream pool (pxg_pool_bytes)."""
    return int(_gpu().pxg_pool_bytes(max_events, max_words, profile))


PRODUCT_KEYS = ("xx", "yy", "adc", "moduleInd", "clus", "pdigi", "rawIdArr", "errors", "clusInModule", "moduleId",
                "clusModuleStart", "xl", "yl", "xerr", "yerr", "xg", "yg", "zg", "rg", "iphi", "charge", "xsize",
                "ysize", "detInd", "hitsLayerStart", "phiOff", "phiBins")


def stored_products(p):
    """The products of one event (a Stream.process / event_products dict with upto=5)
    trimmed to what the event stores: the containers' tails past them are not products."""
    nt, nv = int(p["nTuples"]), int(p["nVertices"])
    nb = int(p["tupleOff"][nt])
    out = {k: p[k] for k in PRODUCT_KEYS}
    out["moduleStart"] = p["moduleStart"][: int(p["nModules"]) + 1]
    out["tupleOff"] = p["tupleOff"][: nt + 1]
    out["tupleBins"] = p["tupleBins"][:nb]
    out["detIndices"] = p["detIndices"][:nb]
    for k in ("quality", "pt", "eta", "chi2", "idv"):
        out[k] = p[k][:nt]
    out["state"] = p["state"][: 5 * nt]
    out["covariance"] = p["covariance"][: 15 * nt]
    for k in ("zv", "wv", "chi2v", "ptv2", "sortInd", "ndof"):
        out[k] = p[k][:nv]
    out["counts"] = np.array([p["nDigis"], p["nModules"], p["nClusters"], nt, nv], np.int64)
    return out


def differing_products(a, b):
    """Names of the stored products in which two events differ (floats compared as bit
    patterns); empty when the two are identical."""
    ua, ub = stored_products(a), stored_products(b)
    bad = []
    for k in ua:
        x, y = np.asarray(ua[k]), np.asarray(ub[k])
        if x.dtype.kind == "f":
            x, y = x.view(np.uint32), y.view(np.uint32)
        if x.shape != y.shape or not np.array_equal(x, y):
            bad.append(k)
    return bad

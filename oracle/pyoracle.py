"""ORACLE -- test infrastructure only.

ctypes binding of the CPU restatement of the serial chain (lib/liboracle.so)
and of the reference's own stage A/B code built from the reference sources
(_ref/libref.so).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may use this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "lib", "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")

_DTYPES = {1: np.uint8, 2: np.uint16, 3: np.int16, 4: np.uint32, 5: np.int32, 6: np.float32, 7: np.float64}
_libs = {}


def _load(path, prefix):
    if path not in _libs:
        if not os.path.exists(path):
            raise OSError(f"{path} is not built (make -C oracle)")
        lib = C.CDLL(path)
        vp = C.c_void_p
        getattr(lib, prefix + "_conditions_create").restype = vp
        getattr(lib, prefix + "_conditions_create").argtypes = [vp] * 10 + [C.c_uint32]
        getattr(lib, prefix + "_conditions_destroy").argtypes = [vp]
        getattr(lib, prefix + "_process").restype = vp
        getattr(lib, prefix + "_process").argtypes = [vp, vp, C.c_size_t, C.c_int]
        getattr(lib, prefix + "_event_destroy").argtypes = [vp]
        getattr(lib, prefix + "_field").restype = vp
        getattr(lib, prefix + "_field").argtypes = [vp, C.c_char_p, C.POINTER(C.c_size_t), C.POINTER(C.c_int)]
        getattr(lib, prefix + "_scalar").restype = C.c_uint32
        getattr(lib, prefix + "_scalar").argtypes = [vp, C.c_char_p]
        getattr(lib, prefix + "_atan2s7").restype = C.c_int16
        getattr(lib, prefix + "_atan2s7").argtypes = [C.c_float, C.c_float]
        if prefix == "orc":
            lib.orc_conditions_set_profile.restype = C.c_int
            lib.orc_conditions_set_profile.argtypes = [vp, C.c_int]
            lib.orc_fit_helix.restype = C.c_int
            lib.orc_fit_helix.argtypes = [C.c_int, vp, vp, C.c_double, vp, vp, vp, vp, vp]
            lib.orc_invert_spd.argtypes = [vp, vp, C.c_int]
            lib.orc_fit_helix_riemann.restype = C.c_int
            lib.orc_fit_helix_riemann.argtypes = [C.c_int, vp, vp, C.c_double, vp, vp, vp, vp, vp]
            lib.orc_circle_fit_riemann.restype = C.c_int
            lib.orc_circle_fit_riemann.argtypes = [C.c_int, vp, vp, C.c_double, vp, vp, vp, vp, vp]
            lib.orc_eigen3.argtypes = [vp, C.c_int, vp, vp]
            lib.orc_rmath.restype = C.c_double
            lib.orc_rmath.argtypes = [C.c_int, C.c_double, C.c_double]
            lib.orc_conditions_set_riemann.restype = C.c_int
            lib.orc_conditions_set_riemann.argtypes = [vp, C.c_int]
            lib.orc_conditions_set_bfield.restype = C.c_int
            lib.orc_conditions_set_bfield.argtypes = [vp, C.c_float]
            lib.orc_conditions_set_vertex_ptmin.restype = C.c_int
            lib.orc_conditions_set_vertex_ptmin.argtypes = [vp, C.c_float]
            lib.orc_vertices.restype = C.c_int
            lib.orc_vertices.argtypes = [C.c_uint32] + [vp] * 10
            lib.orc_vertices_algo.restype = C.c_int
            lib.orc_vertices_algo.argtypes = [C.c_int, C.c_uint32] + [vp] * 10
            lib.orc_conditions_set_vertex_algo.restype = C.c_int
            lib.orc_conditions_set_vertex_algo.argtypes = [vp, C.c_int]
            lib.orc_conditions_set_fishbone.restype = C.c_int
            lib.orc_conditions_set_fishbone.argtypes = [vp, C.c_int, C.c_int]
            lib.orc_circle.argtypes = [C.c_float] * 6 + [C.POINTER(C.c_float)] * 2
            lib.orc_fit_helix_d.restype = C.c_int
            lib.orc_fit_helix_d.argtypes = [C.c_int, vp, vp, C.c_double] + [vp] * 6
            lib.orc_mult_scatt.restype = C.c_double
            lib.orc_mult_scatt.argtypes = [C.c_double] * 4
            lib.orc_cluster_tracks.restype = C.c_int
            lib.orc_cluster_tracks.argtypes = [C.c_int, C.c_uint32, vp, vp, vp, vp]
            lib.orc_set_mutation.argtypes = [C.c_int]
            lib.orc_set_variant.argtypes = [C.c_int]
            lib.orc_circle_debug.restype = C.c_int
            lib.orc_circle_debug.argtypes = [C.c_int, vp, vp, C.c_double, vp, vp]
            lib.orc_fast_fit.restype = C.c_int
            lib.orc_fast_fit.argtypes = [C.c_int, vp, vp]
            lib.orc_translate_karimaki.argtypes = [vp, vp, C.c_double, C.c_double]
            lib.orc_run_many.restype = C.c_uint32
            lib.orc_run_many.argtypes = [vp, vp, vp, C.c_uint32, C.c_int, C.c_int, C.c_uint32]
        else:
            for f in ("ref_divu52", "ref_localX", "ref_localY"):
                getattr(lib, f).restype = C.c_uint16
                getattr(lib, f).argtypes = [C.c_uint16]
        _libs[path] = lib
    return _libs[path]


def ref_available():
    return os.path.exists(REF_LIB)


class _Chain:
    """Serial chain over one detector's conditions; subclasses pick the library."""

    prefix = None
    path = None

    def __init__(self, conditions):
        self.lib = _load(self.path, self.prefix)
        self._keep = {k: np.ascontiguousarray(v) for k, v in conditions.items() if isinstance(v, np.ndarray)}
        p = lambda k: self._keep[k].ctypes.data_as(C.c_void_p)  # noqa: E731
        fed = self._keep["fed_ids"]
        self._h = getattr(self.lib, self.prefix + "_conditions_create")(
            p("cabling"), p("mod_to_unp"), p("gain"), p("gain_data"), p("cpe_common"), p("cpe_det"),
            p("cpe_layer"), p("cpe_average"), p("beamspot"), p("fed_ids"), len(fed))

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            getattr(self.lib, self.prefix + "_conditions_destroy")(h)

    def process(self, raw, upto=3):
        """Runs the chain on one raw record; returns a dict of numpy arrays and scalars."""
        raw = np.ascontiguousarray(raw, dtype=np.uint8)
        e = getattr(self.lib, self.prefix + "_process")(self._h, raw.ctypes.data_as(C.c_void_p), raw.size, upto)
        if not e:
            raise RuntimeError("raw record rejected")
        try:
            out = {}
            for name in FIELDS:
                n, dt = C.c_size_t(), C.c_int()
                ptr = getattr(self.lib, self.prefix + "_field")(e, name.encode(), C.byref(n), C.byref(dt))
                if not dt.value:
                    continue
                dtype = _DTYPES[dt.value]
                count = n.value
                if count == 0:
                    out[name] = np.zeros(0, dtype=dtype)
                else:
                    buf = (C.c_uint8 * (count * np.dtype(dtype).itemsize)).from_address(ptr)
                    out[name] = np.frombuffer(bytes(buf), dtype=dtype).copy()
            for s in SCALARS:
                out[s] = int(getattr(self.lib, self.prefix + "_scalar")(e, s.encode()))
            return out
        finally:
            getattr(self.lib, self.prefix + "_event_destroy")(e)

    def atan2s7(self, y, x):
        return int(getattr(self.lib, self.prefix + "_atan2s7")(y, x))


FIELDS = ["words", "fedIds", "xx", "yy", "adc", "moduleInd", "clus", "pdigi", "rawIdArr", "errors",
          "moduleStart", "clusInModule", "moduleId", "clusModuleStart", "xl", "yl", "xerr", "yerr", "xg", "yg",
          "zg", "rg", "iphi", "charge", "xsize", "ysize", "detInd", "hitsLayerStart", "phiOff", "phiBins",
          "avgCorrected"]
# stage C/D products (only with upto >= 4): cells as parallel arrays plus CSR
# neighbour/track lists, tuples as CSR (tupleOff/tupleBins), fits per tuple,
# vertices per track / per vertex
CA_FIELDS = ["cellInner", "cellOuter", "cellDoubletId", "cellLayerPair", "cellUsed", "cellNeighborOff",
             "cellNeighbors", "cellTrackOff", "cellTracks", "outerHitOff", "outerHitCells", "tupleOff", "tupleBins",
             "quality", "multOff", "multBins", "detIndices", "hitToTupleOff", "hitToTupleBins", "state",
             "covariance", "pt", "eta", "chi2", "idv", "zv", "wv", "chi2v", "ptv2", "ndof", "sortInd", "stats"]
FIELDS = FIELDS + CA_FIELDS
AB_SCALARS = ["nWords", "nModules", "nDigis", "nClusters", "nHits", "fedCounter"]
SCALARS = AB_SCALARS + ["nCells", "nTuples", "caStatus", "nVertices", "nVerticesFound", "vtxTracks", "nNeighborLists",
                        "nTrackLists"]


class Oracle(_Chain):
    prefix = "orc"
    path = ORACLE_LIB

    # vertex clustering algorithms (gpuVertexFinder.h:46-60)
    VERTEX_DENSITY, VERTEX_DBSCAN, VERTEX_ITERATIVE = 0, 1, 2

    def __init__(self, conditions, profile=0, vertex_algo=0, early_fishbone=True, late_fishbone=False,
                 riemann_fit=False, bfield=None, vertex_ptmin=None):
        super().__init__(conditions)
        if bfield is not None and self.lib.orc_conditions_set_bfield(self._h, bfield) != 0:
            raise ValueError(f"bad field {bfield}")
        if vertex_ptmin is not None and self.lib.orc_conditions_set_vertex_ptmin(self._h, vertex_ptmin) != 0:
            raise ValueError(f"bad vertex pt cut {vertex_ptmin}")
        self.lib.orc_conditions_set_fishbone(self._h, int(early_fishbone), int(late_fishbone))
        self.lib.orc_conditions_set_riemann(self._h, int(riemann_fit))
        if self.lib.orc_conditions_set_profile(self._h, profile) != 0:
            raise ValueError(f"unknown capacity profile {profile}")
        if self.lib.orc_conditions_set_vertex_algo(self._h, vertex_algo) != 0:
            raise ValueError(f"unknown vertex algorithm {vertex_algo}")

    def fit_helix(self, hits, ge, B):
        """Broken-line fit of one track: hits (3, n) global xyz in cm (double), ge (6, n) global
        errors (xx, xy, yy, xz, yz, zz; float); returns state[5], cov[15], pt, eta, chi2."""
        hits = np.ascontiguousarray(hits, dtype=np.float64)
        ge = np.ascontiguousarray(ge, dtype=np.float32)
        n = hits.shape[1]
        st = np.zeros(5, np.float32)
        cv = np.zeros(15, np.float32)
        sc = np.zeros(3, np.float32)
        a = lambda x: x.ctypes.data_as(C.c_void_p)  # noqa: E731
        rc = self.lib.orc_fit_helix(n, a(hits), a(ge), B, a(st), a(cv), a(sc[0:]), a(sc[1:]), a(sc[2:]))
        if rc != 0:
            raise ValueError(f"fit_helix: unsupported number of hits {n}")
        return st, cv, float(sc[0]), float(sc[1]), float(sc[2])

    def fit_helix_d(self, hits, ge, B):
        """Broken-line fit in double precision before the float conversion:
        (circle (phi, d, k), its 3x3 covariance, line (cot theta, zip), its 2x2 covariance,
        chi2 of the circle, chi2 of the line)."""
        hits = np.ascontiguousarray(hits, dtype=np.float64)
        ge = np.ascontiguousarray(ge, dtype=np.float32)
        n = hits.shape[1]
        c, cc, ln, lc, x2 = np.zeros(3), np.zeros(9), np.zeros(2), np.zeros(4), np.zeros(2)
        a = lambda v: v.ctypes.data_as(C.c_void_p)  # noqa: E731
        rc = self.lib.orc_fit_helix_d(n, a(hits), a(ge), B, a(c), a(cc), a(ln), a(lc), a(x2[0:]), a(x2[1:]))
        if rc != 0:
            raise ValueError("fit_helix_d: 3 to 5 hits")
        return c, cc.reshape(3, 3), ln, lc.reshape(2, 2), x2[0], x2[1]

    def fast_fit(self, hits):
        hits = np.ascontiguousarray(hits, dtype=np.float64)
        ff = np.zeros(4)
        if self.lib.orc_fast_fit(hits.shape[1], hits.ctypes.data_as(C.c_void_p), ff.ctypes.data_as(C.c_void_p)):
            raise ValueError("fast_fit: 3 or 4 hits")
        return ff

    def circle_debug(self, hits, ge, B, ff=None):
        """The circle fit's intermediate values: dict with d, e (corrected first two hits),
        par_mid, cov_mid ((phi, d, k) and covariance at the chord midpoint), ms (variance added
        to phi at the first hit), S0, S1, R, slope.  ff: a fast fit (X0, Y0, R, tan theta) to
        linearise about instead of the hits' own."""
        hits = np.ascontiguousarray(hits, dtype=np.float64)
        ge = np.ascontiguousarray(ge, dtype=np.float32)
        out = np.zeros(21)
        a = lambda v: v.ctypes.data_as(C.c_void_p)  # noqa: E731
        ffp = None if ff is None else np.ascontiguousarray(ff, dtype=np.float64)
        if self.lib.orc_circle_debug(hits.shape[1], a(hits), a(ge), B, None if ffp is None else a(ffp), a(out)) != 0:
            raise ValueError("circle_debug: 3 or 4 hits")
        return {"d": out[0:2], "e": out[2:4], "par_mid": out[4:7], "cov_mid": out[7:16].reshape(3, 3),
                "ms": out[16], "S0": out[17], "S1": out[18], "R": out[19], "slope": out[20]}

    def translate_karimaki(self, par, cov, x0, y0):
        par = np.array(par, dtype=np.float64)
        cov = np.array(cov, dtype=np.float64).reshape(9).copy()
        a = lambda v: v.ctypes.data_as(C.c_void_p)  # noqa: E731
        self.lib.orc_translate_karimaki(a(par), a(cov), float(x0), float(y0))
        return par, cov.reshape(3, 3)

    def mult_scatt(self, length, B, R, slope):
        return float(self.lib.orc_mult_scatt(length, B, R, slope))

    def cluster_tracks(self, zt, ezt2, algo=0):
        """The vertex finder's track clustering alone: (cluster of each track, 9997 for
        noise; neighbour counts; number of clusters)."""
        zt = np.ascontiguousarray(zt, dtype=np.float32)
        ezt2 = np.ascontiguousarray(ezt2, dtype=np.float32)
        iv = np.zeros(zt.size, np.int32)
        nn = np.zeros(zt.size, np.int32)
        a = lambda v: v.ctypes.data_as(C.c_void_p)  # noqa: E731
        found = self.lib.orc_cluster_tracks(algo, zt.size, a(zt), a(ezt2), a(iv), a(nn))
        return iv, nn, found

    def fit_helix_riemann(self, hits, ge, B):
        """Riemann fit of one track, same arguments and results as fit_helix."""
        hits = np.ascontiguousarray(hits, dtype=np.float64)
        ge = np.ascontiguousarray(ge, dtype=np.float32)
        n = hits.shape[1]
        st = np.zeros(5, np.float32)
        cv = np.zeros(15, np.float32)
        sc = np.zeros(3, np.float32)
        a = lambda x: x.ctypes.data_as(C.c_void_p)  # noqa: E731
        rc = self.lib.orc_fit_helix_riemann(n, a(hits), a(ge), B, a(st), a(cv), a(sc[0:]), a(sc[1:]), a(sc[2:]))
        if rc != 0:
            raise ValueError(f"fit_helix_riemann: unsupported number of hits {n}")
        return st, cv, float(sc[0]), float(sc[1]), float(sc[2])

    def circle_fit_riemann(self, hits, ge, B):
        """The Riemann circle fit alone: fast fit (X0, Y0, R, tan theta), circle (X0, Y0, R),
        its 3x3 covariance, charge and chi2."""
        hits = np.ascontiguousarray(hits, dtype=np.float64)
        ge = np.ascontiguousarray(ge, dtype=np.float32)
        n = hits.shape[1]
        fast, par, cov = np.zeros(4), np.zeros(3), np.zeros((3, 3))
        q = np.zeros(1, np.int32)
        chi2 = np.zeros(1, np.float32)
        a = lambda x: x.ctypes.data_as(C.c_void_p)  # noqa: E731
        if self.lib.orc_circle_fit_riemann(n, a(hits), a(ge), B, a(fast), a(par), a(cov), a(q), a(chi2)) != 0:
            raise ValueError(f"circle_fit_riemann: unsupported number of hits {n}")
        return fast, par, cov, int(q[0]), float(chi2[0])

    def eigen3(self, m, use_float=False):
        """Closed-form eigen decomposition of a symmetric 3x3 (lower triangle read):
        ascending eigenvalues and the eigenvectors as columns."""
        m = np.ascontiguousarray(m, dtype=np.float64)
        ev, vec = np.zeros(3), np.zeros((3, 3))
        a = lambda x: x.ctypes.data_as(C.c_void_p)  # noqa: E731
        self.lib.orc_eigen3(a(m), int(use_float), a(ev), a(vec))
        return ev, vec

    def rmath(self, fn, x, y=0.0):
        """The Riemann fit's own elementary functions: fn 'atan2' (of y, x), 'atan', 'sin',
        'cos' or 'log'."""
        code = {"atan2": 0, "atan": 1, "sin": 2, "cos": 3, "log": 4}[fn]
        return self.lib.orc_rmath(code, float(x), float(y))

    def invert_spd(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        out = np.zeros_like(m)
        self.lib.orc_invert_spd(m.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), m.shape[0])
        return out

    def vertices(self, zt, ezt2, ptt2, algo=0):
        """Vertex finder on given tracks (algo: 0 density, 1 DBSCAN, 2 iterative); returns dict of
        per-track vertex ids and per-vertex arrays."""
        zt, ezt2, ptt2 = (np.ascontiguousarray(a, dtype=np.float32) for a in (zt, ezt2, ptt2))
        nt = zt.size
        iv = np.zeros(nt, np.int16)
        out = {k: np.zeros(1024, np.float32) for k in ("zv", "wv", "chi2", "ptv2")}
        ndof = np.zeros(1024, np.int32)
        srt = np.zeros(1024, np.uint16)
        a = lambda x: x.ctypes.data_as(C.c_void_p)  # noqa: E731
        nv = self.lib.orc_vertices_algo(algo, nt, a(zt), a(ezt2), a(ptt2), a(iv), a(out["zv"]), a(out["wv"]),
                                        a(out["chi2"]), a(out["ptv2"]), a(ndof), a(srt))
        if nv < 0:
            raise ValueError(f"unknown vertex algorithm {algo}")
        res = {k: v[:nv] for k, v in out.items()}
        res.update(iv=iv, ndof=ndof[:nv], sortInd=srt[:nv], nv=nv)
        return res

    def circle(self, x1, y1, x2, y2, x3, y3):
        """Signed curvature c of the circle through 3 points and dca0 = c * (distance of closest approach to (0,0))."""
        c, d = C.c_float(), C.c_float()
        self.lib.orc_circle(x1, y1, x2, y2, x3, y3, C.byref(c), C.byref(d))
        return c.value, d.value

    def run_many(self, raws, upto=3, threads=1, repeat=1):
        """Processes the raw records `repeat` times over `threads` workers; returns events done."""
        blob = np.concatenate(raws).astype(np.uint8)
        offs = np.zeros(len(raws) + 1, dtype=np.uint64)
        offs[1:] = np.cumsum([len(r) for r in raws])
        return int(self.lib.orc_run_many(self._h, blob.ctypes.data_as(C.c_void_p), offs.ctypes.data_as(C.c_void_p),
                                         len(raws), upto, threads, repeat))


class Reference(_Chain):
    prefix = "ref"
    path = REF_LIB

    def divu52(self, n):
        return int(self.lib.ref_divu52(n))

    def localX(self, n):
        return int(self.lib.ref_localX(n))

    def localY(self, n):
        return int(self.lib.ref_localY(n))


class mutation:
    """Context manager planting one of the oracle's mistakes (oracle.h kMut*) for the
    tests that show their checks catch it."""
    NONE, XXI0, CHI2_NDF, VARBETA_ONE_SIDE, NO_FIRST_MS_CIRCLE, NO_SLOPE_FACTOR, NO_FIRST_MS_LINE, LINE_WEIGHT, \
        DENSITY_TIE, DENSITY_NO_CHI2, CU_BORDER, PHI_JACOBIAN = range(12)

    def __init__(self, which):
        self.which = which

    def __enter__(self):
        _load(ORACLE_LIB, "orc").orc_set_mutation(self.which)
        return self

    def __exit__(self, *exc):
        _load(ORACLE_LIB, "orc").orc_set_mutation(0)
        return False


class variant:
    """Context manager switching the oracle away from one of the reference's own expressions
    (oracle.h kVar*), for the tests that show what those expressions do."""
    HALF_CHARGE_REAL, PHI_JACOBIAN_SIGN = 1, 2

    def __init__(self, bits):
        self.bits = bits

    def __enter__(self):
        _load(ORACLE_LIB, "orc").orc_set_variant(self.bits)
        return self

    def __exit__(self, *exc):
        _load(ORACLE_LIB, "orc").orc_set_variant(0)
        return False

"""ctypes binding of the synthetic event source (include/pxgen.h)."""
import ctypes as C
import os

import numpy as np

from ._lib import load_library

PRESETS = {"serial100": 0, "pu50": 1, "pu20_80": 2, "pu_poisson60": 3, "pu200": 4}


class Config(C.Structure):
    _fields_ = [
        ("pu_mode", C.c_int32), ("pu_lo", C.c_int32), ("pu_hi", C.c_int32), ("pu_mean", C.c_double),
        ("tracks_fixed", C.c_int32), ("tracks_mean", C.c_double), ("pt_mode", C.c_int32),
        ("pt_lo", C.c_double), ("pt_hi", C.c_double), ("eta_max", C.c_double), ("vz_sigma", C.c_double),
        ("noise_frac", C.c_double), ("noise_q_lo", C.c_double), ("noise_q_hi", C.c_double),
        ("error_rate", C.c_double), ("seed", C.c_uint64),
    ]


class Truth(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_uint32), ("n_tracks", C.c_uint32), ("n_crossings", C.c_uint32),
        ("n_signal_pixels", C.c_uint32), ("n_noise_pixels", C.c_uint32), ("n_words", C.c_uint32),
        ("n_error_words", C.c_uint32), ("vertex_z", C.c_float * 1024),
    ]


def _lib():
    lib = load_library("libpxgen.so")
    if not getattr(lib, "_pxgen_typed", False):
        vp, u32p = C.c_void_p, C.POINTER(C.c_uint32)
        lib.pxgen_preset.argtypes = [C.c_int, C.POINTER(Config)]
        lib.pxgen_detector_create.restype = vp
        lib.pxgen_detector_create.argtypes = [C.c_uint64]
        lib.pxgen_detector_destroy.argtypes = [vp]
        for name in ("pxgen_cabling", "pxgen_gain", "pxgen_cpe_common", "pxgen_cpe_layer",
                     "pxgen_cpe_average", "pxgen_beamspot"):
            getattr(lib, name).restype = vp
            getattr(lib, name).argtypes = [vp]
        for name in ("pxgen_mod_to_unp", "pxgen_gain_data", "pxgen_cpe_det", "pxgen_fed_ids"):
            getattr(lib, name).restype = vp
            getattr(lib, name).argtypes = [vp, u32p]
        lib.pxgen_write_conditions.argtypes = [vp, C.c_char_p]
        lib.pxgen_event_create.restype = vp
        lib.pxgen_event_create.argtypes = [vp, C.POINTER(Config), C.c_uint32]
        lib.pxgen_event_destroy.argtypes = [vp]
        lib.pxgen_event_raw_size.restype = C.c_size_t
        lib.pxgen_event_raw_size.argtypes = [vp]
        lib.pxgen_event_raw.restype = vp
        lib.pxgen_event_raw.argtypes = [vp]
        lib.pxgen_event_truth.argtypes = [vp, C.POINTER(Truth)]
        lib.pxgen_write_dataset.argtypes = [vp, C.POINTER(Config), C.c_uint32, C.c_char_p]
        lib._pxgen_typed = True
    return lib


def preset(name_or_index, **overrides):
    cfg = Config()
    idx = PRESETS[name_or_index] if isinstance(name_or_index, str) else int(name_or_index)
    if _lib().pxgen_preset(idx, C.byref(cfg)) != 0:
        raise ValueError(f"unknown preset {name_or_index!r}")
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


class Detector:
    """Synthetic detector geometry and conditions (reference binary layouts)."""

    def __init__(self, seed=20261202):
        self._lib = _lib()
        self._h = self._lib.pxgen_detector_create(seed)
        if not self._h:
            raise RuntimeError("pxgen_detector_create failed")

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.pxgen_detector_destroy(h)

    @property
    def handle(self):
        return self._h

    def _blob(self, fn, nbytes):
        p = getattr(self._lib, fn)(self._h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(nbytes,)).copy()

    def _sized(self, fn, itemsize):
        n = C.c_uint32()
        p = getattr(self._lib, fn)(self._h, C.byref(n))
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value * itemsize,)).copy(), n.value

    def conditions(self):
        """All conditions as raw byte arrays, keyed by name."""
        c = {}
        c["cabling"] = self._blob("pxgen_cabling", 1440128)
        c["mod_to_unp"], _ = self._sized("pxgen_mod_to_unp", 1)
        c["gain"] = self._blob("pxgen_gain", 24048)
        c["gain_data"], _ = self._sized("pxgen_gain_data", 1)
        c["cpe_common"] = self._blob("pxgen_cpe_common", 16)
        c["cpe_det"], c["n_det"] = self._sized("pxgen_cpe_det", 112)
        c["cpe_layer"] = self._blob("pxgen_cpe_layer", 4 * 11 + 116)
        c["cpe_average"] = self._blob("pxgen_cpe_average", 4 * (6 * 148 + 2))
        c["beamspot"] = self._blob("pxgen_beamspot", 44)
        fed, n = self._sized("pxgen_fed_ids", 4)
        c["fed_ids"] = fed.view(np.uint32)
        return c

    def write_conditions(self, path):
        os.makedirs(path, exist_ok=True)
        if self._lib.pxgen_write_conditions(self._h, path.encode()) != 0:
            raise OSError(f"cannot write conditions into {path}")

    def write_dataset(self, cfg, n, path):
        os.makedirs(path, exist_ok=True)
        if self._lib.pxgen_write_dataset(self._h, C.byref(cfg), n, path.encode()) != 0:
            raise OSError(f"cannot write dataset into {path}")

    def event(self, cfg, index):
        """Returns (raw record bytes as uint8 array, Truth)."""
        e = self._lib.pxgen_event_create(self._h, C.byref(cfg), index)
        try:
            n = self._lib.pxgen_event_raw_size(e)
            p = self._lib.pxgen_event_raw(e)
            raw = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n,)).copy()
            t = Truth()
            self._lib.pxgen_event_truth(e, C.byref(t))
        finally:
            self._lib.pxgen_event_destroy(e)
        return raw, t

import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "pixeltrack-standalone_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def detector():
    from pxtrack import gen
    return gen.Detector()


@pytest.fixture(scope="session")
def conditions(detector):
    return detector.conditions()


@pytest.fixture(scope="session")
def oracle(conditions):
    import pyoracle
    return pyoracle.Oracle(conditions)


@pytest.fixture(scope="session")
def reference(conditions):
    import pyoracle
    if not pyoracle.ref_available():
        pytest.skip("oracle/_ref/libref.so not built (needs the reference sources)")
    return pyoracle.Reference(conditions)


@pytest.fixture(scope="session")
def gpu_ctx(conditions):
    from pxtrack import gpu
    return gpu.Context(conditions)


@pytest.fixture(scope="session")
def gpu_stream(gpu_ctx):
    return gpu_ctx.stream(max_events=1, max_words=300000)


def event(detector, preset, index, **overrides):
    from pxtrack import gen
    raw, truth = detector.event(gen.preset(preset, **overrides), index)
    return raw, truth


def assert_same(a, b, keys, ctx=""):
    """Bit-exact comparison of named arrays (floats compared as bit patterns)."""
    for k in keys:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape, f"{ctx} {k}: shape {x.shape} != {y.shape}"
        if x.dtype.kind == "f":
            x, y = x.view(np.uint32 if x.itemsize == 4 else np.uint64), y.view(np.uint32 if y.itemsize == 4 else np.uint64)
        if not np.array_equal(x, y):
            bad = np.nonzero(x != y)[0]
            raise AssertionError(f"{ctx} {k}: {bad.size} of {x.size} differ, first at {bad[:5]}: "
                                 f"{np.asarray(a[k])[bad[:5]]} vs {np.asarray(b[k])[bad[:5]]}")


ERROR_DTYPE = np.dtype({"names": ["rawId", "word", "errorType", "fedId"], "formats": [np.uint32, np.uint32, np.uint8, np.uint8],
                        "offsets": [0, 4, 8, 9], "itemsize": 12})


def error_records(raw_bytes):
    """PixelErrorCompact records (12 bytes, 2 of them padding) -> structured array of the 4 fields."""
    rec = np.frombuffer(np.ascontiguousarray(raw_bytes).tobytes(), dtype=ERROR_DTYPE)
    out = np.zeros(rec.size, dtype=[("rawId", np.uint32), ("word", np.uint32), ("errorType", np.uint8), ("fedId", np.uint8)])
    for f in out.dtype.names:
        out[f] = rec[f]
    return out


def assert_same_errors(a, b, ctx=""):
    ea, eb = error_records(a), error_records(b)
    assert ea.size == eb.size, f"{ctx}: {ea.size} != {eb.size} error records"
    assert np.array_equal(ea, eb), f"{ctx}: error records differ"

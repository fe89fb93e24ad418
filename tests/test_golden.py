"""Golden fixtures of stages A/B made by the reference's own code (tests/golden/make_golden.py):
the oracle (CPU) and the HIP path (GPU) must reproduce them bit for bit."""
import glob
import os

import numpy as np
import pytest

from conftest import REPO, assert_same, assert_same_errors

GOLDEN = sorted(glob.glob(os.path.join(REPO, "tests", "golden", "stageAB_*.npz")))
FIELDS = ["xx", "yy", "adc", "moduleInd", "clus", "pdigi", "rawIdArr", "moduleStart", "clusInModule",
          "moduleId", "clusModuleStart", "xl", "yl", "xerr", "yerr", "xg", "yg", "zg", "rg", "iphi", "charge",
          "xsize", "ysize", "detInd", "hitsLayerStart", "phiOff", "phiBins"]


def _load(path, conditions):
    import sys
    sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
    from make_golden import conditions_digest
    f = np.load(path, allow_pickle=False)
    if str(f["conditions_sha256"]) != conditions_digest(conditions):
        pytest.fail(f"{os.path.basename(path)} was made with other conditions: regenerate it")
    expected = {k: f[f"out_{k}"] for k in FIELDS + ["errors"]}
    expected.update({k[len("scalar_"):]: int(f[k]) for k in f.files if k.startswith("scalar_")})
    return f["raw"], expected


def test_fixtures_present():
    assert len(GOLDEN) >= 3


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_oracle_matches_golden(conditions, oracle, path):
    raw, want = _load(path, conditions)
    got = oracle.process(raw, upto=3)
    for k in ("nModules", "nDigis", "nClusters", "nHits"):
        assert got[k] == want[k], k
    assert_same(want, got, FIELDS, os.path.basename(path))
    assert_same_errors(want["errors"], got["errors"], os.path.basename(path))


@pytest.mark.gpu
@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_gpu_matches_golden(conditions, gpu_ctx, gpu_stream, path):
    from pxtrack import gpu
    raw, want = _load(path, conditions)
    words, fed = gpu.unpack_raw(raw, gpu_ctx.fed_ids)
    got = gpu_stream.process(words, fed, upto=3)
    n = want["nDigis"]
    got_digis = {k: got[k][:n] for k in ("xx", "yy", "adc", "moduleInd", "clus", "pdigi", "rawIdArr")}
    got.update(got_digis)
    # module starts are defined up to the number of module runs
    m = want["nModules"] + 1
    assert np.array_equal(want["moduleStart"][:m], got["moduleStart"][:m]), "moduleStart"
    assert_same(want, got, [f for f in FIELDS if f != "moduleStart"], os.path.basename(path))
    assert_same_errors(want["errors"], got["errors"], os.path.basename(path))

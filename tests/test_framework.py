"""The job driver with the reference's module interface (pixeltrack-amd): command
line, raw.bin and conditions files, the throughput line, and per-event counts
equal to the oracle's when the four producers run on the GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import REPO

EXE = os.path.join(REPO, "pixeltrack-standalone_amd", "lib", "pixeltrack-amd")
LINE = re.compile(r"Processed (\d+) events in ([0-9.e+-]+) seconds, throughput ([0-9.e+-]+) events/s")


@pytest.fixture(scope="module")
def dataset(detector, tmp_path_factory):
    from pxtrack import gen
    d = tmp_path_factory.mktemp("data")
    detector.write_conditions(str(d))
    detector.write_dataset(gen.preset("pu50", pu_lo=8, pu_hi=8), 12, str(d))
    return str(d)


def _run(*args, timeout=300):
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)


def test_help_and_bad_option():
    assert _run("--help").returncode == 0
    r = _run("--bogus")
    assert r.returncode != 0 and "Invalid parameter" in r.stdout


def test_missing_data_is_reported(tmp_path):
    r = _run("--data", str(tmp_path / "nope"))
    assert r.returncode != 0 and "cannot open" in r.stdout


def test_empty_job_reads_raw_and_reports(dataset):
    r = _run("--data", dataset, "--empty", "--numberOfThreads", "3", "--maxEvents", "30")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Processing 30 events, with 3 concurrent events and 3 threads." in r.stdout
    m = LINE.search(r.stdout)
    assert m and int(m.group(1)) == 30


@pytest.mark.gpu
def test_gpu_job_counts_match_oracle(dataset, detector, oracle, tmp_path):
    from pxtrack import gen
    counts = tmp_path / "counts.txt"
    r = _run("--data", dataset, "--numberOfThreads", "2", "--warmupEvents", "2", "--validation", str(counts))
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    m = LINE.search(r.stdout)
    assert m and int(m.group(1)) == 12
    rows = np.loadtxt(counts, dtype=np.int64, ndmin=2)
    rows = rows[rows[:, 0] >= 2]  # timed events 2..13 cycle through the 12 records
    assert len(rows) == 12
    for ev_id, nmod, ndig, nclus, ntup, nloose, nvtx in rows:
        raw, _ = detector.event(gen.preset("pu50", pu_lo=8, pu_hi=8), int(ev_id) % 12)
        o = oracle.process(raw, upto=5)
        nt = o["nTuples"]
        assert (nmod, ndig, nclus, ntup) == (o["nModules"], o["nDigis"], o["nClusters"], nt), ev_id
        assert nloose == int((o["quality"][:nt] == 2).sum()), ev_id
        assert nvtx == o["nVertices"], ev_id

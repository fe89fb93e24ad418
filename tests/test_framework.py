"""The job driver with the reference's module interface (pixeltrack-amd): command
line, raw.bin and conditions files, the throughput line, per-event counts
equal to the oracle's when the four producers run on the GPU, the
CountValidator against expected counts in the reference's digicluster.bin /
tracks.bin / vertices.bin layout, and the HistoValidator's histograms against
the same histograms of the oracle's products."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import REPO

EXE = os.path.join(REPO, "pixeltrack-standalone_amd", "lib", "pixeltrack-amd")
LINE = re.compile(r"Processed (\d+) events in ([0-9.e+-]+) seconds, throughput ([0-9.e+-]+) events/s")


CFG = dict(pu_lo=8, pu_hi=8)
N_EVENTS = 12


@pytest.fixture(scope="module")
def dataset(detector, tmp_path_factory):
    from pxtrack import gen
    d = tmp_path_factory.mktemp("data")
    detector.write_conditions(str(d))
    detector.write_dataset(gen.preset("pu50", **CFG), N_EVENTS, str(d))
    return str(d)


@pytest.fixture(scope="module")
def oracle_products(detector, oracle):
    from pxtrack import gen
    return [oracle.process(detector.event(gen.preset("pu50", **CFG), i)[0], upto=5) for i in range(N_EVENTS)]


def write_expected_counts(d, products, tweak=None):
    """The validation files in the reference's layout (bin/Source.cc:58-68): per event
    u32 modules, digis, clusters in digicluster.bin; u32 tracks in tracks.bin; u32
    vertices in vertices.bin.  `tweak(i, row)` may change an event's row."""
    dc, tr, vx = [], [], []
    for i, o in enumerate(products):
        row = [o["nModules"], o["nDigis"], o["nClusters"], o["nTuples"], o["nVertices"]]
        if tweak:
            tweak(i, row)
        dc += row[:3]
        tr.append(row[3])
        vx.append(row[4])
    np.array(dc, dtype=np.uint32).tofile(os.path.join(d, "digicluster.bin"))
    np.array(tr, dtype=np.uint32).tofile(os.path.join(d, "tracks.bin"))
    np.array(vx, dtype=np.uint32).tofile(os.path.join(d, "vertices.bin"))


def _run(*args, timeout=300):
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)


def test_help_and_bad_option():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--runForMinutes", "--transfer", "--validation", "--histogram"):
        assert flag in r.stdout
    r = _run("--bogus")
    assert r.returncode != 0 and "Invalid parameter" in r.stdout
    r = _run("--maxEvents", "3", "--runForMinutes", "1")
    assert r.returncode != 0 and "please give only one of them" in r.stdout


def test_validation_needs_the_count_files(dataset, tmp_path):
    import shutil
    d = tmp_path / "nocounts"
    shutil.copytree(dataset, d)
    for f in ("digicluster.bin", "tracks.bin", "vertices.bin"):
        if (d / f).exists():
            (d / f).unlink()
    r = _run("--data", str(d), "--validation", "--maxEvents", "2")
    assert r.returncode != 0 and "digicluster.bin" in r.stdout


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
    r = _run("--data", dataset, "--numberOfThreads", "2", "--warmupEvents", "2", "--counts", str(counts))
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    m = LINE.search(r.stdout)
    assert m and int(m.group(1)) == 12
    rows = np.loadtxt(counts, dtype=np.int64, ndmin=2)
    rows = rows[rows[:, 0] >= 2]  # timed events 2..13 cycle through the 12 records
    assert len(rows) == 12
    for ev_id, nmod, ndig, nclus, ntup, nloose, nvtx in rows:
        raw, _ = detector.event(gen.preset("pu50", **CFG), int(ev_id) % N_EVENTS)
        o = oracle.process(raw, upto=5)
        nt = o["nTuples"]
        assert (nmod, ndig, nclus, ntup) == (o["nModules"], o["nDigis"], o["nClusters"], nt), ev_id
        assert nloose == int((o["quality"][:nt] == 2).sum()), ev_id
        assert nvtx == o["nVertices"], ev_id


@pytest.mark.gpu
def test_gpu_job_count_validator(dataset, oracle_products, tmp_path):
    import shutil
    d = tmp_path / "data"
    shutil.copytree(dataset, d)
    write_expected_counts(str(d), oracle_products)
    r = _run("--data", str(d), "--numberOfThreads", "3", "--validation", "--maxEvents", "24")
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "CountValidator: all 24 events passed validation" in r.stdout

    # a wrong expected cluster count in one record fails the events made from it;
    # a vertex count off by one is within the reference's tolerance
    def tweak(i, row):
        if i == 5:
            row[2] += 1
        if i == 7:
            row[4] += 1
    write_expected_counts(str(d), oracle_products, tweak)
    r = _run("--data", str(d), "--numberOfThreads", "2", "--validation", "--maxEvents", "24")
    print(r.stdout, flush=True)
    assert r.returncode != 0
    assert "CountValidator: 2 events failed validation" in r.stdout
    assert "N(clusters) is" in r.stdout
    assert "Average absolute vertex difference" not in r.stdout  # only printed when all pass


@pytest.mark.gpu
def test_gpu_job_histograms(detector, dataset, oracle_products, gpu_ctx, tmp_path):
    """HistoValidator histograms of three runs of the same events, compared bin by bin
    with the histogram-ratio tool (pxtrack.histos, plot-histograms.py:97-196): the oracle
    ("serial", filled in Python), the job driver's plain per-event path ("mi355x", written
    by its HistoValidator) and the batched path ("batched", one launch over all events).
    Exact products give equal histograms; the fitted quantities, equal only within the
    fits' tolerance, may move an entry across a bin edge."""
    from pxtrack import gen, gpu, histos
    r = subprocess.run([EXE, "--data", dataset, "--numberOfThreads", "2", "--histogram"], capture_output=True,
                       text=True, timeout=300, cwd=str(tmp_path))
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    got = histos.read(str(tmp_path / "histograms_mi355x.txt"))
    assert set(got) == set(histos.BINNING) and len(got) == 33
    serial = histos.Histograms()
    for o in oracle_products:
        serial.fill(o)
    serial.write(str(tmp_path / "histograms_serial.txt"))
    raws = [detector.event(gen.preset("pu50", **CFG), i)[0] for i in range(N_EVENTS)]
    inputs = [gpu.unpack_raw(raw, gpu_ctx.fed_ids) for raw in raws]
    s = gpu_ctx.stream(max_events=N_EVENTS, max_words=int(sum(len(w) for w, _ in inputs)))
    s.load(inputs)
    s.run(0, N_EVENTS, upto=5)
    s.fetch_counts(0, N_EVENTS)
    batched = histos.Histograms()
    for e in range(N_EVENTS):
        batched.fill(s.event_products(e, upto=5))
    s.close()
    batched.write(str(tmp_path / "histograms_batched.txt"))
    for name, (lo, hi, counts) in got.items():
        nb, blo, bhi = histos.BINNING[name]
        assert len(counts) == nb + 2 and lo == pytest.approx(blo, rel=1e-5) and hi == pytest.approx(bhi, rel=1e-5)
        exp = serial.counts[name]
        if name in histos.APPROX:
            assert counts.sum() == exp.sum(), name
            assert np.abs(counts - exp).sum() <= 4, (name, np.nonzero(counts - exp))
        else:
            assert np.array_equal(counts, exp), (name, np.nonzero(counts - exp))
    files = [str(tmp_path / f"histograms_{k}.txt") for k in ("serial", "mi355x", "batched")]
    plots = tmp_path / "plots"
    rc = histos.main(["--den", "serial", *files, "--approx-tolerance", "2e-3", "--plot", str(plots)])
    assert rc == 0
    assert len(list(plots.glob("*.png"))) == 33
    # batched and plain runs give identical products, hence identical histograms
    assert histos.main(["--den", "mi355x", files[1], files[2]]) == 0

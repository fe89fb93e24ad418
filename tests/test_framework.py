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


def _histo(values, nbins, lo, hi):
    """Fixed-range histogram with under/overflow bins in float32 arithmetic, the
    rule of the reference's SimpleAtomicHisto::fill."""
    lo32, hi32 = np.float32(lo), np.float32(hi)
    out = np.zeros(nbins + 2, dtype=np.int64)
    for v in np.asarray(values, dtype=np.float32):
        if v < lo32:
            out[0] += 1
        elif v >= hi32:
            out[-1] += 1
        else:
            i = int(np.float32(np.float32(v - lo32) / np.float32(hi32 - lo32)) * np.float32(nbins))
            out[min(i, nbins - 1) + 1] += 1
    return out


@pytest.mark.gpu
def test_gpu_job_histograms(dataset, oracle_products, tmp_path):
    r = subprocess.run([EXE, "--data", dataset, "--numberOfThreads", "2", "--histogram"], capture_output=True,
                       text=True, timeout=300, cwd=str(tmp_path))
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    got = {}
    for line in open(tmp_path / "histograms_mi355x.txt"):
        name, n, lo, hi, *counts = line.split()
        assert int(n) == len(counts)
        got[name] = (float(lo), float(hi), np.array(counts, dtype=np.int64))
    assert len(got) == 33
    exp = {}
    def add(name, values, nbins, lo, hi):
        h = _histo(values, nbins, lo, hi)
        exp[name] = exp.get(name, 0) + h
    for o in oracle_products:
        nt = o["nTuples"]
        loose = (o["quality"][:nt] >= 2)
        sizes = np.diff(o["tupleOff"][:nt + 1].astype(np.int64))
        st = o["state"][:5 * nt].reshape(-1, 5)
        nv = o["nVertices"]
        add("digi_n", [o["nDigis"]], 100, 0, 1e5)
        add("digi_adc", o["adc"][:o["nDigis"]], 250, 0, 5e4)
        add("module_n", [o["nModules"]], 100, 1500, 2000)
        add("cluster_n", [o["nClusters"]], 200, 5000, 25000)
        add("cluster_per_module_n", o["clusInModule"][:o["nModules"]], 110, 0, 110)
        nh = o["nHits"]
        add("hit_n", [nh], 200, 5000, 25000)
        for key, name, nb, lo, hi in (("xl", "hit_lx", 200, -1, 1), ("yl", "hit_ly", 800, -4, 4),
                                      ("xerr", "hit_lex", 100, 0, 5e-5), ("yerr", "hit_ley", 100, 0, 1e-4),
                                      ("xg", "hit_gx", 200, -20, 20), ("yg", "hit_gy", 200, -20, 20),
                                      ("zg", "hit_gz", 600, -60, 60), ("rg", "hit_gr", 200, 0, 20),
                                      ("charge", "hit_charge", 400, 0, 4e6), ("xsize", "hit_sizex", 800, 0, 800),
                                      ("ysize", "hit_sizey", 800, 0, 800)):
            add(name, o[key][:nh], nb, lo, hi)
        add("track_n", [int(loose.sum())], 150, 0, 15000)
        add("track_nhits", sizes[loose], 3, 3, 6)
        add("track_quality", o["quality"][:nt][loose], 6, 0, 6)
        add("track_chi2", o["chi2"][:nt][loose], 100, 0, 40)
        add("track_pt", o["pt"][:nt][loose], 400, 0, 400)
        add("track_eta", o["eta"][:nt][loose], 100, -3, 3)
        add("track_phi", st[loose, 0], 100, -3.15, 3.15)
        add("track_tip", st[loose, 1], 100, -1, 1)
        add("track_tip_zoom", st[loose, 1], 100, -0.05, 0.05)
        add("track_zip", st[loose, 4], 100, -15, 15)
        add("track_zip_zoom", st[loose, 4], 100, -0.1, 0.1)
        add("vertex_n", [nv], 60, 0, 60)
        add("vertex_z", o["zv"][:nv], 100, -15, 15)
        add("vertex_chi2", o["chi2v"][:nv], 100, 0, 40)
        add("vertex_ndof", o["ndof"][:nv], 170, 0, 170)
        add("vertex_pt2", o["ptv2"][:nv], 100, 0, 4000)
    # floats the device computes within 1e-5 of the oracle (fits, vertex fits) may sit on
    # the other side of a bin edge: same totals, at most a few entries moved
    approx = {"track_chi2", "track_pt", "track_eta", "track_phi", "track_tip", "track_tip_zoom", "track_zip",
              "track_zip_zoom", "vertex_z", "vertex_chi2", "vertex_pt2"}
    for name, h in exp.items():
        g = got[name][2]
        assert len(g) == len(h), name
        if name in approx:
            assert g.sum() == h.sum(), name
            assert np.abs(g - h).sum() <= 4, (name, np.nonzero(g - h))
        else:
            assert np.array_equal(g, h), (name, np.nonzero(g - h))
    assert set(exp) == set(got)

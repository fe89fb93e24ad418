"""The histogram-ratio tool (pxtrack.histos): the fill rule of SimpleAtomicHisto
(plugin-Validation/SimpleAtomicHisto.h:20-41) against a scalar float32 loop, the
text format round trip, and the comparison of plot-histograms.py:97-196 on the
oracle's products: equal runs compare equal, a moved entry and a changed
fitted value are reported, and the command line exits 1 on a difference."""
import numpy as np
import pytest

from conftest import event


def _scalar_fill(values, nbins, lo, hi):
    lo32, hi32 = np.float32(lo), np.float32(hi)
    out = [0] * (nbins + 2)
    for v in np.asarray(values, dtype=np.float32):
        if v < lo32:
            i = 0
        elif v >= hi32:
            i = nbins + 1
        else:
            i = int(np.float32(np.float32(v - lo32) / np.float32(hi32 - lo32)) * np.float32(nbins))
            i = min(i, nbins - 1) + 1
        out[i] += 1
    return np.array(out)


def test_bin_counts_follow_the_float32_fill_rule():
    from pxtrack import histos
    rng = np.random.default_rng(5)
    for nb, lo, hi in ((100, -3.15, 3.15), (3, 3, 6), (100, 0, 5e-5), (400, 0, 4e6)):
        span = hi - lo
        v = np.concatenate([rng.uniform(lo - 0.1 * span, hi + 0.1 * span, 5000),
                            [lo, hi, np.nextafter(np.float32(hi), np.float32(lo)), np.float32(lo) - 1e-7]])
        # bin edges themselves, where float32 rounding decides
        v = np.concatenate([v, lo + np.arange(nb + 1) * span / nb])
        assert np.array_equal(histos.bin_counts(v, nb, lo, hi), _scalar_fill(v, nb, lo, hi)), (nb, lo, hi)


@pytest.fixture(scope="module")
def products(detector, oracle):
    return [oracle.process(event(detector, "pu50", i, pu_lo=8, pu_hi=8)[0], upto=5) for i in range(3)]


def test_round_trip_and_comparison(products, tmp_path):
    from pxtrack import histos
    a = histos.Histograms()
    for p in products:
        a.fill(p)
    assert sum(a.counts["digi_n"]) == 3 and a.counts["track_n"].sum() == 3
    pa = tmp_path / "histograms_serial.txt"
    a.write(str(pa))
    back = histos.read(str(pa))
    assert set(back) == set(histos.BINNING)
    for name, (lo, hi, c) in back.items():
        assert np.array_equal(c, a.counts[name]), name
        assert lo == pytest.approx(histos.BINNING[name][1], rel=1e-6, abs=1e-12)
    # the same products again: every ratio is 1
    b = histos.Histograms()
    for p in products:
        b.fill(p)
    pb = tmp_path / "histograms_again.txt"
    b.write(str(pb))
    assert histos.main(["--den", "serial", str(pa), str(pb)]) == 0
    # one hit's charge in another bin: an exact histogram differs
    q = dict(products[0])
    q["charge"] = q["charge"].copy()
    q["charge"][0] += 50000
    c = histos.Histograms()
    for p in [q] + products[1:]:
        c.fill(p)
    pc = tmp_path / "histograms_moved.txt"
    c.write(str(pc))
    r = histos.compare(c.counts["hit_charge"], a.counts["hit_charge"], "hit_charge")
    assert not r["ok"] and r["differing_bins"] == 2
    assert histos.main(["--den", "serial", str(pa), str(pc)]) == 1
    # a fitted value moved by one entry is within the fitted quantities' tolerance only
    # when few entries move
    f = dict(products[0])
    f["pt"] = f["pt"] * np.float32(1.5)
    d = histos.Histograms()
    for p in [f] + products[1:]:
        d.fill(p)
    r = histos.compare(d.counts["track_pt"], a.counts["track_pt"], "track_pt", approx_tolerance=1e-3)
    assert not r["ok"] and r["moved_fraction"] > 1e-3
    one = histos.compare(a.counts["track_pt"], a.counts["track_pt"], "track_pt")
    assert one["ok"] and one["max_dev"] == 0


def test_plots(products, tmp_path):
    from pxtrack import histos
    a = histos.Histograms()
    for p in products:
        a.fill(p)
    pa, pb = tmp_path / "histograms_x.txt", tmp_path / "histograms_y.txt"
    a.write(str(pa))
    a.write(str(pb))
    assert histos.main(["--den", "x", str(pa), str(pb), "--plot", str(tmp_path / "plots")]) == 0
    assert len(list((tmp_path / "plots").glob("*.png"))) == 33

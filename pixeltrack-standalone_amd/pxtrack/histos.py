"""Product histograms and their comparison across runs.

The 33 histograms of the job driver's HistoValidator (the reference's
plugin-Validation/HistoValidator.cc:29-151, binning of :29-66) filled from one
event's products in the dict layout of `gpu.Stream.process` /
`Stream.event_products`, the text format the job driver writes
(`histograms_<label>.txt`: one line per histogram, "name nbins+2 min max
underflow counts... overflow"), and the per-bin ratio comparison of
plot-histograms.py:97-196 (bins with an empty denominator are skipped).

Command line (compares the files' histograms with the denominator's):

    python -m pxtrack.histos --den serial histograms_serial.txt histograms_plain.txt histograms_batched.txt
        [--tolerance 0] [--approx-tolerance 1e-3] [--plot DIR]

prints one line per histogram and numerator (bins compared, the largest
|ratio - 1|, the bins that differ, the under/overflow ratios) and exits 1 when a
histogram differs by more than the tolerance: --tolerance for the histograms of
exact products, --approx-tolerance (a fraction of the entries moved to other
bins) for the fitted track and vertex quantities, whose values agree only within
the fits' tolerance and may cross a bin edge.  --plot writes the distributions
and ratios as PNG files when matplotlib is importable.
"""
import argparse
import os
import re
import sys

import numpy as np

# name -> (bins, min, max): HistoValidator.cc:29-66
BINNING = {
    "digi_n": (100, 0, 1e5), "digi_adc": (250, 0, 5e4), "module_n": (100, 1500, 2000),
    "cluster_n": (200, 5000, 25000), "cluster_per_module_n": (110, 0, 110), "hit_n": (200, 5000, 25000),
    "hit_lx": (200, -1, 1), "hit_ly": (800, -4, 4), "hit_lex": (100, 0, 5e-5), "hit_ley": (100, 0, 1e-4),
    "hit_gx": (200, -20, 20), "hit_gy": (200, -20, 20), "hit_gz": (600, -60, 60), "hit_gr": (200, 0, 20),
    "hit_charge": (400, 0, 4e6), "hit_sizex": (800, 0, 800), "hit_sizey": (800, 0, 800),
    "track_n": (150, 0, 15000), "track_nhits": (3, 3, 6), "track_chi2": (100, 0, 40), "track_pt": (400, 0, 400),
    "track_eta": (100, -3, 3), "track_phi": (100, -3.15, 3.15), "track_tip": (100, -1, 1),
    "track_tip_zoom": (100, -0.05, 0.05), "track_zip": (100, -15, 15), "track_zip_zoom": (100, -0.1, 0.1),
    "track_quality": (6, 0, 6), "vertex_n": (60, 0, 60), "vertex_z": (100, -15, 15), "vertex_chi2": (100, 0, 40),
    "vertex_ndof": (170, 0, 170), "vertex_pt2": (100, 0, 4000),
}
# fitted quantities: equal between implementations only within the fits' tolerance
APPROX = {"track_chi2", "track_pt", "track_eta", "track_phi", "track_tip", "track_tip_zoom", "track_zip",
          "track_zip_zoom", "vertex_z", "vertex_chi2", "vertex_pt2"}


def bin_counts(values, nbins, lo, hi):
    """Counts of a fixed-range histogram with under/overflow bins, in float32 arithmetic
    (SimpleAtomicHisto::fill, plugin-Validation/SimpleAtomicHisto.h:20-41): below min
    to bin 0, at or above max to the last bin, a value that rounds onto the upper
    edge into the last regular bin."""
    v = np.asarray(values, dtype=np.float32).ravel()
    lo32, hi32 = np.float32(lo), np.float32(hi)
    out = np.zeros(nbins + 2, dtype=np.int64)
    under, over = v < lo32, v >= hi32
    out[0] = int(under.sum())
    out[-1] = int(over.sum())
    mid = v[~under & ~over]
    i = (((mid - lo32) / (hi32 - lo32)) * np.float32(nbins)).astype(np.int64)
    i = np.minimum(i, nbins - 1)
    out[1:-1] += np.bincount(i, minlength=nbins)[:nbins]
    return out


class Histograms:
    """The 33 histograms, filled event by event."""

    def __init__(self):
        self.counts = {k: np.zeros(b[0] + 2, dtype=np.int64) for k, b in BINNING.items()}

    def add(self, name, values):
        self.counts[name] += bin_counts(values, *BINNING[name])

    def fill(self, p):
        """One event's products (digis, clusters, hits, tracks, vertices) as in HistoValidator::produce
        (HistoValidator.cc:83-151): tracks are the stored tuples with hits and quality loose."""
        nd, nm = int(p["nDigis"]), int(p["nModules"])
        self.add("digi_n", [nd])
        self.add("digi_adc", p["adc"][:nd])
        self.add("module_n", [nm])
        self.add("cluster_n", [int(p["nClusters"])])
        self.add("cluster_per_module_n", p["clusInModule"][:nm])
        nh = int(p["nHits"])
        self.add("hit_n", [nh])
        for key, name in (("xl", "hit_lx"), ("yl", "hit_ly"), ("xerr", "hit_lex"), ("yerr", "hit_ley"),
                          ("xg", "hit_gx"), ("yg", "hit_gy"), ("zg", "hit_gz"), ("rg", "hit_gr"),
                          ("charge", "hit_charge"), ("xsize", "hit_sizex"), ("ysize", "hit_sizey")):
            self.add(name, p[key][:nh])
        nt = int(p["nTuples"])
        sizes = np.diff(np.asarray(p["tupleOff"][:nt + 1], dtype=np.int64))
        loose = (sizes > 0) & (np.asarray(p["quality"][:nt]) >= 2)
        state = np.asarray(p["state"][:5 * nt]).reshape(-1, 5)  # per tuple: phi, tip, q/pt, cot(theta), zip
        self.add("track_n", [int(loose.sum())])
        self.add("track_nhits", sizes[loose])
        self.add("track_chi2", p["chi2"][:nt][loose])
        self.add("track_pt", p["pt"][:nt][loose])
        self.add("track_eta", p["eta"][:nt][loose])
        self.add("track_phi", state[loose, 0])
        self.add("track_tip", state[loose, 1])
        self.add("track_tip_zoom", state[loose, 1])
        self.add("track_zip", state[loose, 4])
        self.add("track_zip_zoom", state[loose, 4])
        self.add("track_quality", np.asarray(p["quality"][:nt])[loose])
        nv = int(p["nVertices"])
        self.add("vertex_n", [nv])
        self.add("vertex_z", p["zv"][:nv])
        self.add("vertex_chi2", p["chi2v"][:nv])
        self.add("vertex_ndof", p["ndof"][:nv])
        self.add("vertex_pt2", p["ptv2"][:nv])
        return self

    def write(self, path):
        with open(path, "w") as f:
            for name in sorted(self.counts):
                nb, lo, hi = BINNING[name]
                # the bounds as a C++ stream prints a float (6 significant digits)
                f.write(f"{name} {nb + 2} {float(np.float32(lo)):g} {float(np.float32(hi)):g} "
                        + " ".join(str(int(c)) for c in self.counts[name]) + "\n")


def read(path):
    """{name: (min, max, counts with under/overflow)} of a histograms_<label>.txt file."""
    out = {}
    with open(path) as f:
        for line in f:
            fields = line.split()
            if not fields:
                continue
            name, n, lo, hi = fields[0], int(fields[1]), float(fields[2]), float(fields[3])
            counts = np.array(fields[4:], dtype=np.int64)
            if len(counts) != n:
                raise ValueError(f"{path}: {name} has {len(counts)} counts, header says {n}")
            out[name] = (lo, hi, counts)
    return out


def label_of(path):
    m = re.search(r"histograms_(?P<label>.*)\.txt$", os.path.basename(path))
    return m.group("label") if m else os.path.basename(path)


def ratio(num, den):
    """Per-bin ratio over the regular bins with a nonzero denominator (plot-histograms.py:108-115):
    (bin indices, ratios)."""
    n, d = num[1:-1].astype(np.float64), den[1:-1].astype(np.float64)
    keep = d != 0
    return np.nonzero(keep)[0], n[keep] / d[keep]


def compare(num, den, name, tolerance=0.0, approx_tolerance=1e-3):
    """Summary of one histogram of a numerator run against the denominator's."""
    idx, r = ratio(num, den)
    diff = np.abs(num - den)
    moved = diff.sum() / 2 / max(1, den.sum())
    maxdev = float(np.abs(r - 1).max()) if r.size else 0.0
    empty_num = int(((den[1:-1] == 0) & (num[1:-1] != 0)).sum())
    if name in APPROX:
        ok = num.sum() == den.sum() and moved <= approx_tolerance
    else:
        ok = maxdev <= tolerance and empty_num == 0 and (tolerance > 0 or np.array_equal(num, den))
    under = num[0] / den[0] if den[0] else (0.0 if num[0] == 0 else float("inf"))
    over = num[-1] / den[-1] if den[-1] else (0.0 if num[-1] == 0 else float("inf"))
    return {"name": name, "ok": bool(ok), "bins": int(r.size), "max_dev": maxdev,
            "differing_bins": int((diff[1:-1] != 0).sum()), "moved_fraction": float(moved),
            "entries_num": int(num.sum()), "entries_den": int(den.sum()), "underflow_ratio": under,
            "overflow_ratio": over}


def plot(dirname, name, runs, den_label):
    """Distributions and ratios to the denominator (plot-histograms.py:146-196), when matplotlib exists."""
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
    except ImportError:
        return False
    os.makedirs(dirname, exist_ok=True)
    lo, hi, den = runs[den_label]
    nb = len(den) - 2
    x = lo + np.arange(nb) * (hi - lo) / nb
    fig, (a1, a2) = plt.subplots(2, 1, sharex=True, figsize=(7, 6))
    for label, (_, _, c) in runs.items():
        a1.bar(x, c[1:-1], width=(hi - lo) / nb, align="edge", alpha=0.5, label=label, log=True)
        if label != den_label:
            i, r = ratio(c, den)
            a2.plot(x[i], r, ".", label=f"{label}/{den_label}")
    a1.set_ylabel("entries")
    a1.legend()
    a2.set_xlabel(name)
    a2.set_ylabel("ratio")
    a2.grid()
    fig.savefig(os.path.join(dirname, f"{name}.png"))
    plt.close(fig)
    return True


def main(argv=None):
    ap = argparse.ArgumentParser(description="Compare HistoValidator histograms of several runs bin by bin.")
    ap.add_argument("files", nargs="+", help="histograms_<label>.txt files")
    ap.add_argument("--den", required=True, help="label of the denominator run")
    ap.add_argument("--tolerance", type=float, default=0.0,
                    help="largest |ratio - 1| of a bin for the histograms of exact products (default 0)")
    ap.add_argument("--approx-tolerance", type=float, default=1e-3,
                    help="largest fraction of entries in other bins for the fitted quantities")
    ap.add_argument("--plot", metavar="DIR", help="write PNG plots of the distributions and ratios")
    args = ap.parse_args(argv)
    runs = {label_of(f): read(f) for f in args.files}
    if args.den not in runs:
        sys.exit(f"no histograms_{args.den}.txt among the files")
    bad = 0
    den = runs[args.den]
    for label, h in runs.items():
        if label == args.den:
            continue
        for name in sorted(set(den) | set(h)):
            if name not in den or name not in h:
                print(f"{label}/{args.den} {name}: missing in {'the numerator' if name not in h else 'the denominator'}")
                bad += 1
                continue
            c = compare(h[name][2], den[name][2], name, args.tolerance, args.approx_tolerance)
            bad += not c["ok"]
            print(f"{label}/{args.den} {name}: {'ok ' if c['ok'] else 'DIFF'} bins {c['bins']} "
                  f"max|r-1| {c['max_dev']:.3g} differing {c['differing_bins']} moved {c['moved_fraction']:.2e} "
                  f"entries {c['entries_num']}/{c['entries_den']} under {c['underflow_ratio']:.3g} "
                  f"over {c['overflow_ratio']:.3g}")
    if args.plot:
        names = set.intersection(*(set(h) for h in runs.values()))
        made = [plot(args.plot, name, {lab: h[name] for lab, h in runs.items()}, args.den) for name in sorted(names)]
        if not any(made):
            print("matplotlib is not importable: no plots written", file=sys.stderr)
    print(f"{bad} histogram(s) differ beyond the tolerance")
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())

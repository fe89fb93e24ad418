"""Pins of the oracle's broken-line fit (BrokenLine.h) and density vertex
clustering (gpuClusterTracksByDensity.h) against what the mathematics fixes,
with checks chosen so that a plausible transcription mistake fails one of them.

* MultScatt (BrokenLine.h:46-59) against the Highland formula in its PDG form
  with the reference's material model (0.06 radiation lengths per 16 cm, a
  geometry factor of 0.7, pT capped at 20 GeV).
* The fitted covariance against J V J^T, J the finite-difference Jacobian of
  the fitted parameters with respect to the hit coordinates and V the hit
  covariance, where scattering is negligible.
* chi2 / (2N - 5) averages 1 and the parameter pulls are N(0, 1) on noisy
  helices without scattering.
* The pulls are N(0, 1) on helices that scatter the way the broken line
  models it (a kink at the first hit with the first segment's variance, kinks
  at the inner hits with the variances of their two segments), where
  scattering dominates the hit errors.
* Density clustering brute-forced from its definition on overlapping vertices
  generated like test/VertexFinder_t.h:30-80.

Each check is also run with one planted mistake in the oracle
(pyoracle.mutation) and must then fail.
"""
import math

import numpy as np
import pytest

import pyoracle

B = float(np.float32(0.0114256972711507))  # the reference's field constant (GeV/cm)
RADII = [2.9, 6.8, 10.9, 16.0]
M = pyoracle.mutation


# ---- multiple scattering ----------------------------------------------------------------

def highland_theta0_sq(p_gev, x_over_x0):
    """Highland / Lynch-Dahl (PDG): theta0 = 13.6 MeV / (beta c p) z sqrt(x/X0)
    [1 + 0.038 ln(x z^2 / (X0 beta^2))], beta = 1, z = 1 (planar-angle RMS)."""
    theta0 = 0.0136 / p_gev * math.sqrt(x_over_x0) * (1 + 0.038 * math.log(x_over_x0))
    return theta0 * theta0


def bl_kink_variance(length, R, slope):
    """The broken line's planar-angle variance of a segment: 0.7 x Highland theta0^2 for
    0.06 X0 per 16 cm, momentum from pT = min(20 GeV, B R) and the slope cot(theta)."""
    pt = min(20.0, B * R)
    p = pt * math.sqrt(1 + slope * slope)
    return 0.7 * highland_theta0_sq(p, abs(length) * 0.06 / 16.0)


def _ms_grid_ok(oracle):
    for length in (0.3, 1.0, 4.0, 9.0, 15.0):
        for R in (30.0, 120.0, 1500.0, 5000.0):  # 0.34 ... 57 GeV (capped at 20)
            for slope in (0.0, -0.6, 2.2):
                got = oracle.mult_scatt(length, B, R, slope)
                exp = bl_kink_variance(length, R, slope)
                if abs(got - exp) > 1e-12 * exp:
                    return False
    return True


def test_mult_scatt_is_highland(oracle):
    assert _ms_grid_ok(oracle)
    with M(M.XXI0):
        assert not _ms_grid_ok(oracle)


# ---- helix generation --------------------------------------------------------------------

def hit_covariance(x, y, s_rphi, s_z):
    """Global 3x3 covariance (xx, xy, yy, xz, yz, zz) of a barrel hit measured along the layer's
    r-phi direction and along z."""
    r = math.hypot(x, y)
    tx, ty = -y / r, x / r
    return np.array([s_rphi ** 2 * tx * tx, s_rphi ** 2 * tx * ty, s_rphi ** 2 * ty * ty, 0.0, 0.0, s_z ** 2])


def propagate(pt, phi0, cot, z0, q, n, rng=None, s_rphi=0.0, s_z=0.0, scatter=False):
    """Hits of a helix from (0, 0, z0) on the first n barrel layers; with the field along +z a
    positive charge turns clockwise (dphi/ds = -q/R).  Optionally the hits are smeared by the
    measurement errors and the track scatters at the hits as the broken line models it: a
    kink at the first hit with the first segment's variance, at each inner hit with the sum
    of its two segments' variances (planar angles, one in the transverse plane, one in the
    plane of the track and the beam line)."""
    R = pt / B
    kappa = -q / R
    x, y, z, phi, c = 0.0, 0.0, z0, phi0, cot
    pts = []
    truth_s = []
    for L in range(n):
        rL = RADII[L]

        def at(s):
            return (x + (math.sin(phi + kappa * s) - math.sin(phi)) / kappa,
                    y - (math.cos(phi + kappa * s) - math.cos(phi)) / kappa)
        lo, hi = 0.0, 2.0 * rL
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            xm, ym = at(mid)
            if xm * xm + ym * ym < rL * rL:
                lo = mid
            else:
                hi = mid
        s = 0.5 * (lo + hi)
        x, y = at(s)
        z = z + c * s
        phi = phi + kappa * s
        pts.append([x, y, z])
        truth_s.append(s * math.sqrt(1 + c * c))
        if scatter and L < n - 1:
            pass  # kinks are applied below, once the segment lengths are known
        if scatter:
            pts[-1].append((phi, c))
    hits = np.array([p[:3] for p in pts]).T
    return hits, R


def scattered_helix(pt, phi0, cot, z0, q, n, rng, s_rphi, s_z):
    """Measured hits of a helix that scatters at the hits as the broken line models it (see
    propagate); the kink variances use the unscattered helix's segment lengths."""
    clean, R = propagate(pt, phi0, cot, z0, q, n)
    S = np.concatenate([[0.0], np.cumsum(np.sqrt(np.sum(np.diff(clean, axis=1) ** 2, axis=0)))])
    var = np.zeros(n)
    var[0] = bl_kink_variance(S[1] - S[0], R, cot)
    for i in range(1, n - 1):
        var[i] = bl_kink_variance(S[i + 1] - S[i], R, cot) + bl_kink_variance(S[i] - S[i - 1], R, cot)
    kappa = -q / R
    x, y, z, phi, c = 0.0, 0.0, z0, phi0, cot
    hits = np.zeros((3, n))
    ge = np.zeros((6, n), np.float32)
    for L in range(n):
        rL = RADII[L]

        def at(s):
            return (x + (math.sin(phi + kappa * s) - math.sin(phi)) / kappa,
                    y - (math.cos(phi + kappa * s) - math.cos(phi)) / kappa)
        lo, hi = 0.0, 2.0 * rL
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            xm, ym = at(mid)
            if xm * xm + ym * ym < rL * rL:
                lo = mid
            else:
                hi = mid
        s = 0.5 * (lo + hi)
        x, y = at(s)
        z = z + c * s
        phi = phi + kappa * s
        r = math.hypot(x, y)
        g1, g2 = rng.normal(size=2)
        hits[:, L] = (x - y / r * s_rphi * g1, y + x / r * s_rphi * g1, z + s_z * g2)
        ge[:, L] = hit_covariance(x, y, s_rphi, s_z)
        if L < n - 1 and var[L] > 0:
            d1, d2 = rng.normal(scale=math.sqrt(var[L]), size=2)
            phi += d1 * math.sqrt(1 + c * c)  # transverse kink: planar angle / sin(theta)
            th = math.atan2(1.0, c) + d2
            c = math.cos(th) / math.sin(th)
    return hits, ge


# ---- covariance against propagated hit errors -----------------------------------------

def fitted(oracle, hits, ge):
    c, cc, ln, lc, _, _ = oracle.fit_helix_d(hits, ge, B)
    return np.array([c[0], c[1], c[2], ln[0], ln[1]]), cc, lc


def hit_cov_matrix(ge):
    n = ge.shape[1]
    V = np.zeros((3 * n, 3 * n))
    for i in range(n):
        xx, xy, yy, xz, yz, zz = (float(v) for v in ge[:, i])
        V[3 * i:3 * i + 3, 3 * i:3 * i + 3] = [[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]]
    return V


def jacobian(f, hits, h=1e-4, angles=(0,)):
    """Central finite-difference Jacobian of f(hits) (a vector) with respect to the 3N hit
    coordinates, x, y, z of each hit in turn."""
    n = hits.shape[1]
    cols = []
    for i in range(n):
        for a in range(3):
            hp, hm = hits.copy(), hits.copy()
            hp[a, i] += h
            hm[a, i] -= h
            d = f(hp) - f(hm)
            for k in angles:
                d[k] = math.remainder(d[k], 2 * math.pi)
            cols.append(d / (2 * h))
    return np.array(cols).T


def barrel_errors(hits, s_rphi, s_z):
    n = hits.shape[1]
    return np.array([hit_covariance(hits[0, i], hits[1, i], s_rphi, s_z) for i in range(n)], np.float32).T.copy()


def close(got, exp, rtol):
    scale = np.sqrt(np.outer(np.abs(np.diag(exp)), np.abs(np.diag(exp))))
    return bool(np.all(np.abs(got - exp) <= rtol * scale))


HELICES = [(30.0, 0.4, 0.3, 1.0, 1), (50.0, -2.2, 1.3, -4.0, -1), (25.0, 2.9, -2.1, 2.0, 1)]


def _line_cov_ok(oracle, n):
    for pt, phi0, cot, z0, q in HELICES:
        hits, _ = propagate(pt, phi0, cot, z0, q, n)
        ge = barrel_errors(hits, 50e-4, 1000e-4)  # scattering negligible next to 1 mm in z
        J = jacobian(lambda x: fitted(oracle, x, ge)[0][3:], hits, angles=())
        if not close(fitted(oracle, hits, ge)[2], J @ hit_cov_matrix(ge) @ J.T, 0.005):
            return False
    return True


@pytest.mark.parametrize("n", [3, 4])
def test_line_covariance_is_propagated_hit_errors(oracle, n):
    """Line fit (BrokenLine.h:401-490): where scattering is negligible, the covariance of
    (cot theta, zip) equals J V J^T within 0.5% per element (relative to sqrt(C_ii C_jj)),
    J the finite-difference Jacobian with respect to the hit coordinates."""
    assert _line_cov_ok(oracle, n)
    with M(M.LINE_WEIGHT):
        assert not _line_cov_ok(oracle, n)


# ---- the circle fit's covariance, stage by stage -------------------------------------------

def circle_centre(par):
    """Centre of the Karimaki circle (phi, d, k) relative to its origin: the circle
    (k/2)(x^2 + y^2) - (1 + k d)(x sin phi - y cos phi) + k d^2 / 2 + d = 0."""
    phi, d, k = par
    return (1.0 / k + d) * np.array([math.sin(phi), -math.cos(phi)])


def translate_circle(par, x0, y0):
    """(phi, d, k) of the same circle seen from a new origin, in which the old origin has
    coordinates (x0, y0): from the centre, with the direction of travel kept."""
    phi, d, k = par
    c = circle_centre(par) + np.array([x0, y0])
    a = math.copysign(math.hypot(c[0], c[1]), 1.0 / k + d)
    n = c / a
    return np.array([math.atan2(n[0], -n[1]), a - 1.0 / k, k])


def par_jacobian(f, par):
    """Central finite-difference Jacobian of f with respect to (phi, d, k); the steps keep the
    rounding of d = |centre| - 1/k (two terms of size 1/k) well below the result."""
    h = (1e-7, 1e-6, 1e-5 * abs(par[2]))
    cols = []
    for a in range(3):
        p, m = np.array(par, float), np.array(par, float)
        p[a] += h[a]
        m[a] -= h[a]
        d = f(p) - f(m)
        d[0] = math.remainder(d[0], 2 * math.pi)
        cols.append(d / (2 * h[a]))
    return np.array(cols).T


def test_translate_karimaki_is_circle_translation(oracle):
    """TranslateKarimaki (BrokenLine.h:85-108) moves the circle's origin like the geometric
    translation (through the centre) and propagates the covariance with its Jacobian."""
    rng = np.random.default_rng(5)
    for _ in range(40):
        par = np.array([rng.uniform(-3, 3), rng.uniform(-0.3, 0.3), rng.choice([-1, 1]) * rng.uniform(1e-3, 2e-2)])
        a = rng.normal(size=(3, 3))
        cov = a @ a.T * 1e-6 + np.diag([1e-6, 1e-5, 1e-9])
        x0, y0 = rng.uniform(-8, 8, size=2)
        got_par, got_cov = oracle.translate_karimaki(par, cov, x0, y0)
        exp_par = translate_circle(par, x0, y0)
        assert abs(math.remainder(got_par[0] - exp_par[0], 2 * math.pi)) < 1e-10
        assert abs(got_par[1] - exp_par[1]) < 1e-9 and got_par[2] == par[2]
        T = par_jacobian(lambda p: translate_circle(p, x0, y0), par)
        assert close(got_cov, T @ cov @ T.T, 1e-4)


def _circle_ok(oracle, n, pt, phi0, cot, z0, q, s_rphi, s_z, with_hit_errors):
    hits, _ = propagate(pt, phi0, cot, z0, q, n)
    ge = barrel_errors(hits, s_rphi, s_z)
    dbg = oracle.circle_debug(hits, ge, B)
    ok = True
    if with_hit_errors:
        # (phi, k) at the chord midpoint of the corrected first two hits carry the propagated
        # hit errors once d(phi)/d(u0) has the sign of the derivative of atan2(e - d); the
        # reference's expression (BrokenLine.h:348) has the opposite sign, which flips the
        # sign of the u0 terms of phi's variance and covariance
        J = jacobian(lambda x: oracle.circle_debug(x, ge, B)["par_mid"][[0, 2]], hits)
        P = J @ hit_cov_matrix(ge) @ J.T
        with pyoracle.variant(pyoracle.variant.PHI_JACOBIAN_SIGN):
            fixed = oracle.circle_debug(hits, ge, B)["cov_mid"][np.ix_([0, 2], [0, 2])]
        ok &= close(fixed, P, 0.005)
        ok &= not close(dbg["cov_mid"][np.ix_([0, 2], [0, 2])], P, 0.05)  # the reference's sign
        ok &= abs(dbg["cov_mid"][2, 2] - P[1, 1]) <= 0.005 * P[1, 1]     # k does not depend on it
    # the sagitta at the midpoint enters with the reference's (q / 2) in integer arithmetic: 0
    ok &= bool(np.all(dbg["cov_mid"][1] == 0) and np.all(dbg["cov_mid"][:, 1] == 0))
    # scattering of the first segment added to phi at the first hit
    ok &= abs(dbg["ms"] - (1 + dbg["slope"] ** 2) * bl_kink_variance(dbg["S1"] - dbg["S0"], dbg["R"],
                                                                    dbg["slope"])) <= 1e-12 * dbg["ms"]
    # moved to the first hit, the scattering added, moved to the beam line (origin)
    e_d = dbg["e"] - dbg["d"]
    T1 = par_jacobian(lambda p: translate_circle(p, 0.5 * e_d[0], 0.5 * e_d[1]), dbg["par_mid"])
    at_first = translate_circle(dbg["par_mid"], 0.5 * e_d[0], 0.5 * e_d[1])
    T2 = par_jacobian(lambda p: translate_circle(p, dbg["d"][0], dbg["d"][1]), at_first)
    C = T1 @ dbg["cov_mid"] @ T1.T
    C[0, 0] += dbg["ms"]
    C = T2 @ C @ T2.T
    got = fitted(oracle, hits, ge)[1]
    ok &= close(got, C, 1e-4)
    return ok


@pytest.mark.parametrize("n", [3, 4])
def test_circle_covariance_stage_by_stage(oracle, n):
    """Circle fit (BrokenLine.h:256-382): (phi, k) at the chord midpoint carry the propagated
    hit errors (scattering negligible: pT above the 20 GeV cap, 100 um errors) up to the
    reference's sign of d(phi)/d(u0); the covariance at the beam line is that one moved to the
    first hit, plus the first segment's scattering ((1 + cot^2) x the Highland kink
    variance), moved to the origin, with translation Jacobians of the geometric circle
    translation taken at the fitted points.  The same composition holds where scattering
    dominates (1 GeV, 10 um errors)."""
    for h in HELICES:
        assert _circle_ok(oracle, n, *h, 100e-4, 200e-4, True)
    low = [(1.0, 0.4, 1.2, 1.0, 1), (0.8, -2.0, -0.7, -3.0, -1)]
    for h in low:
        assert _circle_ok(oracle, n, *h, 10e-4, 20e-4, False)
    with M(M.NO_FIRST_MS_CIRCLE):
        assert not all(_circle_ok(oracle, n, *h, 10e-4, 20e-4, False) for h in low)
    with M(M.XXI0):
        assert not all(_circle_ok(oracle, n, *h, 10e-4, 20e-4, False) for h in low)
    for mut in (M.CU_BORDER, M.PHI_JACOBIAN):
        with M(mut):
            assert not all(_circle_ok(oracle, n, *h, 100e-4, 200e-4, True) for h in HELICES), f"mutation {mut}"


# ---- chi2 and pulls --------------------------------------------------------------------------

def pulls(oracle, cfg, ntracks, n, s_rphi, s_z, seed, scatter):
    """Fitted minus true parameters over their fitted errors, for tracks whose charge the fit
    measures right (a few nearly straight tracks have their charge flipped by the hit
    errors); returns the pulls, chi2 / ndf and the fraction of flipped tracks."""
    rng = np.random.default_rng(seed)
    pl, chi2, flipped = [], [], 0
    for k in range(ntracks):
        pt, phi0, cot, z0, q = cfg(rng)
        if scatter:
            hits, ge = scattered_helix(pt, phi0, cot, z0, q, n, rng, s_rphi, s_z)
        else:
            clean, _ = propagate(pt, phi0, cot, z0, q, n)
            ge = barrel_errors(clean, s_rphi, s_z)
            hits = clean.copy()
            for i in range(n):
                x, y = clean[0, i], clean[1, i]
                r = math.hypot(x, y)
                g1, g2 = rng.normal(size=2)
                hits[:, i] += (-y / r * s_rphi * g1, x / r * s_rphi * g1, s_z * g2)
        st, cov, _, _, c2 = oracle.fit_helix(hits, ge, B)
        st = np.asarray(st, np.float64)
        t = np.array([phi0, 0.0, q / pt, cot, z0])
        if st[2] * t[2] < 0:
            flipped += 1
            continue
        d = st - t
        d[0] = math.remainder(d[0], 2 * math.pi)
        pl.append(d / np.sqrt(np.asarray(cov, np.float64)[[0, 5, 9, 12, 14]]))
        chi2.append(c2)
    return np.array(pl), np.array(chi2), flipped / ntracks


def _high_pt(rng):
    return (rng.uniform(20.0, 30.0), rng.uniform(-math.pi, math.pi), rng.uniform(-1.5, 1.5),
            rng.uniform(-5, 5), rng.choice([-1, 1]))


def _low_pt(rng):
    return (rng.uniform(0.8, 1.2), rng.uniform(-math.pi, math.pi), rng.choice([-1.0, 1.0]) * rng.uniform(0.9, 1.5),
            rng.uniform(-5, 5), rng.choice([-1, 1]))


def _no_scatter_ok(oracle, n, ntracks=2500):
    pl, chi2, flipped = pulls(oracle, _high_pt, ntracks, n, 15e-4, 40e-4, seed=n, scatter=False)
    w = pl.std(axis=0)
    # cot(theta) and zip; q/pt is tested where scattering dominates (below), since at these
    # momenta the sagitta is small enough that the hit errors flip the charge of a few
    # percent of the triplets, and removing them cuts the tails of its pulls; phi and tip
    # are pinned stage by stage above (the reference moves the circle between origins held
    # fixed, which leaves the fitted origins' own errors out)
    ok = abs(chi2.mean() - 1) < 0.06 and flipped < 0.05
    ok &= bool(np.all(np.abs(pl.mean(axis=0)[3:]) < 0.08)) and bool(np.all(np.abs(w[3:] - 1) < 0.05))
    return ok, chi2.mean(), w, flipped


@pytest.mark.parametrize("n", [3, 4])
def test_chi2_and_pulls_without_scattering(oracle, n):
    """Noisy helices where scattering is small next to the hit errors (20-30 GeV, 15/40 um):
    chi2 / (2N - 5) averages 1 and the pulls of cot(theta) and zip have mean 0, width 1."""
    ok, m, w, f = _no_scatter_ok(oracle, n)
    assert ok, f"mean chi2/ndf {m:.3f}, pull widths {np.round(w, 3)}, charge flipped {f:.4f}"
    with M(M.CHI2_NDF):
        assert not _no_scatter_ok(oracle, n, 800)[0]
    with M(M.LINE_WEIGHT):
        assert not _no_scatter_ok(oracle, n, 800)[0]


def _scatter_ok(oracle, n, ntracks=2500):
    pl, chi2, flipped = pulls(oracle, _low_pt, ntracks, n, 10e-4, 20e-4, seed=10 + n, scatter=True)
    w = pl.std(axis=0)
    ok = bool(np.all(np.abs(w[2:] - 1) < 0.05)) and bool(np.all(np.abs(w[:2] - 1) < 0.10))
    ok &= bool(np.all(np.abs(pl.mean(axis=0)) < 0.08)) and abs(chi2.mean() - 1) < 0.08 and flipped == 0
    return ok, w, chi2.mean()


@pytest.mark.parametrize("n", [3, 4])
def test_pulls_with_scattering(oracle, n):
    """Helices of about 1 GeV that scatter as the broken line models it, with 10/20 um hit
    errors (scattering dominates): the pulls of q/pt, cot(theta) and zip have width 1 within
    5%, phi and tip within 10%, chi2 / ndf averages 1; a planted mistake in the scattering
    terms breaks that."""
    ok, w, m = _scatter_ok(oracle, n)
    assert ok, f"pull widths {np.round(w, 3)}, mean chi2/ndf {m:.3f}"
    planted = [M.XXI0, M.NO_FIRST_MS_CIRCLE, M.NO_SLOPE_FACTOR, M.NO_FIRST_MS_LINE]
    if n == 4:
        planted.append(M.VARBETA_ONE_SIDE)  # only quadruplets have a kink between two measured segments
    for mut in planted:
        with M(mut):
            okm, wm, _ = _scatter_ok(oracle, n, 800)
            assert not okm, f"mutation {mut} not caught: pull widths {np.round(wm, 3)}"


# ---- density clustering of the vertex finder ---------------------------------------------------

def vertex_finder_event(rng, nav):
    """Tracks as test/VertexFinder_t.h:30-80 generates them: Poisson(nav) vertices at
    z ~ 3.5 N(0, 1), Poisson(10) tracks each with errors U(0.005, 0.025) cm, and
    2 Poisson(10) noise tracks uniform in (-13, 13) cm with error 0.03 cm."""
    zt, ez = [], []
    for zv in 3.5 * rng.normal(size=rng.poisson(nav)):
        for _ in range(rng.poisson(10)):
            e = rng.uniform(0.005, 0.025)
            zt.append(zv + e * rng.normal())
            ez.append(e * e)
    for _ in range(2 * rng.poisson(10)):
        zt.append(rng.uniform(-13, 13))
        ez.append(0.03 * 0.03)
    return np.array(zt, np.float32), np.array(ez, np.float32)


def brute_density_clusters(zt, ezt2, eps=np.float32(0.07), errmax2=np.float32(0.01 * 0.01), chi2max=np.float32(9.0),
                           min_t=2):
    """clusterTracksByDensity (gpuClusterTracksByDensity.h:20-207) from its definition, over all
    pairs: tracks whose 1 mm z bins (int(10 z), clamped to a byte) are adjacent are
    neighbours within eps and the chi2 cut; a track with a small enough error counts its
    neighbours; every track links to its nearest neighbour of higher count (or equal count
    and lower z) within the cuts; the chains end at roots, which with at least min_t
    neighbours seed clusters numbered by root index; the other roots are noise (9997).
    Ties of the float32 distance go to the last candidate in the reference's histogram order."""
    nt = zt.size
    izt = np.clip(np.trunc(zt.astype(np.float64) * 10).astype(np.int64), -128, 127)
    dist = np.abs(zt[:, None] - zt[None, :])  # float32, as the reference computes it
    near = (np.abs(izt[:, None] - izt[None, :]) <= 1) & (dist <= eps) & \
        (dist * dist <= chi2max * (ezt2[:, None] + ezt2[None, :]))
    np.fill_diagonal(near, False)
    nn = np.where(ezt2 <= errmax2, near.sum(axis=1), 0)
    parent = np.arange(nt)
    for i in range(nt):
        denser = near[i] & ((nn > nn[i]) | ((nn == nn[i]) & (zt < zt[i])))
        cand = np.nonzero(denser)[0]
        if cand.size:
            d = dist[i, cand]
            best = cand[d == d.min()]
            # float32 distances can tie: the reference takes the last of them in its histogram's
            # order, bins ascending and, within a bin, track index descending (the bins are
            # filled by decrementing counters, HistoContainer.h:34-48)
            parent[i] = max(best, key=lambda j: (izt[j], -j))
    root = parent.copy()
    for _ in range(nt):
        nxt = root[root]
        if np.array_equal(nxt, root):
            break
        root = nxt
    seeds = [r for r in range(nt) if root[r] == r and nn[r] >= min_t]
    label = {r: k for k, r in enumerate(seeds)}
    iv = np.array([label.get(root[i], 9997) for i in range(nt)], np.int32)
    return iv, nn.astype(np.int32), len(seeds)


def _density_ok(oracle, events):
    for zt, ez in events:
        iv, nn, found = oracle.cluster_tracks(zt, ez, algo=0)
        biv, bnn, bfound = brute_density_clusters(zt, ez)
        if found != bfound or not np.array_equal(nn, bnn) or not np.array_equal(iv, biv):
            return False
    return True


def test_density_clustering_against_brute_force(oracle):
    """The vertex finder's default clustering on overlapping vertices (30, 50 and 70 vertices
    along 7 cm of beam line, tracks closer than eps from neighbouring vertices), against its
    definition evaluated over all track pairs; the tie rule and the chi2 cut matter."""
    rng = np.random.default_rng(2026)
    events = [vertex_finder_event(rng, nav) for nav in (30, 50, 70) for _ in range(3)]
    assert _density_ok(oracle, events)
    sizes = [oracle.cluster_tracks(zt, ez)[2] for zt, ez in events]
    assert min(sizes) > 10
    with M(M.DENSITY_TIE):
        assert not _density_ok(oracle, events)
    with M(M.DENSITY_NO_CHI2):
        assert not _density_ok(oracle, events)

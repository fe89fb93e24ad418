"""The oracle's Riemann fit (RiemannFit.h, the reference's off-by-default
alternative to the broken-line fit) pinned without the reference: Eigen is not
available, so its numbers are pinned by closed forms and by linear error
propagation, not against the reference itself (parity unpinned).

* the closed-form 3x3 symmetric eigen solver against numpy's LAPACK solver
  (double and float) and on degenerate spectra;
* the fast fit against the circumcircle of its three points and the helix pitch;
* noiseless helices: the full fit returns the generating parameters;
* the circle and line covariances against J V J^T, with J the finite-difference
  Jacobian of the fitted parameters with respect to the hit coordinates
  (momentum high and hit errors large, so that multiple scattering is small);
* the reference's own fit inputs (testRiemannFit.cc:51-87): the Riemann and the
  broken-line fits agree within their errors.
"""
import math

import numpy as np
import pytest

F = np.float32
B = float(F(0.0114256972711507))


@pytest.fixture(scope="module")
def orc(conditions):
    import pyoracle
    return pyoracle.Oracle(conditions)


def _helix_points(pt, phi0, cot, z0, q, radii, B):
    """Points of a helix from (0,0,z0) with transverse momentum pt (GeV) in field B (GeV/cm units);
    a positive charge turns clockwise seen from +z."""
    R = pt / B
    cx, cy = q * R * math.sin(phi0), -q * R * math.cos(phi0)
    pts = []
    for rr in radii:
        a = 2 * math.asin(rr / (2 * R))
        ang = math.atan2(-cy, -cx) - q * a
        pts.append((cx + R * math.cos(ang), cy + R * math.sin(ang), z0 + R * a * cot))
    return np.array(pts).T


def _sym(rng, spread):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q @ np.diag(spread) @ q.T


@pytest.mark.parametrize("seed", range(40))
def test_eigen3_double_matches_lapack(orc, seed):
    rng = np.random.default_rng(seed)
    spread = rng.choice([1e-6, 1e-2, 1.0, 3.0, 1e3], size=3) * rng.uniform(0.5, 2, size=3)
    m = _sym(rng, spread) + rng.uniform(-2, 2) * np.eye(3)
    ev, vec = orc.eigen3(m)
    ref_ev, ref_vec = np.linalg.eigh(m)
    scale = np.abs(m).max()
    assert np.all(np.diff(ev) >= 0)
    assert np.allclose(ev, ref_ev, atol=1e-13 * scale)
    # each eigenvector, and its agreement with LAPACK's up to sign: the closed form
    # loses accuracy as the square of the relative gap to the nearest eigenvalue
    for k in range(3):
        gap = min(abs(ref_ev[k] - ref_ev[j]) for j in range(3) if j != k) / scale
        if gap < 1e-6:
            continue
        v, r = vec[:, k], ref_vec[:, k]
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        assert np.abs(m @ v - ev[k] * v).max() < 1e-13 * scale / gap**2
        assert min(np.abs(v - r).max(), np.abs(v + r).max()) < 1e-13 / gap**2


@pytest.mark.parametrize("seed", range(20))
def test_eigen3_float_matches_lapack(orc, seed):
    rng = np.random.default_rng(100 + seed)
    m = _sym(rng, rng.uniform(0.1, 10, size=3) * np.array([1, 3, 9]))
    ev, vec = orc.eigen3(m, use_float=True)
    ref_ev, ref_vec = np.linalg.eigh(m.astype(np.float32).astype(np.float64))
    assert np.allclose(ev, ref_ev, atol=2e-5 * np.abs(m).max())
    v, r = vec[:, 0], ref_vec[:, 0]
    assert min(np.abs(v - r).max(), np.abs(v + r).max()) < 1e-4


@pytest.mark.parametrize("use_float", [False, True])
def test_eigen3_degenerate_spectra(orc, use_float):
    ev, vec = orc.eigen3(np.eye(3) * 2.5, use_float)
    assert np.allclose(ev, 2.5) and np.allclose(vec, np.eye(3))
    rng = np.random.default_rng(7)
    for spread in ([1, 1, 4], [2, 5, 5], [0, 0, 1e-3]):
        m = _sym(rng, np.array(spread, float))
        ev, vec = orc.eigen3(m, use_float)
        # a double root is found to the square root of the precision
        tol = 1e-3 if use_float else 1e-7
        assert np.allclose(np.sort(ev), np.sort(spread), atol=tol * max(spread))
        assert np.allclose(m @ vec, vec * ev, atol=10 * tol * max(spread))
        assert np.allclose(np.abs(np.linalg.det(vec)), 1, atol=10 * tol)


@pytest.mark.parametrize("n", [3, 4, 5])
def test_fast_fit_is_the_circumcircle(orc, n):
    radii = [2.9, 6.8, 10.9, 16.0, 21.0][:n]
    for pt, phi0, cot, q in ((0.8, 0.4, 0.7, 1), (3.0, -2.2, -1.5, -1), (12.0, 1.6, 0.05, 1)):
        hits = _helix_points(pt, phi0, cot, 0.3, q, radii, B)
        ge = np.full((6, n), 1e-6, np.float32)
        fast = orc.circle_fit_riemann(hits, ge, B)[0]
        for i in (0, n // 2, n - 1):
            d = math.hypot(hits[0, i] - fast[0], hits[1, i] - fast[1])
            assert abs(d / fast[2] - 1) < 1e-10
        assert abs(fast[2] / (pt / B) - 1) < 1e-9
        # tan(theta) from the arc between the first and the last hit: 1 / cot(theta),
        # signed by the sense of rotation
        assert abs(abs(fast[3]) - 1 / abs(cot)) < 1e-8 * (1 / abs(cot))
        assert np.sign(fast[3]) == -q * np.sign(cot)


@pytest.mark.parametrize("n", [3, 4, 5])
@pytest.mark.parametrize("pt,phi0,cot,z0,q", [(1.0, 0.3, 0.4, 1.2, 1), (5.0, -2.0, -1.1, -3.0, -1),
                                              (0.7, 2.5, 2.0, 0.5, 1), (20.0, 1.0, 0.1, 0.0, -1)])
def test_riemann_recovers_noiseless_helix(orc, n, pt, phi0, cot, z0, q):
    hits = _helix_points(pt, phi0, cot, z0, q, [2.9, 6.8, 10.9, 16.0, 21.0][:n], B)
    ge = np.zeros((6, n), np.float32)
    ge[0] = ge[2] = 1e-7
    ge[5] = 4e-6
    st, cov, fpt, eta, chi2 = orc.fit_helix_riemann(hits, ge, B)
    assert abs(math.remainder(st[0] - phi0, 2 * math.pi)) < 1e-5
    assert abs(st[1]) < 1e-5
    assert abs(st[2] * pt * q - 1) < 1e-5
    assert abs(st[3] - cot) < 1e-5
    assert abs(st[4] - z0) < 1e-5
    assert abs(fpt / pt - 1) < 1e-5
    assert abs(eta - math.asinh(cot)) < 1e-5
    assert 0 <= chi2 < 1e-3
    C5 = np.zeros((5, 5))
    C5[np.triu_indices(5)] = cov
    C5 = C5 + np.triu(C5, 1).T
    assert np.all(np.linalg.eigvalsh(C5[:3, :3]) > -1e-9 * np.abs(C5[:3, :3]).max())
    assert np.all(np.linalg.eigvalsh(C5[3:, 3:]) > 0)


def _jacobian(f, x0, steps):
    cols = []
    for k in range(x0.size):
        d = np.zeros_like(x0)
        d.flat[k] = steps.flat[k]
        cols.append((f(x0 + d) - f(x0 - d)) / (2 * steps.flat[k]))
    return np.array(cols).T


@pytest.mark.parametrize("n,pt,err", [(3, 20.0, 1e-3), (4, 20.0, 1e-3), (5, 20.0, 1e-3), (3, 1.0, 3e-2),
                                      (3, 3.0, 1e-2), (4, 3.0, 3e-2)])
def test_circle_covariance_is_linear_propagation(orc, n, pt, err):
    # within 1.1% (scattering, and the reference's second-order terms, account for up
    # to 0.8%); a sign error in the x-y covariance terms of the paraboloid points moves
    # the low-momentum cases by 1.5%
    """Hit errors large for the momentum (0.3-0.6 mm at 20 GeV, 1-3 mm at 1-3 GeV):
    scattering adds about 1% or less to the hit variances, so the circle covariance
    is J V J^T of the hit errors alone.
    (Where scattering dominates, the reference's propagation departs from J V J^T by
    design: its covariance of the scaled points adds the scale factor's variance
    instead of projecting it out, RiemannFit.h:604-607; those terms stay unpinned.)"""
    rng = np.random.default_rng(n)
    for trial in range(4):
        phi0, cot, q = rng.uniform(-3, 3), rng.uniform(-2, 2), rng.choice([-1, 1])
        hits = _helix_points(pt, phi0, cot, 0.0, q, [2.9, 6.8, 10.9, 16.0, 21.0][:n], B)
        ge = np.zeros((6, n), np.float32)
        ge[0] = rng.uniform(1, 4, n) * err
        ge[2] = rng.uniform(1, 4, n) * err
        ge[1] = rng.uniform(-0.5, 0.5, n) * np.sqrt(ge[0] * ge[2])
        ge[5] = 1e-4
        par0 = orc.circle_fit_riemann(hits, ge, B)[1]

        def par(xy):
            h = hits.copy()
            h[:2] = xy.reshape(2, n)
            return orc.circle_fit_riemann(h, ge, B)[1]

        J = _jacobian(par, hits[:2].copy(), np.full((2, n), 1e-6))
        V = np.zeros((2 * n, 2 * n))
        for i in range(n):
            V[i, i], V[i + n, i + n] = ge[0, i], ge[2, i]
            V[i, i + n] = V[i + n, i] = ge[1, i]
        expect = J @ V @ J.T
        cov = orc.circle_fit_riemann(hits, ge, B)[2]
        assert np.allclose(cov, cov.T, rtol=1e-9, atol=0)
        sig = np.sqrt(np.diag(expect))
        rel = (cov - expect) / np.outer(sig, sig)
        assert np.abs(rel).max() < 0.011, (trial, rel, par0)


@pytest.mark.parametrize("n", [3, 4])
def test_line_covariance_is_linear_propagation(orc, n):
    """z errors only (transverse errors negligible, pt 20 GeV): the (cot theta, zip)
    covariance is J V J^T of the z errors."""
    rng = np.random.default_rng(10 + n)
    for trial in range(4):
        phi0, cot, q = rng.uniform(-3, 3), rng.uniform(-2, 2), rng.choice([-1, 1])
        hits = _helix_points(20.0, phi0, cot, rng.uniform(-5, 5), q, [2.9, 6.8, 10.9, 16.0][:n], B)
        ge = np.zeros((6, n), np.float32)
        ge[0] = ge[2] = 1e-12
        ge[5] = rng.uniform(1, 4, n) * 1e-4

        def line(z):
            h = hits.copy()
            h[2] = z
            st = orc.fit_helix_riemann(h, ge, B)[0]
            return np.array([st[3], st[4]], np.float64)

        J = _jacobian(line, hits[2].copy(), np.full(n, 1e-3))
        expect = J @ np.diag(ge[5].astype(np.float64)) @ J.T
        cov = orc.fit_helix_riemann(hits, ge, B)[1]
        got = np.array([[cov[12], cov[13]], [cov[13], cov[14]]])
        sig = np.sqrt(np.diag(expect))
        assert np.abs((got - expect) / np.outer(sig, sig)).max() < 0.02, (trial, got, expect)


def _reference_inputs(n):
    """The fit inputs of the reference's testRiemannFit.cc:51-87 (hits in cm, global
    errors xx, xy, yy, xz, yz, zz): a five-hit track, and a three/four-hit one."""
    if n == 5:
        hits = np.array([[2.934787, 6.314229, 8.936963, 10.360559, 12.856387],
                         [0.773211, 1.816356, 2.765734, 3.330824, 4.422212],
                         [-10.980247, -23.162731, -32.759060, -38.061260, -47.518867]])
        ge = np.array([[1.424715e-07, -4.996975e-07, 1.752614e-06, 3.660689e-11, 1.644638e-09, 7.346080e-05],
                       [6.899177e-08, -1.873414e-07, 5.087101e-07, -2.078806e-10, -2.210498e-11, 4.346079e-06],
                       [1.406273e-06, 4.042467e-07, 6.391180e-07, -3.141497e-07, 6.513821e-08, 1.163863e-07],
                       [1.176358e-06, 2.154100e-07, 5.072816e-07, -8.161219e-08, 1.437878e-07, 5.951832e-08],
                       [2.852843e-05, 7.956492e-06, 3.117701e-06, -1.060541e-06, 8.777413e-09, 1.426417e-07]],
                      np.float32).T
        return hits, ge
    hits = np.array([[1.98645, 4.72598, 7.65632, 11.3151], [2.18002, 4.88864, 7.75845, 11.3134],
                     [2.46338, 6.99838, 11.808, 17.793]])[:, :n]
    ge = np.zeros((6, 4), np.float32)
    ge[0] = [7.14652e-06, 2.15789e-06, 1.63328e-06, 6.27919e-06]
    ge[2] = [6.10348e-06, 2.08211e-06, 1.61672e-06, 6.28081e-06]
    ge[5] = [5.184e-05, 1.444e-05, 6.25e-06, 3.136e-05]
    ge[1] = [-5.60077e-06, -1.11936e-06, -6.24945e-07, -5.28e-06]
    return hits, np.ascontiguousarray(ge[:, :n])


@pytest.mark.parametrize("n", [3, 4, 5])
def test_riemann_and_broken_line_agree_on_reference_inputs(orc, n):
    hits, ge = _reference_inputs(n)
    Bt = 0.0113921  # testRiemannFit.cc:91
    sr, cr, ptr, etar, chi2r = orc.fit_helix_riemann(hits, ge, Bt)
    sb, cb, ptb, etab, chi2b = orc.fit_helix(hits, ge, Bt)
    assert np.all(np.isfinite(sr)) and np.all(np.isfinite(cr)) and chi2r >= 0
    for k, d in ((0, 0), (1, 5), (2, 9), (3, 12), (4, 14)):
        sig = math.sqrt(max(cr[d], 0) + max(cb[d], 0))
        assert abs(sr[k] - sb[k]) < 3 * sig + 1e-6, (k, sr[k], sb[k], sig)
    assert abs(ptr / ptb - 1) < 0.1


def test_chain_with_riemann_fit(detector, conditions):
    """The chain with useRiemannFit (CAHitNtupletGeneratorOnGPU.cc:168-172): the same
    tuples as with the broken-line fit, and on them the two fits agree within their
    errors for nearly every track."""
    import pyoracle
    from conftest import event
    bl = pyoracle.Oracle(conditions)
    rf = pyoracle.Oracle(conditions, riemann_fit=True)
    raw, _ = event(detector, "pu50", 2, pu_lo=10, pu_hi=10)
    a, b = bl.process(raw, upto=4), rf.process(raw, upto=4)
    nt = a["nTuples"]
    assert b["nTuples"] == nt and np.array_equal(a["tupleBins"], b["tupleBins"])
    sa, sb = a["state"][:5 * nt].reshape(-1, 5), b["state"][:5 * nt].reshape(-1, 5)
    ca, cb = a["covariance"][:15 * nt].reshape(-1, 15), b["covariance"][:15 * nt].reshape(-1, 15)
    # the tuples of the multiplicity container are fitted, the others keep zeros
    fitted = np.nonzero(ca[:, 0] > 0)[0]
    assert fitted.size > 0.3 * nt and np.array_equal(fitted, np.nonzero(cb[:, 0] > 0)[0])
    assert np.all(np.isfinite(sb[fitted]))
    pulls = []
    for k, d in ((0, 0), (1, 5), (2, 9), (3, 12), (4, 14)):
        sig = np.sqrt(np.maximum(ca[fitted, d], 0) + np.maximum(cb[fitted, d], 0))
        pulls.append(np.abs(sa[fitted, k] - sb[fitted, k]) / sig)
    pulls = np.array(pulls)
    assert np.mean(pulls.max(axis=0) < 3) > 0.95, np.percentile(pulls.max(axis=0), [50, 90, 99])
    assert np.mean(b["chi2"][fitted] >= 0) == 1


def _ulps(a, b):
    return abs(a - b) / np.spacing(abs(b)) if b != 0 else abs(a) / np.spacing(1e-300)


def test_elementary_functions(orc):
    """The fit's own atan2, atan, sin, cos and log (IEEE basic operations only, the same
    code on the device) against the C library: within 4 units in the last place over
    the argument ranges of the fit, exact at the special points."""
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-4, 4, 4000), rng.uniform(-1e-3, 1e-3, 500), [0.0, 1.0, -1.0, math.pi / 4]])
    for x in xs:
        assert _ulps(orc.rmath("sin", x), math.sin(x)) <= 4 or abs(orc.rmath("sin", x) - math.sin(x)) < 1e-300, x
        assert _ulps(orc.rmath("cos", x), math.cos(x)) <= 4, x
    for x in np.concatenate([rng.uniform(-50, 50, 3000), 10.0 ** rng.uniform(-8, 8, 1000), [0.0, 1.0, -1.0]]):
        assert _ulps(orc.rmath("atan", x), math.atan(x)) <= 4, x
    for _ in range(3000):
        y, x = rng.normal(size=2) * 10.0 ** rng.uniform(-3, 3, size=2)
        assert _ulps(orc.rmath("atan2", x, y), math.atan2(y, x)) <= 4, (y, x)
    for y, x in ((0.0, 1.0), (1.0, 0.0), (-1.0, 0.0), (0.0, -1.0), (-0.0, -1.0), (2.0, -2.0)):
        assert orc.rmath("atan2", x, y) == math.atan2(y, x), (y, x)
    for x in np.concatenate([10.0 ** rng.uniform(-12, 6, 4000), [1.0, 2.0, 0.5, math.e]]):
        assert _ulps(orc.rmath("log", x), math.log(x)) <= 4, x
    assert orc.rmath("log", 1.0) == 0.0 and orc.rmath("log", 0.0) == -math.inf

// ORACLE -- test infrastructure only (see oracle.h).
// Stage C, fit part: the broken-line helix fit of every ntuplet (triplets,
// quadruplets, and pentuplets fitted on their first four hits), restated from
// plugin-PixelTriplets/BrokenLine.h:46-490 and BrokenLineFitOnGPU.h:24-182 in
// double precision.  The reference evaluates these expressions with Eigen
// (not available here); products and sums are taken here in increasing index
// order, and the symmetric positive-definite systems are inverted through an
// LDL^T (square-root-free Cholesky) factorisation instead of the reference's
// unrolled formulas, so the fitted values agree with the reference only to
// rounding (parity unpinned; pinned by invariants in the tests).
#include <cmath>
#include <cstring>

#include "oracle.h"

namespace orc {

  namespace {
    inline double sq(double a) { return a * a; }

    // inverse of a symmetric positive-definite n x n matrix (row-major, n <= 6)
    void invertSPD(const double* a, double* inv, int n) {
      double L[36] = {0}, D[6];
      for (int j = 0; j < n; ++j) {
        double d = a[j * n + j];
        for (int k = 0; k < j; ++k)
          d -= L[j * n + k] * L[j * n + k] * D[k];
        D[j] = d;
        L[j * n + j] = 1.0;
        for (int i = j + 1; i < n; ++i) {
          double v = a[i * n + j];
          for (int k = 0; k < j; ++k)
            v -= L[i * n + k] * L[j * n + k] * D[k];
          L[i * n + j] = v / d;
        }
      }
      // inverse of unit lower triangular L
      double Li[36] = {0};
      for (int j = 0; j < n; ++j) {
        Li[j * n + j] = 1.0;
        for (int i = j + 1; i < n; ++i) {
          double v = 0;
          for (int k = j; k < i; ++k)
            v -= L[i * n + k] * Li[k * n + j];
          Li[i * n + j] = v;
        }
      }
      // A^-1 = L^-T D^-1 L^-1
      for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
          double v = 0;
          for (int k = i; k < n; ++k)
            v += Li[k * n + i] * Li[k * n + j] / D[k];
          inv[i * n + j] = inv[j * n + i] = v;
        }
    }

    // BrokenLine.h:46-59
    double multScatt(double length, double B, double R, double slope) {
      double pt2 = std::min(20., B * R);
      pt2 *= pt2;
      const double XXI_0 = g_mutation == kMutXXI0 ? 0.12 / 16. : 0.06 / 16.;
      constexpr double fact = 0.7 * (13.6 / 1000.) * (13.6 / 1000.);
      return fact / (pt2 * (1. + sq(slope))) * (std::abs(length) * XXI_0) *
             sq(1. + 0.038 * std::log(std::abs(length) * XXI_0));
    }

    struct Rot2 {
      double r00, r01, r10, r11;
    };
    // BrokenLine.h:68-75
    Rot2 rotation(double slope) {
      Rot2 r;
      r.r00 = 1. / std::sqrt(1. + sq(slope));
      r.r01 = slope * r.r00;
      r.r10 = -r.r01;
      r.r11 = r.r00;
      return r;
    }

    inline double cross2(double ax, double ay, double bx, double by) { return ax * by - ay * bx; }

    // c = a * b (3x3), c = a * b^T
    void mul3(const double* a, const double* b, double* c) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          c[i * 3 + j] = a[i * 3 + 0] * b[0 * 3 + j] + a[i * 3 + 1] * b[1 * 3 + j] + a[i * 3 + 2] * b[2 * 3 + j];
    }
    void mul3T(const double* a, const double* b, double* c) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          c[i * 3 + j] = a[i * 3 + 0] * b[j * 3 + 0] + a[i * 3 + 1] * b[j * 3 + 1] + a[i * 3 + 2] * b[j * 3 + 2];
    }
    // cov <- J cov J^T
    void sandwich3(const double* J, double* cov) {
      double t[9], r[9];
      mul3(J, cov, t);
      mul3T(t, J, r);
      std::memcpy(cov, r, sizeof(r));
    }
    void sandwich2(const double* J, double* cov) {
      double t[4], r[4];
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          t[i * 2 + j] = J[i * 2 + 0] * cov[0 * 2 + j] + J[i * 2 + 1] * cov[1 * 2 + j];
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          r[i * 2 + j] = t[i * 2 + 0] * J[j * 2 + 0] + t[i * 2 + 1] * J[j * 2 + 1];
      std::memcpy(cov, r, sizeof(r));
    }

    struct CircleFit {
      double par[3];
      double cov[9];
      int q;
      float chi2;
    };
    // intermediate values of the circle fit, for the tests that pin its covariance:
    // the corrected first two hits d and e, (phi, d, k) and their covariance with the chord
    // midpoint as the origin, the scattering variance added at the first hit, the first
    // two hits' path lengths S0, S1, and the fast fit's radius and slope
    struct CircleDebug {
      double d[2], e[2], parMid[3], covMid[9], ms, S0, S1, R, slope;
    };
    struct LineFit {
      double par[2];
      double cov[4];
      double chi2;
    };

    // BrokenLine.h:85-108
    void translateKarimaki(CircleFit& c, double x0, double y0) {
      const double p0 = c.par[0], p1 = c.par[1], p2 = c.par[2];
      double DP = x0 * std::cos(p0) + y0 * std::sin(p0);
      double DO = x0 * std::sin(p0) - y0 * std::cos(p0) + p1;
      double uu = 1 + p2 * p1;
      double C = -p2 * y0 + uu * std::cos(p0);
      double BB = p2 * x0 + uu * std::sin(p0);
      double A = 2. * DO + p2 * (sq(DO) + sq(DP));
      double U = std::sqrt(1. + p2 * A);
      double xi = 1. / (sq(BB) + sq(C));
      double v = 1. + p2 * DO;
      double lambda = (0.5 * A) / (U * sq(1. + U));
      double mu = 1. / (U * (1. + U)) + p2 * lambda;
      double zeta = sq(DO) + sq(DP);
      const double J[9] = {xi * uu * v, -xi * sq(p2) * DP, xi * DP, 2. * mu * uu * DP, 2. * mu * v,
                           mu * zeta - lambda * A, 0, 0, 1.};
      c.par[0] = std::atan2(BB, C);
      c.par[1] = A / (1 + U);
      sandwich3(J, c.cov);
    }

    // BrokenLine.h:173-203: n x n band matrix, row-major
    void matrixCu(int n, const double* w, const double* S, const double* VB, double* out) {
      double C[25] = {0};
      for (int i = 0; i < n; ++i) {
        double cii = w[i];
        if (i > 1)
          cii += 1. / (VB[i - 1] * sq(S[i] - S[i - 1]));
        if (i > 0 && i < n - 1)
          cii += (1. / VB[i]) * sq((S[i + 1] - S[i - 1]) / ((S[i + 1] - S[i]) * (S[i] - S[i - 1])));
        if (i < n - 2)
          cii += 1. / (VB[i + 1] * sq(S[i + 1] - S[i]));
        C[i * n + i] = cii;
        if (i > 0 && i < n - 1)
          C[i * n + i + 1] =
              1. / (VB[i] * (S[i + 1] - S[i])) * (-(S[i + 1] - S[i - 1]) / ((S[i + 1] - S[i]) * (S[i] - S[i - 1])));
        if (i < n - 2)
          C[i * n + i + 1] += 1. / (VB[i + 1] * (S[i + 1] - S[i])) *
                              (-(S[i + 2] - S[i]) / ((S[i + 2] - S[i + 1]) * (S[i + 1] - S[i])));
        if (i < n - 2)
          C[i * n + i + 2] = 1. / (VB[i + 1] * (S[i + 2] - S[i + 1]) * (S[i + 1] - S[i]));
        C[i * n + i] *= 0.5;
      }
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          out[i * n + j] = C[i * n + j] + C[j * n + i];
    }

    template <int N>
    struct BLData {
      int q;
      double radii[2][N];
      double s[N], S[N], Z[N], VarBeta[N];
    };

    // BrokenLine.h:216-237; hits[3][N] (x, y, z rows)
    template <int N>
    void fastFit(const double (&h)[3][N], double* r) {
      constexpr int n = N;
      const double ax = h[0][n / 2] - h[0][0], ay = h[1][n / 2] - h[1][0];
      const double bx = h[0][n - 1] - h[0][n / 2], by = h[1][n - 1] - h[1][n / 2];
      const double cx = h[0][0] - h[0][n - 1], cy = h[1][0] - h[1][n - 1];
      const double a2 = ax * ax + ay * ay, b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
      double tmp = 0.5 / cross2(cx, cy, ax, ay);
      r[0] = h[0][0] - (ay * c2 + cy * a2) * tmp;
      r[1] = h[1][0] + (ax * c2 + cx * a2) * tmp;
      r[2] = std::sqrt(a2 * b2 * c2) / (2. * std::abs(cross2(bx, by, ax, ay)));
      const double dx = h[0][0] - r[0], dy = h[1][0] - r[1];
      const double ex = h[0][n - 1] - r[0], ey = h[1][n - 1] - r[1];
      r[3] = r[2] * std::atan2(cross2(dx, dy, ex, ey), dx * ex + dy * ey) / (h[2][n - 1] - h[2][0]);
    }

    // BrokenLine.h:119-162
    template <int N>
    void prepare(const double (&h)[3][N], const double* ff, double B, BLData<N>& d) {
      constexpr int n = N;
      double dx = h[0][1] - h[0][0], dy = h[1][1] - h[1][0];
      double ex = h[0][n - 1] - h[0][n - 2], ey = h[1][n - 1] - h[1][n - 2];
      d.q = cross2(dx, dy, ex, ey) > 0 ? -1 : 1;
      const double slope = -d.q / ff[3];
      const Rot2 R = rotation(slope);
      for (int i = 0; i < n; ++i) {
        d.radii[0][i] = h[0][i] - ff[0];
        d.radii[1][i] = h[1][i] - ff[1];
      }
      const double nrm = std::sqrt(ff[0] * ff[0] + ff[1] * ff[1]);
      ex = -ff[2] * ff[0] / nrm;
      ey = -ff[2] * ff[1] / nrm;
      for (int i = 0; i < n; ++i) {
        dx = d.radii[0][i];
        dy = d.radii[1][i];
        d.s[i] = d.q * ff[2] * std::atan2(cross2(dx, dy, ex, ey), dx * ex + dy * ey);
      }
      for (int i = 0; i < n; ++i) {
        const double si = d.s[i], zi = h[2][i];
        d.S[i] = R.r00 * si + R.r01 * zi;
        d.Z[i] = R.r10 * si + R.r11 * zi;
      }
      d.VarBeta[0] = d.VarBeta[n - 1] = 0;
      for (int i = 1; i < n - 1; ++i)
        d.VarBeta[i] = multScatt(d.S[i + 1] - d.S[i], B, ff[2], slope) +
                       (g_mutation == kMutVarBetaOneSide ? 0. : multScatt(d.S[i] - d.S[i - 1], B, ff[2], slope));
    }

    // BrokenLine.h:401-490; ge[6][N] float (xx, xy, yy, xz, yz, zz)
    template <int N>
    void lineFit(const float (&ge)[6][N], const double* ff, double B, const BLData<N>& d, LineFit& L) {
      constexpr int n = N;
      const double slope = -d.q / ff[3];
      const Rot2 R = rotation(slope);
      double w[N], ru[N];
      for (int i = 0; i < n; ++i) {
        const double V[3][3] = {{ge[0][i], ge[1][i], ge[3][i]}, {ge[1][i], ge[2][i], ge[4][i]},
                                {ge[3][i], ge[4][i], ge[5][i]}};
        const double tmp = 1. / std::sqrt(d.radii[0][i] * d.radii[0][i] + d.radii[1][i] * d.radii[1][i]);
        const double J[2][3] = {{d.radii[1][i] * tmp, -d.radii[0][i] * tmp, 0}, {0, 0, 1.}};
        // (R J) V J^T R^T, element (1,1)
        double RJ[2][3];
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 3; ++b)
            RJ[a][b] = (a == 0 ? R.r00 : R.r10) * J[0][b] + (a == 0 ? R.r01 : R.r11) * J[1][b];
        double RJV[3];
        for (int b = 0; b < 3; ++b)
          RJV[b] = RJ[1][0] * V[0][b] + RJ[1][1] * V[1][b] + RJ[1][2] * V[2][b];
        double RJVJt[2];
        for (int b = 0; b < 2; ++b)
          RJVJt[b] = RJV[0] * J[b][0] + RJV[1] * J[b][1] + RJV[2] * J[b][2];
        const double m11 = RJVJt[0] * R.r10 + RJVJt[1] * R.r11;
        w[i] = g_mutation == kMutLineWeight ? 1. / V[2][2] : 1. / m11;
      }
      for (int i = 0; i < n; ++i)
        ru[i] = w[i] * d.Z[i];
      double Cu[N * N], I[N * N];
      matrixCu(n, w, d.S, d.VarBeta, Cu);
      invertSPD(Cu, I, n);
      double u[N];
      for (int i = 0; i < n; ++i) {
        double v = 0;
        for (int k = 0; k < n; ++k)
          v += I[i * n + k] * ru[k];
        u[i] = v;
      }
      L.par[0] = (u[1] - u[0]) / (d.S[1] - d.S[0]);
      L.par[1] = u[0];
      const double idiff = 1. / (d.S[1] - d.S[0]);
      L.cov[0] = (I[0] - 2 * I[1] + I[n + 1]) * sq(idiff) +
                 (g_mutation == kMutNoFirstMSLine ? 0. : multScatt(d.S[1] - d.S[0], B, ff[2], slope));
      L.cov[1] = L.cov[2] = (I[1] - I[0]) * idiff;
      L.cov[3] = I[0];
      double J[4] = {1., 0, -d.S[0], 1.};
      L.par[1] += -L.par[0] * d.S[0];
      sandwich2(J, L.cov);
      const double tmp = R.r00 - L.par[0] * R.r01;
      J[3] = 1. / tmp;
      J[0] = J[3] * J[3];
      J[1] = 0;
      J[2] = L.par[1] * R.r01 * J[0];
      L.par[1] = L.par[1] * J[3];
      L.par[0] = (R.r01 + L.par[0] * R.r00) * J[3];
      sandwich2(J, L.cov);
      L.chi2 = 0;
      for (int i = 0; i < n; ++i) {
        L.chi2 += w[i] * sq(d.Z[i] - u[i]);
        if (i > 0 && i < n - 1)
          L.chi2 += sq(u[i - 1] / (d.S[i] - d.S[i - 1]) -
                       u[i] * (d.S[i + 1] - d.S[i - 1]) / ((d.S[i + 1] - d.S[i]) * (d.S[i] - d.S[i - 1])) +
                       u[i + 1] / (d.S[i + 1] - d.S[i])) /
                    d.VarBeta[i];
      }
    }

    // BrokenLine.h:256-382
    template <int N>
    void circleFit(const double (&h)[3][N], const float (&ge)[6][N], const double* ff, double B, BLData<N>& d,
                   CircleFit& c, CircleDebug* dbg = nullptr) {
      constexpr int n = N;
      c.q = d.q;
      const double slope = -c.q / ff[3];
      for (int i = 0; i < n; ++i)
        d.VarBeta[i] *= g_mutation == kMutNoSlopeFactor ? 1. : 1. + sq(slope);
      for (int i = 0; i < n; ++i)
        d.Z[i] = std::sqrt(d.radii[0][i] * d.radii[0][i] + d.radii[1][i] * d.radii[1][i]) - ff[2];
      double w[N];
      for (int i = 0; i < n; ++i) {
        const double V00 = ge[0][i], V01 = ge[1][i], V11 = ge[2][i];
        const Rot2 RR = rotation(-d.radii[0][i] / d.radii[1][i]);
        // (RR V RR^T)(1,1)
        const double a0 = RR.r10 * V00 + RR.r11 * V01;
        const double a1 = RR.r10 * V01 + RR.r11 * V11;
        w[i] = 1. / (a0 * RR.r10 + a1 * RR.r11);
      }
      constexpr int m = N + 1;
      double ru[m];
      ru[n] = 0;
      for (int i = 0; i < n; ++i)
        ru[i] = w[i] * d.Z[i];
      double Cn[N * N], C[m * m] = {0}, I[m * m];
      matrixCu(n, w, d.s, d.VarBeta, Cn);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          C[i * m + j] = Cn[i * n + j];
      const double* s = d.s;
      const double* VB = d.VarBeta;
      C[n * m + n] = 0;
      for (int i = 0; i < n; ++i) {
        double cin = 0;
        if (i > 0 && i < n - 1)
          cin += -(s[i + 1] - s[i - 1]) * (s[i + 1] - s[i - 1]) / (2. * VB[i] * (s[i + 1] - s[i]) * (s[i] - s[i - 1]));
        if (i > 1 && g_mutation != kMutCuBorder)
          cin += (s[i] - s[i - 2]) / (2. * VB[i - 1] * (s[i] - s[i - 1]));
        if (i < n - 2)
          cin += (s[i + 2] - s[i]) / (2. * VB[i + 1] * (s[i + 1] - s[i]));
        C[i * m + n] = C[n * m + i] = cin;
        if (i > 0 && i < n - 1)
          C[n * m + n] += sq(s[i + 1] - s[i - 1]) / (4. * VB[i]);
      }
      invertSPD(C, I, m);
      double u[m];
      for (int i = 0; i < m; ++i) {
        double v = 0;
        for (int k = 0; k < m; ++k)
          v += I[i * m + k] * ru[k];
        u[i] = v;
      }
      for (int k = 0; k < 2; ++k) {
        const double nr = std::sqrt(d.radii[0][k] * d.radii[0][k] + d.radii[1][k] * d.radii[1][k]);
        d.radii[0][k] /= nr;
        d.radii[1][k] /= nr;
      }
      const double dx = h[0][0] + (-d.Z[0] + u[0]) * d.radii[0][0];
      const double dy = h[1][0] + (-d.Z[0] + u[0]) * d.radii[1][0];
      const double ex = h[0][1] + (-d.Z[1] + u[1]) * d.radii[0][1];
      const double ey = h[1][1] + (-d.Z[1] + u[1]) * d.radii[1][1];
      const double emdx = ex - dx, emdy = ey - dy;
      const double tmp1 = emdx * emdx + emdy * emdy;
      c.par[0] = std::atan2(emdy, emdx);
      c.par[1] = -c.q * (ff[2] - std::sqrt(sq(ff[2]) - 0.25 * tmp1));
      c.par[2] = c.q * (1. / ff[2] + u[n]);
      // integer division, as in the reference (BrokenLine.h:351,353): always 0, so the
      // sagitta's variance at the chord midpoint drops out of the covariance; the
      // kVarHalfChargeReal switch evaluates q / 2 in real arithmetic for the tests
      const double qHalf = (g_variant & kVarHalfChargeReal) ? 0.5 * c.q : double(c.q / 2);
      // d(phi)/d(u0) as the reference writes it (BrokenLine.h:348); the derivative of
      // atan2(e - d) has the opposite sign, which the kVarPhiJacobianSign variant applies
      const double dPhi0 = (g_mutation == kMutPhiJacobian ? d.radii[0][0] * emdx - emdy * d.radii[1][0]
                                                          : d.radii[1][0] * emdx - emdy * d.radii[0][0]) / tmp1;
      const double J[9] = {(g_variant & kVarPhiJacobianSign) ? -dPhi0 : dPhi0,
                           (d.radii[1][1] * emdx - emdy * d.radii[0][1]) / tmp1,
                           0,
                           qHalf * (emdx * d.radii[0][0] + emdy * d.radii[1][0]) / std::sqrt(sq(2 * ff[2]) - tmp1),
                           qHalf * (emdx * d.radii[0][1] + emdy * d.radii[1][1]) / std::sqrt(sq(2 * ff[2]) - tmp1),
                           0,
                           0,
                           0,
                           double(c.q)};
      c.cov[0] = I[0 * m + 0];
      c.cov[1] = I[0 * m + 1];
      c.cov[2] = I[0 * m + n];
      c.cov[3] = I[1 * m + 0];
      c.cov[4] = I[1 * m + 1];
      c.cov[5] = I[1 * m + n];
      c.cov[6] = I[n * m + 0];
      c.cov[7] = I[n * m + 1];
      c.cov[8] = I[n * m + n];
      sandwich3(J, c.cov);
      if (dbg) {
        *dbg = {{dx, dy}, {ex, ey}, {c.par[0], c.par[1], c.par[2]}, {}, 0, d.S[0], d.S[1], ff[2], slope};
        std::memcpy(dbg->covMid, c.cov, sizeof(c.cov));
      }
      translateKarimaki(c, 0.5 * emdx, 0.5 * emdy);
      const double msFirst = (1 + sq(slope)) * multScatt(d.S[1] - d.S[0], B, ff[2], slope);
      if (dbg)
        dbg->ms = msFirst;
      if (g_mutation != kMutNoFirstMSCircle)
        c.cov[0] += msFirst;
      translateKarimaki(c, dx, dy);
      c.chi2 = 0;
      for (int i = 0; i < n; ++i) {
        c.chi2 += w[i] * sq(d.Z[i] - u[i]);
        if (i > 0 && i < n - 1)
          c.chi2 += sq(u[i - 1] / (s[i] - s[i - 1]) - u[i] * (s[i + 1] - s[i - 1]) / ((s[i + 1] - s[i]) * (s[i] - s[i - 1])) +
                       u[i + 1] / (s[i + 1] - s[i]) + (s[i + 1] - s[i - 1]) * u[n] / 2) /
                    VB[i];
      }
    }

    template <int N>
    void fitOne(Event& ev, const Conditions& c, uint32_t tkid, double B) {
      auto& C = ev.ca;
      double h[3][N];
      float ge[6][N];
      const uint32_t first = C.tupleOff[tkid];
      for (int i = 0; i < N; ++i) {
        const uint16_t hit = C.tupleBins[first + i];
        const auto& r = c.cpeDet[ev.detInd[hit]].frame;
        // SOAFrame::toGlobal(cxx, cxy, cyy, gl) (DataFormats/SOARotation.h:103-111)
        const float cxx = ev.xerr[hit], cxy = 0, cyy = ev.yerr[hit];
        ge[0][i] = r.R11 * (r.R11 * cxx + r.R21 * cxy) + r.R21 * (r.R11 * cxy + r.R21 * cyy);
        ge[1][i] = r.R11 * (r.R12 * cxx + r.R22 * cxy) + r.R21 * (r.R12 * cxy + r.R22 * cyy);
        ge[2][i] = r.R12 * (r.R12 * cxx + r.R22 * cxy) + r.R22 * (r.R12 * cxy + r.R22 * cyy);
        ge[3][i] = r.R11 * (r.R13 * cxx + r.R23 * cxy) + r.R21 * (r.R13 * cxy + r.R23 * cyy);
        ge[4][i] = r.R12 * (r.R13 * cxx + r.R23 * cxy) + r.R22 * (r.R13 * cxy + r.R23 * cyy);
        ge[5][i] = r.R13 * (r.R13 * cxx + r.R23 * cxy) + r.R23 * (r.R13 * cxy + r.R23 * cyy);
        h[0][i] = ev.xg[hit];
        h[1][i] = ev.yg[hit];
        h[2][i] = ev.zg[hit];
      }
      double ff[4];
      fastFit<N>(h, ff);
      BLData<N> d;
      prepare<N>(h, ff, B, d);
      LineFit L;
      lineFit<N>(ge, ff, B, d, L);
      CircleFit Ci;
      circleFit<N>(h, ge, ff, B, d, Ci);
      // TrajectoryStateSoA::copyFromCircle (CUDADataFormats/TrajectoryStateSoA.h:20-38)
      const float b = 1.f / float(B);
      float* st = C.state[tkid].data();
      st[0] = float(Ci.par[0]);
      st[1] = float(Ci.par[1]);
      st[2] = float(Ci.par[2]);
      st[3] = float(L.par[0]);
      st[4] = float(L.par[1]);
      st[2] *= b;
      float* cv = C.cov[tkid].data();
      cv[0] = float(Ci.cov[0]);
      cv[1] = float(Ci.cov[1]);
      cv[2] = b * float(Ci.cov[2]);
      cv[4] = cv[3] = 0;
      cv[5] = float(Ci.cov[4]);
      cv[6] = b * float(Ci.cov[5]);
      cv[8] = cv[7] = 0;
      cv[9] = b * b * float(Ci.cov[8]);
      cv[11] = cv[10] = 0;
      cv[12] = float(L.cov[0]);
      cv[13] = float(L.cov[1]);
      cv[14] = float(L.cov[3]);
      C.pt[tkid] = float(B) / float(std::abs(Ci.par[2]));
      C.eta[tkid] = asinhf(float(L.par[0]));
      C.chi2[tkid] = float((Ci.chi2 + L.chi2) / (2 * N - (g_mutation == kMutChi2Ndf ? 4 : 5)));
    }
  }  // namespace

  // a18/a19: HelixFitOnGPU::launchBrokenLineKernelsOnCPU (BrokenLineFitOnGPU.cc:3-68), fit5as4
  void fitTracks(const Conditions& c, Event& ev) {
    auto& C = ev.ca;
    C.state.assign(ev.caps.maxTuples, {0, 0, 0, 0, 0});
    C.cov.assign(ev.caps.maxTuples, {});
    C.pt.assign(ev.caps.maxTuples, 0);
    C.eta.assign(ev.caps.maxTuples, 0);
    C.chi2.assign(ev.caps.maxTuples, 0);
    if (ev.nHits == 0)
      return;
    const double B = double(c.bField);
    for (uint32_t nh = 3; nh <= 5; ++nh) {
      for (uint32_t k = C.multOff[nh]; k < C.multOff[nh + 1]; ++k) {
        const uint32_t tkid = C.multBins[k];
        if (nh == 3)
          fitOne<3>(ev, c, tkid, B);
        else
          fitOne<4>(ev, c, tkid, B);
      }
    }
  }

  // exposed for the fixture tests of the reference's testBrokenLineFit inputs
  int fitHelix(int n, const double* hits, const float* ge, double B, float* state, float* cov, float* pt, float* eta,
               float* chi2) {
    Event ev;
    auto& C = ev.ca;
    C.state.assign(1, {});
    C.cov.assign(1, {});
    C.pt.assign(1, 0);
    C.eta.assign(1, 0);
    C.chi2.assign(1, 0);
    // direct entry: hits[3][n], ge[6][n]
    auto run = [&](auto tag) {
      constexpr int N = decltype(tag)::value;
      double h[3][N];
      float g[6][N];
      for (int r = 0; r < 3; ++r)
        for (int i = 0; i < N; ++i)
          h[r][i] = hits[r * n + i];
      for (int r = 0; r < 6; ++r)
        for (int i = 0; i < N; ++i)
          g[r][i] = ge[r * n + i];
      double ff[4];
      fastFit<N>(h, ff);
      BLData<N> d;
      prepare<N>(h, ff, B, d);
      LineFit L;
      lineFit<N>(g, ff, B, d, L);
      CircleFit Ci;
      circleFit<N>(h, g, ff, B, d, Ci);
      const float b = 1.f / float(B);
      state[0] = float(Ci.par[0]);
      state[1] = float(Ci.par[1]);
      state[2] = float(Ci.par[2]) * b;
      state[3] = float(L.par[0]);
      state[4] = float(L.par[1]);
      cov[0] = float(Ci.cov[0]);
      cov[1] = float(Ci.cov[1]);
      cov[2] = b * float(Ci.cov[2]);
      cov[3] = cov[4] = 0;
      cov[5] = float(Ci.cov[4]);
      cov[6] = b * float(Ci.cov[5]);
      cov[7] = cov[8] = 0;
      cov[9] = b * b * float(Ci.cov[8]);
      cov[10] = cov[11] = 0;
      cov[12] = float(L.cov[0]);
      cov[13] = float(L.cov[1]);
      cov[14] = float(L.cov[3]);
      *pt = float(B) / float(std::abs(Ci.par[2]));
      *eta = asinhf(float(L.par[0]));
      *chi2 = float((Ci.chi2 + L.chi2) / (2 * N - (g_mutation == kMutChi2Ndf ? 4 : 5)));
    };
    if (n == 3)
      run(std::integral_constant<int, 3>{});
    else if (n == 4)
      run(std::integral_constant<int, 4>{});
    else if (n == 5)
      run(std::integral_constant<int, 5>{});
    else
      return -1;
    return 0;
  }

  // the fit's double-precision results before the float conversion: circle (phi, d, k) and
  // its covariance at the origin, line (cot theta, zip) and its covariance, the two chi2
  int fitHelixD(int n, const double* hits, const float* ge, double B, double* circle, double* ccov, double* line,
                double* lcov, double* chi2c, double* chi2l) {
    auto run = [&](auto tag) {
      constexpr int N = decltype(tag)::value;
      double h[3][N];
      float g[6][N];
      for (int r = 0; r < 3; ++r)
        for (int i = 0; i < N; ++i)
          h[r][i] = hits[r * n + i];
      for (int r = 0; r < 6; ++r)
        for (int i = 0; i < N; ++i)
          g[r][i] = ge[r * n + i];
      double ff[4];
      fastFit<N>(h, ff);
      BLData<N> d;
      prepare<N>(h, ff, B, d);
      LineFit L;
      lineFit<N>(g, ff, B, d, L);
      CircleFit Ci;
      circleFit<N>(h, g, ff, B, d, Ci);
      for (int k = 0; k < 3; ++k)
        circle[k] = Ci.par[k];
      for (int k = 0; k < 9; ++k)
        ccov[k] = Ci.cov[k];
      line[0] = L.par[0];
      line[1] = L.par[1];
      for (int k = 0; k < 4; ++k)
        lcov[k] = L.cov[k];
      *chi2c = Ci.chi2;
      *chi2l = L.chi2;
    };
    if (n == 3)
      run(std::integral_constant<int, 3>{});
    else if (n == 4)
      run(std::integral_constant<int, 4>{});
    else if (n == 5)
      run(std::integral_constant<int, 5>{});
    else
      return -1;
    return 0;
  }

  double multScattForTest(double length, double B, double R, double slope) { return multScatt(length, B, R, slope); }

  int fastFitForTest(int n, const double* hits, double* ff) {
    auto run = [&](auto tag) {
      constexpr int N = decltype(tag)::value;
      double h[3][N];
      for (int r = 0; r < 3; ++r)
        for (int i = 0; i < N; ++i)
          h[r][i] = hits[r * n + i];
      fastFit<N>(h, ff);
    };
    if (n == 3)
      run(std::integral_constant<int, 3>{});
    else if (n == 4)
      run(std::integral_constant<int, 4>{});
    else
      return -1;
    return 0;
  }

  // the circle fit's intermediate values (CircleDebug, 21 doubles in field order); with ffIn
  // the fit linearises about that fast fit (X0, Y0, R, tan theta) instead of the hits' own
  int circleDebug(int n, const double* hits, const float* ge, double B, const double* ffIn, double* out) {
    auto run = [&](auto tag) {
      constexpr int N = decltype(tag)::value;
      double h[3][N];
      float g[6][N];
      for (int r = 0; r < 3; ++r)
        for (int i = 0; i < N; ++i)
          h[r][i] = hits[r * n + i];
      for (int r = 0; r < 6; ++r)
        for (int i = 0; i < N; ++i)
          g[r][i] = ge[r * n + i];
      double ff[4];
      fastFit<N>(h, ff);
      if (ffIn)
        std::memcpy(ff, ffIn, sizeof(ff));
      BLData<N> d;
      prepare<N>(h, ff, B, d);
      CircleFit Ci;
      CircleDebug dbg;
      circleFit<N>(h, g, ff, B, d, Ci, &dbg);
      static_assert(sizeof(CircleDebug) == 21 * sizeof(double), "debug layout");
      std::memcpy(out, &dbg, sizeof(dbg));
    };
    if (n == 3)
      run(std::integral_constant<int, 3>{});
    else if (n == 4)
      run(std::integral_constant<int, 4>{});
    else
      return -1;
    return 0;
  }

  // TranslateKarimaki on (phi, d, k) and its covariance (row-major 3x3)
  void translateKarimakiForTest(double* par, double* cov, double x0, double y0) {
    CircleFit c{};
    std::memcpy(c.par, par, sizeof(c.par));
    std::memcpy(c.cov, cov, sizeof(c.cov));
    translateKarimaki(c, x0, y0);
    std::memcpy(par, c.par, sizeof(c.par));
    std::memcpy(cov, c.cov, sizeof(c.cov));
  }

  void invertSPDForTest(const double* a, double* inv, int n) {
    if (n >= 1 && n <= 6)  // the fits invert at most (4+2) x (4+2)
      invertSPD(a, inv, n);
  }

}  // namespace orc

// ORACLE -- test infrastructure only (see oracle.h).
// Stage C, alternative fit: the Riemann helix fit of every ntuplet (triplets,
// quadruplets, and pentuplets fitted on their first four hits), restated from
// plugin-PixelTriplets/RiemannFit.h:30-885, FitUtils.h:102-190 and the three
// kernels of RiemannFitOnGPU.h:20-187 (launch order RiemannFitOnGPU.cc:3-113)
// in double precision.
//
// The reference evaluates these expressions with Eigen 3 (not available
// here): matrix products and sums are taken here in increasing index order,
// and the small symmetric positive-definite matrices are inverted through an
// LDL^T factorisation instead of choleskyInversion.h's unrolled formulas.
// The minimum-eigenvalue eigenvector of the 3x3 matrices uses the published
// closed-form algorithm of Eigen's SelfAdjointEigenSolver::computeDirect
// (Eigen 3.3/3.4, Eigen/src/Eigenvalues/SelfAdjointEigenSolver.h:
// direct_selfadjoint_eigenvalues<..., 3, false>), restated in `eigen3Direct`,
// in double for the fit and in float for the numerical Jacobian
// (min_eigen3D_fast).  The elementary functions (atan2, atan, sin, cos, log) are
// the fit's own, built from IEEE basic operations (namespace rm, a few units in
// the last place; in float evaluated in double and rounded), so that the device
// fit computes the same bits.  The fitted values agree with the reference only
// to rounding (parity unpinned; pinned by invariants and closed forms in the
// tests).
#include <cmath>
#include <cstring>
#include <limits>

#include "oracle.h"

namespace orc {

  namespace {
    inline double sq(double a) { return a * a; }

    // ---- elementary functions from IEEE basic operations only -----------------------
    // Written for the Riemann fit so that host and device compute the same bits: its
    // numerical Jacobian (float eigenvectors of a perturbed matrix, RiemannFit.h:732-739)
    // and its smallest eigenvalue (a difference of O(1) terms) turn last-bit differences
    // of library functions into visible ones.  Accuracy: a few units in the last place.
    namespace rm {
      constexpr double kPi = 3.141592653589793, kPio2 = 1.5707963267948966, kPio6 = 0.5235987755982989;
      constexpr double kSqrt3 = 1.7320508075688772, kTanPi12 = 0.2679491924311227;
      constexpr double kTwoOverPi = 0.6366197723675814;
      // pi/2 in three parts, the first two of 30 significant bits
      constexpr double kPio2a = 1.570796325802803, kPio2b = 9.920935791635221e-10, kPio2c = 5.170182981794105e-19;
      // ln 2 in two parts, the first of 32 significant bits
      constexpr double kLn2Hi = 0.6931471803691238, kLn2Lo = 1.9082149292705877e-10;

      // log x = k ln 2 + 2 atanh(f), x = 2^k m with m in [1/sqrt(2), sqrt(2)), f = (m-1)/(m+1):
      // the odd series of atanh to f^25 (|f| < 0.172)
      inline double log(double x) {
        if (!(x > 0) || x == HUGE_VAL)
          return x == 0 ? -HUGE_VAL : (x == HUGE_VAL ? x : NAN);
        int k;
        double m = std::frexp(x, &k);
        if (m < 0.7071067811865476) {
          m *= 2;
          --k;
        }
        const double f = (m - 1) / (m + 1), f2 = f * f;
        constexpr double c[13] = {1.0,        1.0 / 3,  1.0 / 5,  1.0 / 7,  1.0 / 9,  1.0 / 11, 1.0 / 13,
                                  1.0 / 15,   1.0 / 17, 1.0 / 19, 1.0 / 21, 1.0 / 23, 1.0 / 25};
        double s = c[12];
        for (int j = 11; j >= 0; --j)
          s = s * f2 + c[j];
        return k * kLn2Hi + (2 * f * s + k * kLn2Lo);
      }

      // atan: |x| > 1 through pi/2 - atan(1/x), then |x| > tan(pi/12) through
      // pi/6 + atan((x sqrt3 - 1) / (x + sqrt3)), then the odd series to t^31 (t^2 < 0.072)
      inline double atan(double x) {
        double a = std::fabs(x);
        const bool inv = a > 1;
        if (inv)
          a = 1 / a;
        const bool red = a > kTanPi12;
        const double t = red ? (a * kSqrt3 - 1) / (a + kSqrt3) : a;
        const double t2 = t * t;
        constexpr double c[16] = {1.0,       -1.0 / 3,  1.0 / 5,   -1.0 / 7,  1.0 / 9,   -1.0 / 11,
                                  1.0 / 13,  -1.0 / 15, 1.0 / 17,  -1.0 / 19, 1.0 / 21,  -1.0 / 23,
                                  1.0 / 25,  -1.0 / 27, 1.0 / 29,  -1.0 / 31};
        double s = c[15];
        for (int j = 14; j >= 0; --j)
          s = s * t2 + c[j];
        double r = t * s;
        if (red)
          r = kPio6 + r;
        if (inv)
          r = kPio2 - r;
        return x < 0 ? -r : r;
      }

      inline double atan2(double y, double x) {
        if (x > 0)
          return atan(y / x);
        if (x < 0) {
          const double r = atan(y / x);
          return std::signbit(y) ? r - kPi : r + kPi;
        }
        if (x == 0) {
          if (y > 0)
            return kPio2;
          if (y < 0)
            return -kPio2;
          return std::signbit(x) ? (std::signbit(y) ? -kPi : kPi) : y;
        }
        return x + y;  // NaN
      }

      // x = q pi/2 + r with |r| <= pi/4, then the Taylor series of sin and cos to r^21
      inline void sincos(double x, double& s, double& c) {
        const double q = std::rint(x * kTwoOverPi);
        const double r = ((x - q * kPio2a) - q * kPio2b) - q * kPio2c;
        const double r2 = r * r;
        constexpr double cs[11] = {1.0,
                                   -1.0 / 6,
                                   1.0 / 120,
                                   -1.0 / 5040,
                                   1.0 / 362880,
                                   -1.0 / 39916800,
                                   1.0 / 6227020800.0,
                                   -1.0 / 1307674368000.0,
                                   1.0 / 355687428096000.0,
                                   -1.0 / 121645100408832000.0,
                                   1.0 / 51090942171709440000.0};
        constexpr double cc[11] = {1.0,
                                   -1.0 / 2,
                                   1.0 / 24,
                                   -1.0 / 720,
                                   1.0 / 40320,
                                   -1.0 / 3628800,
                                   1.0 / 479001600,
                                   -1.0 / 87178291200.0,
                                   1.0 / 20922789888000.0,
                                   -1.0 / 6402373705728000.0,
                                   1.0 / 2432902008176640000.0};
        double ps = cs[10], pc = cc[10];
        for (int j = 9; j >= 0; --j) {
          ps = ps * r2 + cs[j];
          pc = pc * r2 + cc[j];
        }
        const double sr = r * ps, cr = pc;
        const long long qi = ((static_cast<long long>(q) % 4) + 4) % 4;
        s = qi == 0 ? sr : (qi == 1 ? cr : (qi == 2 ? -sr : -cr));
        c = qi == 0 ? cr : (qi == 1 ? -sr : (qi == 2 ? -cr : sr));
      }
      inline double sin(double x) {
        double s, c;
        sincos(x, s, c);
        return s;
      }
      inline double cos(double x) {
        double s, c;
        sincos(x, s, c);
        return c;
      }
    }  // namespace rm


    // inverse of a symmetric positive-definite n x n matrix (row-major), LDL^T
    template <int n>
    void invertSPDn(const double* a, double* inv) {
      double L[n * n] = {}, D[n], Li[n * n] = {};
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
      for (int j = 0; j < n; ++j) {
        Li[j * n + j] = 1.0;
        for (int i = j + 1; i < n; ++i) {
          double v = 0;
          for (int k = j; k < i; ++k)
            v -= L[i * n + k] * Li[k * n + j];
          Li[i * n + j] = v;
        }
      }
      for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
          double v = 0;
          for (int k = i; k < n; ++k)
            v += Li[k * n + i] * Li[k * n + j] / D[k];
          inv[i * n + j] = inv[j * n + i] = v;
        }
    }

    // ---- Eigen's closed-form 3x3 symmetric eigen solver ---------------------------
    // Angle functions: in double for T = double; for T = float evaluated in double and
    // rounded to float.
    inline double angAtan2(double y, double x) { return rm::atan2(y, x); }
    inline double angCos(double a) { return rm::cos(a); }
    inline double angSin(double a) { return rm::sin(a); }
    inline float angAtan2(float y, float x) { return float(rm::atan2(double(y), double(x))); }
    inline float angCos(float a) { return float(rm::cos(double(a))); }
    inline float angSin(float a) { return float(rm::sin(double(a))); }

    // roots of the characteristic polynomial, ascending (computeRoots)
    template <typename T>
    void eigenRoots(const T m[3][3], T r[3]) {
      const T inv3 = T(1) / T(3);
      const T sqrt3 = std::sqrt(T(3));
      const T c0 = m[0][0] * m[1][1] * m[2][2] + T(2) * m[1][0] * m[2][0] * m[2][1] - m[0][0] * m[2][1] * m[2][1] -
                   m[1][1] * m[2][0] * m[2][0] - m[2][2] * m[1][0] * m[1][0];
      const T c1 = m[0][0] * m[1][1] - m[1][0] * m[1][0] + m[0][0] * m[2][2] - m[2][0] * m[2][0] + m[1][1] * m[2][2] -
                   m[2][1] * m[2][1];
      const T c2 = m[0][0] + m[1][1] + m[2][2];
      const T c2over3 = c2 * inv3;
      T aOver3 = (c2 * c2over3 - c1) * inv3;
      if (aOver3 < T(0))
        aOver3 = T(0);
      const T halfB = T(0.5) * (c0 + c2over3 * (T(2) * c2over3 * c2over3 - c1));
      T q = aOver3 * aOver3 * aOver3 - halfB * halfB;
      if (q < T(0))
        q = T(0);
      const T rho = std::sqrt(aOver3);
      const T theta = angAtan2(std::sqrt(q), halfB) * inv3;
      const T cosT = angCos(theta), sinT = angSin(theta);
      r[0] = c2over3 - rho * (cosT + sqrt3 * sinT);
      r[1] = c2over3 - rho * (cosT - sqrt3 * sinT);
      r[2] = c2over3 + T(2) * rho * cosT;
    }

    template <typename T>
    void cross3(const T a[3], const T b[3], T c[3]) {
      c[0] = a[1] * b[2] - a[2] * b[1];
      c[1] = a[2] * b[0] - a[0] * b[2];
      c[2] = a[0] * b[1] - a[1] * b[0];
    }

    // kernel of a rank-2 matrix (extract_kernel): the larger of the two cross products
    // of the column with the largest diagonal entry with the other two columns
    template <typename T>
    void extractKernel(const T m[3][3], T res[3], T rep[3]) {
      int i0 = 0;
      T best = std::fabs(m[0][0]);
      for (int i = 1; i < 3; ++i)
        if (std::fabs(m[i][i]) > best) {
          best = std::fabs(m[i][i]);
          i0 = i;
        }
      for (int r = 0; r < 3; ++r)
        rep[r] = m[r][i0];
      T ca[3], cb[3], c1[3], c2[3];
      for (int r = 0; r < 3; ++r) {
        ca[r] = m[r][(i0 + 1) % 3];
        cb[r] = m[r][(i0 + 2) % 3];
      }
      cross3(rep, ca, c1);
      cross3(rep, cb, c2);
      const T n0 = c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2];
      const T n1 = c2[0] * c2[0] + c2[1] * c2[1] + c2[2] * c2[2];
      if (n0 > n1) {
        const T s = std::sqrt(n0);
        for (int r = 0; r < 3; ++r)
          res[r] = c1[r] / s;
      } else {
        const T s = std::sqrt(n1);
        for (int r = 0; r < 3; ++r)
          res[r] = c2[r] / s;
      }
    }

    // eigenvalues ascending in ev, eigenvectors as the columns of V (V[row][col]);
    // the lower triangle of a is read
    template <typename T>
    void eigen3Direct(const T a[3][3], T ev[3], T V[3][3]) {
      T m[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          m[i][j] = i >= j ? a[i][j] : a[j][i];
      const T shift = (m[0][0] + m[1][1] + m[2][2]) / T(3);
      for (int i = 0; i < 3; ++i)
        m[i][i] -= shift;
      T scale = T(0);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          scale = std::fabs(m[i][j]) > scale ? std::fabs(m[i][j]) : scale;
      if (scale > T(0))
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            m[i][j] /= scale;
      eigenRoots(m, ev);
      const T eps = std::numeric_limits<T>::epsilon();
      if (ev[2] - ev[0] <= eps) {
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            V[i][j] = i == j ? T(1) : T(0);
      } else {
        T tmp[3][3];
        std::memcpy(tmp, m, sizeof(tmp));
        T d0 = ev[2] - ev[1];
        const T d1 = ev[1] - ev[0];
        int k = 0, l = 2;
        if (d0 > d1) {
          k = 2;
          l = 0;
          d0 = d1;
        }
        T vk[3], vl[3];
        for (int i = 0; i < 3; ++i)
          tmp[i][i] -= ev[k];
        extractKernel(tmp, vk, vl);
        if (d0 <= T(2) * eps * d1) {
          // the two other eigenvalues coincide numerically: the saved column, made
          // unit length after the (literal) update of the restated algorithm
          const T dot = vk[0] * vl[0] + vk[1] * vl[1] + vk[2] * vl[2];
          for (int r = 0; r < 3; ++r)
            vl[r] -= dot * vl[r];
          const T n = vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2];
          if (n > T(0)) {
            const T s = std::sqrt(n);
            for (int r = 0; r < 3; ++r)
              vl[r] /= s;
          }
        } else {
          std::memcpy(tmp, m, sizeof(tmp));
          for (int i = 0; i < 3; ++i)
            tmp[i][i] -= ev[l];
          T dummy[3];
          extractKernel(tmp, vl, dummy);
        }
        for (int r = 0; r < 3; ++r) {
          V[r][k] = vk[r];
          V[r][l] = vl[r];
        }
        T c2[3] = {V[0][2], V[1][2], V[2][2]}, c0[3] = {V[0][0], V[1][0], V[2][0]}, c1[3];
        cross3(c2, c0, c1);
        const T n = c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2];
        const T s = n > T(0) ? std::sqrt(n) : T(1);
        for (int r = 0; r < 3; ++r)
          V[r][1] = c1[r] / s;
      }
      for (int i = 0; i < 3; ++i)
        ev[i] = ev[i] * scale + shift;
    }

    // RiemannFit.h:304-315 min_eigen3D: the eigenvector of the smallest eigenvalue
    // (first one on ties), the eigenvalue in chi2
    void minEigen3D(const double A[3][3], double& chi2, double v[3]) {
      double ev[3], V[3][3];
      eigen3Direct<double>(A, ev, V);
      int k = 0;
      for (int i = 1; i < 3; ++i)
        if (ev[i] < ev[k])
          k = i;
      chi2 = ev[k];
      for (int r = 0; r < 3; ++r)
        v[r] = V[r][k];
    }

    // RiemannFit.h:328-334 min_eigen3D_fast: the same in float
    void minEigen3DFast(const double A[3][3], double v[3]) {
      float a[3][3], ev[3], V[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          a[i][j] = float(A[i][j]);
      eigen3Direct<float>(a, ev, V);
      int k = 0;
      for (int i = 1; i < 3; ++i)
        if (ev[i] < ev[k])
          k = i;
      for (int r = 0; r < 3; ++r)
        v[r] = double(V[r][k]);
    }

    struct RCircle {
      double par[3];  // (X0, Y0, R), then (phi, Tip, q/R) after toPerigee
      double cov[3][3];
      int q;
      float chi2;  // FitResult.h:35
    };
    struct RLine {
      double par[2];  // (cot(theta), Zip)
      double cov[2][2];
      double chi2;
    };

    // RiemannFit.h:30-40
    template <int N>
    void radLenUniform(const double len[N], double rl[N]) {
      constexpr double xx0inv = 0.06 / 16.;
      rl[0] = len[0] * xx0inv;
      for (int j = 1; j < N; ++j)
        rl[j] = std::fabs(len[j] - len[j - 1]) * xx0inv;
    }

    // RiemannFit.h:365-417 fastFit: circle through the first, middle and last hit,
    // tan(theta) from the arc between the first and the last
    template <int N>
    void riemannFastFit(const double h[3][N], double r[4]) {
      constexpr int n = N;
      const double bx0 = h[0][n / 2] - h[0][0], by0 = h[1][n / 2] - h[1][0];
      const double cx0 = h[0][n - 1] - h[0][0], cy0 = h[1][n - 1] - h[1][0];
      const double b2 = bx0 * bx0 + by0 * by0;
      const double c2 = cx0 * cx0 + cy0 * cy0;
      const bool flip = std::fabs(bx0) < std::fabs(by0);
      const double bx = flip ? by0 : bx0, by = flip ? bx0 : by0;
      const double cx = flip ? cy0 : cx0, cy = flip ? cx0 : cy0;
      const double div = 2. * (cx * by - bx * cy);
      const double y0 = (cx * b2 - bx * c2) / div;
      const double x0 = (0.5 * b2 - y0 * by) / bx;
      r[0] = h[0][0] + (flip ? y0 : x0);
      r[1] = h[1][0] + (flip ? x0 : y0);
      r[2] = std::sqrt(sq(x0) + sq(y0));
      const double dx = h[0][0] - r[0], dy = h[1][0] - r[1];
      const double ex = h[0][n - 1] - r[0], ey = h[1][n - 1] - r[1];
      const double dr = r[2] * rm::atan2(dx * ey - dy * ex, dx * ex + dy * ey);
      const double dz = h[2][n - 1] - h[2][0];
      r[3] = dr / dz;
    }

    // RiemannFit.h:120-158 scatter_cov_rad
    template <int N>
    void scatterCovRad(const double h[3][N], const double ff[4], const double rad[N], double B, double S[N][N]) {
      const double pt = std::min(20., ff[2] * B);
      const double p2 = pt * pt * (1. + 1. / sq(ff[3]));
      double theta = rm::atan(ff[3]);
      theta = theta < 0. ? theta + M_PI : theta;
      double s[N], rl[N];
      for (int i = 0; i < N; ++i) {
        const double px = h[0][i] - ff[0], py = h[1][i] - ff[1];
        const double ox = -ff[0], oy = -ff[1];
        const double cross = ox * py - oy * px;
        const double dot = ox * px + oy * py;
        s[i] = std::fabs(rm::atan2(cross, dot) * ff[2]);
      }
      const double f = std::sqrt(1. + 1. / sq(ff[3]));
      double sf[N];
      for (int i = 0; i < N; ++i)
        sf[i] = s[i] * f;
      radLenUniform<N>(sf, rl);
      double sig2[N];
      const double k = 0.000225 / (p2 * sq(rm::sin(theta)));
      for (int i = 0; i < N; ++i)
        sig2[i] = sq(1. + 0.038 * rm::log(rl[i])) * rl[i] * k;
      for (int a = 0; a < N; ++a)
        for (int b = a; b < N; ++b) {
          double v = 0;
          for (int i = 0; i < std::min(a, b); ++i)
            v += (rad[a] - rad[i]) * (rad[b] - rad[i]) * sig2[i];
          S[a][b] = S[b][a] = v;
        }
    }

    // RiemannFit.h:60-110 scatterCovLine: the z' block of the s-z covariance with
    // multiple scattering (only the (1,1) entry of each hit's rotated covariance enters)
    template <int N>
    void scatterCovLine(const double covSz11[N], const double ff[4], const double sArc[N], const double z[N],
                        double B, double ret[N][N]) {
      const double pt = std::min(20., ff[2] * B);
      const double p2 = pt * pt * (1. + 1. / sq(ff[3]));
      double s[N], rl[N], sig2[N];
      for (int i = 0; i < N; ++i)
        s[i] = std::sqrt(sArc[i] * sArc[i] + z[i] * z[i]);
      radLenUniform<N>(s, rl);
      for (int i = 0; i < N; ++i)
        sig2[i] = .000225 / p2 * sq(1. + 0.038 * rm::log(rl[i])) * rl[i];
      for (int a = 0; a < N; ++a)
        for (int b = a; b < N; ++b) {
          double v = a == b ? covSz11[a] : 0.;
          for (int i = 0; i < std::min(a, b); ++i)
            v += std::fabs(s[a] - s[i]) * std::fabs(s[b] - s[i]) * sig2[i];
          ret[a][b] = ret[b][a] = v;
        }
    }

    // RiemannFit.h:442-783 circleFit with error propagation; V is the 2N x 2N hit
    // covariance in the (x..., y...) layout of loadCovariance2D (FitUtils.h:102-122)
    template <int N>
    void riemannCircleFit(const double h[3][N], double V[2 * N][2 * N], const double ff[4], const double rad[N],
                          double B, RCircle& out) {
      constexpr int n = N;
      // weights: prefit-orthogonal hit errors plus scattering, inverted and normalised
      double covRad[N][N] = {}, S[N][N], G[N][N];
      for (int i = 0; i < n; ++i) {  // RiemannFit.h:244-262 cov_carttorad_prefit
        if (rad[i] < 1.e-4) {
          covRad[i][i] = V[i][i];
        } else {
          const double ax = h[0][i], ay = h[1][i];
          const double bx = h[0][i] - ff[0], by = h[1][i] - ff[1];
          const double x2 = ax * bx + ay * by;
          const double y2 = ax * by - ay * bx;
          const double tanc = -y2 / x2;
          const double tanc2 = sq(tanc);
          covRad[i][i] = 1. / (1. + tanc2) * (V[i][i] + V[i + n][i + n] * tanc2 + 2 * V[i][i + n] * tanc);
        }
      }
      scatterCovRad<N>(h, ff, rad, B, S);
      {
        // RiemannFit.h:169-194 cov_radtocart, in its assignment order (on the diagonal
        // blocks' diagonal, (i, i+n) and (i+n, i) both end up with the (i, i+n) value)
        double CC[2 * N][2 * N];
        for (int i = 0; i < n; ++i)
          for (int j = i; j < n; ++j) {
            const double ri = 1. / rad[i], rj = 1. / rad[j];
            CC[i][j] = S[i][j] * h[1][i] * ri * h[1][j] * rj;
            CC[i + n][j + n] = S[i][j] * h[0][i] * ri * h[0][j] * rj;
            CC[i][j + n] = -S[i][j] * h[1][i] * ri * h[0][j] * rj;
            CC[i + n][j] = -S[i][j] * h[0][i] * ri * h[1][j] * rj;
            CC[j][i] = CC[i][j];
            CC[j + n][i + n] = CC[i + n][j + n];
            CC[j + n][i] = CC[i][j + n];
            CC[j][i + n] = CC[i + n][j];
          }
        for (int a = 0; a < 2 * n; ++a)
          for (int b = 0; b < 2 * n; ++b)
            V[a][b] += CC[a][b];
      }
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          covRad[i][j] += S[i][j];
      invertSPDn<N>(&covRad[0][0], &G[0][0]);
      double renorm = 0;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          renorm += G[i][j];
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          G[i][j] *= 1. / renorm;
      double w[N];
      for (int j = 0; j < n; ++j) {
        w[j] = 0;
        for (int i = 0; i < n; ++i)
          w[j] += G[i][j];
      }
      // centre, scale, project on the paraboloid
      double hc[2] = {0, 0};
      for (int i = 0; i < n; ++i) {
        hc[0] += h[0][i];
        hc[1] += h[1][i];
      }
      hc[0] /= n;
      hc[1] /= n;
      double p[3][N], mc[2 * N];
      for (int i = 0; i < n; ++i) {
        p[0][i] = h[0][i] - hc[0];
        p[1][i] = h[1][i] - hc[1];
        mc[i] = p[0][i];
        mc[i + n] = p[1][i];
      }
      double Q = 0;
      for (int k = 0; k < 2 * n; ++k)
        Q += mc[k] * mc[k];
      const double s = std::sqrt(n * 1. / Q);
      for (int i = 0; i < n; ++i) {
        p[0][i] *= s;
        p[1][i] *= s;
        p[2][i] = p[0][i] * p[0][i] + p[1][i] * p[1][i];
      }
      // cost function and its minimum
      double r0[3];
      for (int r = 0; r < 3; ++r) {
        r0[r] = 0;
        for (int i = 0; i < n; ++i)
          r0[r] += p[r][i] * w[i];
      }
      double X[3][N];
      for (int r = 0; r < 3; ++r)
        for (int i = 0; i < n; ++i)
          X[r][i] = p[r][i] - r0[r];
      double A[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          double v = 0;
          for (int i = 0; i < n; ++i) {
            double t = 0;
            for (int j = 0; j < n; ++j)
              t += G[i][j] * X[b][j];
            v += X[a][i] * t;
          }
          A[a][b] = v;
        }
      double chi2, v[3];
      minEigen3D(A, chi2, v);
      if (!(v[2] > 0))
        for (int r = 0; r < 3; ++r)
          v[r] = -v[r];
      const double c = -(v[0] * r0[0] + v[1] * r0[1] + v[2] * r0[2]);
      const double H = std::sqrt(1. - sq(v[2]) - 4. * c * v[2]);
      const double v2x2inv = 1. / (2. * v[2]);
      const double sInv = 1. / s;
      const double uvr[3] = {-v[0] * v2x2inv, -v[1] * v2x2inv, H * v2x2inv};
      out.par[0] = uvr[0] * sInv + hc[0];
      out.par[1] = uvr[1] * sInv + hc[1];
      out.par[2] = uvr[2] * sInv;
      // RiemannFit.h:292-297 charge
      out.q = ((h[0][1] - h[0][0]) * (out.par[1] - h[1][0]) - (h[1][1] - h[1][0]) * (out.par[0] - h[0][0]) > 0) ? -1
                                                                                                                : 1;
      out.chi2 = float(std::fabs(chi2) * renorm / sq(2 * v[2] * uvr[2] * s));

      // ---- error propagation (RiemannFit.h:592-781) ----
      constexpr int m = 2 * N;
      double c2 = 0;  // mc^T V mc
      double vNorm2 = 0;
      for (int a = 0; a < m; ++a) {
        double t = 0;
        for (int b = 0; b < m; ++b) {
          t += V[a][b] * mc[b];
          vNorm2 += V[a][b] * V[a][b];
        }
        c2 += mc[a] * t;
      }
      // covariance of the centred and scaled points, upper triangle mirrored
      double Vcs[m][m];
      const double f = sq(sq(s)) * 1. / (4. * Q * n) * (2. * vNorm2 + 4. * c2);
      for (int a = 0; a < m; ++a)
        for (int b = a; b < m; ++b) {
          Vcs[a][b] = sq(s) * V[a][b] + f * (mc[a] * mc[b]);
          Vcs[b][a] = Vcs[a][b];
        }
      // vcs[0][0], vcs[0][1], vcs[1][0], vcs[1][1] as N x N blocks; vcs01 is the upper
      // block as computed, vcs10 its transpose
      auto v00 = [&](int i, int j) { return Vcs[i][j]; };
      auto v01 = [&](int i, int j) { return Vcs[i][j + n]; };
      auto v10 = [&](int i, int j) { return Vcs[j][i + n]; };
      auto v11 = [&](int i, int j) { return Vcs[i + n][j + n]; };
      // C[i][j] for (00, 01, 02, 11, 12, 22); C[j][i] = C[i][j]^T
      double Cm[6][N][N];
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          const double t0 = p[0][j], t1 = p[1][j];
          Cm[0][i][j] = v00(i, j);
          Cm[1][i][j] = v01(i, j);
          Cm[2][i][j] = 2. * (v00(i, j) * t0 + v01(i, j) * t1);
          Cm[3][i][j] = v11(i, j);
          Cm[4][i][j] = 2. * (v10(i, j) * t0 + v11(i, j) * t1);
        }
      for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
          const double t00 = p[0][i] * p[0][j], t01 = p[0][i] * p[1][j], t11 = p[1][i] * p[1][j];
          const double t10 = p[0][j] * p[1][i];
          const double x = 2. * (v00(i, j) * v00(i, j) + v00(i, j) * v01(i, j) + v11(i, j) * v10(i, j) +
                                 v11(i, j) * v11(i, j)) +
                           4. * (v00(i, j) * t00 + v01(i, j) * t01 + v10(i, j) * t10 + v11(i, j) * t11);
          Cm[5][i][j] = Cm[5][j][i] = x;
        }
      constexpr int nu[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
      auto cIndex = [&](int i, int j) {
        for (int a = 0; a < 6; ++a)
          if (nu[a][0] == i && nu[a][1] == j)
            return a;
        return -1;
      };
      // C0: covariance of the centre of gravity r0
      double C0[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) {
          const int a = cIndex(i, j);
          double t = 0;
          for (int k = 0; k < n; ++k) {
            double u = 0;
            for (int l = 0; l < n; ++l)
              u += Cm[a][k][l] * w[l];
            t += w[k] * u;
          }
          C0[i][j] = C0[j][i] = t;
        }
      // H = I - 1 w^T, s_v = H p^T, D[i][j] = (H C[i][j] H^T) o (w w^T)
      double Hm[N][N], sv[N][3];
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          Hm[i][j] = (i == j ? 1. : 0.) - w[j];
      for (int i = 0; i < n; ++i)
        for (int r = 0; r < 3; ++r) {
          double t = 0;
          for (int k = 0; k < n; ++k)
            t += Hm[i][k] * p[r][k];
          sv[i][r] = t;
        }
      double Dm[6][N][N];
      for (int a = 0; a < 6; ++a) {
        double T[N][N];
        for (int i = 0; i < n; ++i)
          for (int l = 0; l < n; ++l) {
            double t = 0;
            for (int k = 0; k < n; ++k)
              t += Hm[i][k] * Cm[a][k][l];
            T[i][l] = t;
          }
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) {
            double t = 0;
            for (int l = 0; l < n; ++l)
              t += T[i][l] * Hm[j][l];
            Dm[a][i][j] = t * (w[i] * w[j]);
          }
      }
      // D[i][j] for any i, j (the lower ones transposed)
      auto D = [&](int i, int j, int r, int c) {
        return i <= j ? Dm[cIndex(i, j)][r][c] : Dm[cIndex(j, i)][c][r];
      };
      double E[6][6];
      for (int a = 0; a < 6; ++a) {
        const int i = nu[a][0], j = nu[a][1];
        for (int b = a; b < 6; ++b) {
          const int k = nu[b][0], l = nu[b][1];
          double t0[N], t1[N];
          for (int r = 0; r < n; ++r) {
            double x = 0, y = 0;
            if (l == k) {
              for (int q = 0; q < n; ++q) {
                x += D(j, l, r, q) * sv[q][l];
                y += D(i, l, r, q) * sv[q][l];
              }
              x *= 2.;
              y *= 2.;
            } else {
              double x1 = 0, x2 = 0, y1 = 0, y2 = 0;
              for (int q = 0; q < n; ++q) {
                x1 += D(j, l, r, q) * sv[q][k];
                x2 += D(j, k, r, q) * sv[q][l];
                y1 += D(i, l, r, q) * sv[q][k];
                y2 += D(i, k, r, q) * sv[q][l];
              }
              x = x1 + x2;
              y = y1 + y2;
            }
            t0[r] = x;
            t1[r] = i == j ? x : y;
          }
          double e = 0;
          if (i == j) {
            for (int r = 0; r < n; ++r)
              e += sv[r][i] * (t0[r] + t1[r]);
          } else {
            double e0 = 0, e1 = 0;
            for (int r = 0; r < n; ++r) {
              e0 += sv[r][i] * t0[r];
              e1 += sv[r][j] * t1[r];
            }
            e = e0 + e1;
          }
          E[a][b] = E[b][a] = e;
        }
      }
      // J2: numerical Jacobian of the eigenvector, with the float solver
      constexpr double epsilon = 1.e-4;  // FitUtils.h:11
      double J2[3][6];
      for (int a = 0; a < 6; ++a) {
        const int i = nu[a][0], j = nu[a][1];
        double Ad[3][3];
        std::memcpy(Ad, A, sizeof(Ad));
        const double d = std::fabs(A[i][j] * epsilon);
        Ad[i][j] += d;
        if (i != j)
          Ad[j][i] += d;
        double col[3];
        minEigen3DFast(Ad, col);
        const int sign = col[2] > 0 ? 1 : -1;
        for (int r = 0; r < 3; ++r)
          J2[r][a] = (col[r] * sign - v[r]) / d;
      }
      // joint covariance of (v0, v1, v2, c)
      double t0[3][3];
      for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) {
          double x = 0;
          for (int a = 0; a < 6; ++a) {
            double y = 0;
            for (int b = 0; b < 6; ++b)
              y += E[a][b] * J2[q][b];
            x += J2[r][a] * y;
          }
          t0[r][q] = x;
        }
      double Cvc[4][4];
      for (int r = 0; r < 3; ++r) {
        double t1 = 0;
        for (int q = 0; q < 3; ++q) {
          Cvc[r][q] = t0[r][q];
          t1 += -t0[r][q] * r0[q];
        }
        Cvc[r][3] = Cvc[3][r] = t1;
      }
      {
        double cm1 = 0, cm2 = 0, cm3 = 0;
        for (int r = 0; r < 3; ++r)
          for (int q = 0; q < 3; ++q) {
            cm1 += v[r] * C0[r][q] * v[q];
            cm2 += C0[r][q] * t0[r][q];
            cm3 += r0[r] * t0[r][q] * r0[q];
          }
        Cvc[3][3] = cm1 + cm2 + cm3;
      }
      // Jacobian (v0, v1, v2, c) -> (X0, Y0, R)
      const double ti = 1. / H;
      const double J3[3][4] = {{-v2x2inv, 0, v[0] * sq(v2x2inv) * 2., 0},
                               {0, -v2x2inv, v[1] * sq(v2x2inv) * 2., 0},
                               {v[0] * v2x2inv * ti, v[1] * v2x2inv * ti,
                                -H * sq(v2x2inv) * 2. - (2. * c + v[2]) * v2x2inv * ti, -ti}};
      double jq = 0;  // Jq V Jq^T with Jq = mc^T s / n
      for (int a = 0; a < m; ++a) {
        double t = 0;
        for (int b = 0; b < m; ++b)
          t += V[a][b] * (mc[b] * s * 1. / n);
        jq += (mc[a] * s * 1. / n) * t;
      }
      for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) {
          double x = 0;
          for (int a = 0; a < 4; ++a) {
            double y = 0;
            for (int b = 0; b < 4; ++b)
              y += Cvc[a][b] * J3[q][b];
            x += J3[r][a] * y;
          }
          out.cov[r][q] = x * sq(sInv) + (uvr[r] * uvr[q]) * jq;
        }
    }

    // RiemannFit.h:801-937 lineFit
    template <int N>
    void riemannLineFit(const double h[3][N], const float ge[6][N], const RCircle& ci, const double ff[4], double B,
                        RLine& out) {
      constexpr int n = N;
      double theta = -ci.q * rm::atan(ff[3]);
      theta = theta < 0. ? theta + M_PI : theta;
      const double sT = rm::sin(theta), cT = rm::cos(theta);
      const double rot[2][2] = {{sT, cT}, {-cT, sT}};
      double p2[2][N], covSz11[N];
      const double ox = ci.par[0], oy = ci.par[1];
      for (int i = 0; i < n; ++i) {
        const double px = h[0][i] - ox, py = h[1][i] - oy;
        const double cross = (-ox) * py - (-oy) * px;
        const double dot = (-ox) * px + (-oy) * py;
        const double qAtan2 = -ci.q * rm::atan2(cross, dot);
        p2[0][i] = qAtan2 * ci.par[2];
        const double t0 = -ci.q * ci.par[2] * 1. / (sq(dot) + sq(cross));
        const double dX0 = -t0 * ((py + oy) * dot - (px - ox) * cross);
        const double dY0 = t0 * ((px + ox) * dot - (oy - py) * cross);
        const double dR = qAtan2;
        const double dx = t0 * (oy * dot + ox * cross);
        const double dy = t0 * (-ox * dot + oy * cross);
        const double J[2][6] = {{dX0, dY0, dR, dx, dy, 0.}, {0., 0., 0., 0., 0., 1.}};
        double C[6][6] = {};
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            C[a][b] = ci.cov[a][b];
        C[3][3] = ge[0][i];
        C[4][4] = ge[2][i];
        C[5][5] = ge[5][i];
        C[3][4] = C[4][3] = ge[1][i];
        C[3][5] = C[5][3] = ge[3][i];
        C[4][5] = C[5][4] = ge[4][i];
        double M[2][2];
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b) {
            double x = 0;
            for (int k = 0; k < 6; ++k) {
              double y = 0;
              for (int l = 0; l < 6; ++l)
                y += C[k][l] * J[b][l];
              x += J[a][k] * y;
            }
            M[a][b] = x;
          }
        double x = 0;  // (rot M rot^T)(1, 1)
        for (int k = 0; k < 2; ++k)
          for (int l = 0; l < 2; ++l)
            x += rot[1][k] * M[k][l] * rot[1][l];
        covSz11[i] = x;
      }
      for (int i = 0; i < n; ++i)
        p2[1][i] = h[2][i];
      double covMs[N][N];
      scatterCovLine<N>(covSz11, ff, p2[0], p2[1], B, covMs);
      double pr[2][N];
      for (int i = 0; i < n; ++i) {
        pr[0][i] = rot[0][0] * p2[0][i] + rot[0][1] * p2[1][i];
        pr[1][i] = rot[1][0] * p2[0][i] + rot[1][1] * p2[1][i];
      }
      double Vi[N][N];
      invertSPDn<N>(&covMs[0][0], &Vi[0][0]);
      // A = (1..., s'...): A Vi A^T (2x2), inverted
      double Am[2][N];
      for (int i = 0; i < n; ++i) {
        Am[0][i] = 1.;
        Am[1][i] = pr[0][i];
      }
      double AV[2][N];
      for (int a = 0; a < 2; ++a)
        for (int j = 0; j < n; ++j) {
          double x = 0;
          for (int k = 0; k < n; ++k)
            x += Am[a][k] * Vi[k][j];
          AV[a][j] = x;
        }
      double P[2][2], Pi[2][2];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          double x = 0;
          for (int j = 0; j < n; ++j)
            x += AV[a][j] * Am[b][j];
          P[a][b] = x;
        }
      invertSPDn<2>(&P[0][0], &Pi[0][0]);
      double sol[2];
      for (int a = 0; a < 2; ++a) {
        double x = 0;
        for (int b = 0; b < 2; ++b) {
          double y = 0;
          for (int j = 0; j < n; ++j)
            y += AV[b][j] * pr[1][j];
          x += Pi[a][b] * y;
        }
        sol[a] = x;
      }
      const double cf = 1. / (sT - sol[1] * cT);
      const double Jm[2][2] = {{0., cf * cf}, {cf, sol[0] * cT * cf * cf}};
      out.par[0] = cf * (sol[1] * sT + cT);
      out.par[1] = cf * sol[0];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          double x = 0;
          for (int k = 0; k < 2; ++k)
            for (int l = 0; l < 2; ++l)
              x += Jm[a][k] * Pi[k][l] * Jm[b][l];
          out.cov[a][b] = x;
        }
      double res[N];
      for (int i = 0; i < n; ++i)
        res[i] = pr[1][i] - (Am[0][i] * sol[0] + Am[1][i] * sol[1]);
      double chi2 = 0;
      for (int i = 0; i < n; ++i) {
        double t = 0;
        for (int j = 0; j < n; ++j)
          t += Vi[i][j] * res[j];
        chi2 += res[i] * t;
      }
      out.chi2 = chi2;
    }

    // FitUtils.h:159-180 fromCircleToPerigee: (X0, Y0, R) -> (phi, Tip, q/R)
    void toPerigee(RCircle& c) {
      const double t0 = sq(c.par[0]) + sq(c.par[1]);
      const double t1 = std::sqrt(t0);
      const double pak[3] = {rm::atan2(c.q * c.par[0], -c.q * c.par[1]), c.q * (t1 - c.par[2]), c.q / c.par[2]};
      const double t2 = sq(c.par[0]) * 1. / t0;
      const double t3 = 1. / t1 * c.q;
      const double J[3][3] = {{-c.par[1] * t2 * 1. / sq(c.par[0]), t2 * 1. / c.par[0], 0.},
                              {c.par[0] * t3, c.par[1] * t3, double(-c.q)},
                              {0., 0., -c.q / (c.par[2] * c.par[2])}};
      double r[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          double x = 0;
          for (int k = 0; k < 3; ++k) {
            double y = 0;
            for (int l = 0; l < 3; ++l)
              y += c.cov[k][l] * J[b][l];
            x += J[a][k] * y;
          }
          r[a][b] = x;
        }
      std::memcpy(c.cov, r, sizeof(r));
      std::memcpy(c.par, pak, sizeof(pak));
    }

    // one tuple: RiemannFitOnGPU.h:20-187 (kernel_FastFit, kernel_CircleFit,
    // kernel_LineFit) and TrajectoryStateSoA::copyFromCircle
    template <int N>
    void riemannOne(const double h[3][N], const float ge[6][N], double B, float* st, float* cv, float* pt, float* eta,
                    float* chi2) {
      double ff[4];
      riemannFastFit<N>(h, ff);
      double rad[N];
      for (int i = 0; i < N; ++i)
        rad[i] = std::sqrt(h[0][i] * h[0][i] + h[1][i] * h[1][i]);
      double V[2 * N][2 * N] = {};
      for (int i = 0; i < N; ++i) {  // FitUtils.h:102-122 loadCovariance2D
        V[i][i] = ge[0][i];
        V[i + N][i + N] = ge[2][i];
        V[i][i + N] = V[i + N][i] = ge[1][i];
      }
      RCircle ci;
      riemannCircleFit<N>(h, V, ff, rad, B, ci);
      RLine li;
      riemannLineFit<N>(h, ge, ci, ff, B, li);
      toPerigee(ci);
      const float b = 1.f / float(B);
      st[0] = float(ci.par[0]);
      st[1] = float(ci.par[1]);
      st[2] = b * float(ci.par[2]);
      st[3] = float(li.par[0]);
      st[4] = float(li.par[1]);
      cv[0] = float(ci.cov[0][0]);
      cv[1] = float(ci.cov[0][1]);
      cv[2] = b * float(ci.cov[0][2]);
      cv[3] = cv[4] = 0;
      cv[5] = float(ci.cov[1][1]);
      cv[6] = b * float(ci.cov[1][2]);
      cv[7] = cv[8] = 0;
      cv[9] = b * b * float(ci.cov[2][2]);
      cv[10] = cv[11] = 0;
      cv[12] = float(li.cov[0][0]);
      cv[13] = float(li.cov[0][1]);
      cv[14] = float(li.cov[1][1]);
      *pt = float(B / std::fabs(ci.par[2]));
      *eta = asinhf(float(li.par[0]));
      *chi2 = float((double(ci.chi2) + li.chi2) / (2 * N - 5));
    }
  }  // namespace

  // a18/a19 alternative: HelixFitOnGPU::launchRiemannKernelsOnCPU (RiemannFitOnGPU.cc:3-113)
  // with fit5as4 (CAHitNtupletGeneratorOnGPU.cc:53)
  void fitTracksRiemann(const Conditions& c, Event& ev) {
    auto& C = ev.ca;
    C.state.assign(ev.caps.maxTuples, {0, 0, 0, 0, 0});
    C.cov.assign(ev.caps.maxTuples, {});
    C.pt.assign(ev.caps.maxTuples, 0);
    C.eta.assign(ev.caps.maxTuples, 0);
    C.chi2.assign(ev.caps.maxTuples, 0);
    if (ev.nHits == 0)
      return;
    const double B = double(c.bField);
    auto one = [&](auto tag, uint32_t tkid) {
      constexpr int N = decltype(tag)::value;
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
      riemannOne<N>(h, ge, B, C.state[tkid].data(), C.cov[tkid].data(), &C.pt[tkid], &C.eta[tkid], &C.chi2[tkid]);
    };
    for (uint32_t nh = 3; nh <= 5; ++nh)
      for (uint32_t k = C.multOff[nh]; k < C.multOff[nh + 1]; ++k) {
        if (nh == 3)
          one(std::integral_constant<int, 3>{}, C.multBins[k]);
        else
          one(std::integral_constant<int, 4>{}, C.multBins[k]);
      }
  }

  // direct entries for the tests: hits[3][n], ge[6][n]
  int fitHelixRiemann(int n, const double* hits, const float* ge, double B, float* state, float* cov, float* pt,
                      float* eta, float* chi2) {
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
      riemannOne<N>(h, g, B, state, cov, pt, eta, chi2);
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

  // the circle part alone in (X0, Y0, R) with its covariance, and the fast fit
  int circleFitRiemann(int n, const double* hits, const float* ge, double B, double* fast, double* par,
                       double* cov, int* q, float* chi2) {
    auto run = [&](auto tag) {
      constexpr int N = decltype(tag)::value;
      double h[3][N];
      double V[2 * N][2 * N] = {};
      double rad[N];
      for (int r = 0; r < 3; ++r)
        for (int i = 0; i < N; ++i)
          h[r][i] = hits[r * n + i];
      for (int i = 0; i < N; ++i) {
        V[i][i] = ge[0 * n + i];
        V[i + N][i + N] = ge[2 * n + i];
        V[i][i + N] = V[i + N][i] = ge[1 * n + i];
        rad[i] = std::sqrt(h[0][i] * h[0][i] + h[1][i] * h[1][i]);
      }
      riemannFastFit<N>(h, fast);
      RCircle ci;
      riemannCircleFit<N>(h, V, fast, rad, B, ci);
      std::memcpy(par, ci.par, sizeof(ci.par));
      std::memcpy(cov, ci.cov, sizeof(ci.cov));
      *q = ci.q;
      *chi2 = ci.chi2;
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

  // the fit's elementary functions: 0 atan2(y, x), 1 atan(x), 2 sin(x), 3 cos(x), 4 log(x)
  double rmathForTest(int fn, double x, double y) {
    switch (fn) {
      case 0:
        return rm::atan2(y, x);
      case 1:
        return rm::atan(x);
      case 2:
        return rm::sin(x);
      case 3:
        return rm::cos(x);
      default:
        return rm::log(x);
    }
  }

  // eigen3Direct on a symmetric 3x3 (row-major; lower triangle read), in double or float
  void eigen3ForTest(const double* a, int useFloat, double* evals, double* evecs) {
    if (useFloat) {
      float m[3][3], ev[3], V[3][3];
      for (int i = 0; i < 9; ++i)
        m[i / 3][i % 3] = float(a[i]);
      eigen3Direct<float>(m, ev, V);
      for (int i = 0; i < 3; ++i)
        evals[i] = ev[i];
      for (int i = 0; i < 9; ++i)
        evecs[i] = V[i / 3][i % 3];
    } else {
      double m[3][3], V[3][3];
      for (int i = 0; i < 9; ++i)
        m[i / 3][i % 3] = a[i];
      eigen3Direct<double>(m, evals, V);
      for (int i = 0; i < 9; ++i)
        evecs[i] = V[i / 3][i % 3];
    }
  }

}  // namespace orc

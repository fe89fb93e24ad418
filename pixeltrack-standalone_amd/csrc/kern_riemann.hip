// Riemann helix fit on gfx950: the reference's off-by-default alternative to the
// broken-line fit (RiemannFitOnGPU.h:20-187 kernel_FastFit, kernel_CircleFit,
// kernel_LineFit; RiemannFit.h:30-885; FitUtils.h:102-190), one tuple per lane in
// double precision over the entries of the tuple multiplicity container
// (triplets with N = 3, quadruplets and pentuplets on their first 4 hits).
//
// The three reference kernels hand their intermediates (hits, errors, fast fit,
// circle) through global buffers; here one lane runs all three on its tuple
// with the intermediates in registers and private memory.  The small symmetric
// positive-definite matrices are inverted through LDL^T.  The eigenvector of
// the smallest eigenvalue of the 3x3 cost matrix is the closed form of Eigen's
// SelfAdjointEigenSolver::computeDirect (in double; in float for the numerical
// Jacobian).  The elementary functions (atan2, atan, sin, cos, log) are the fit's
// own, from IEEE basic operations, so that the results do not depend on the
// math library's last bits.
#include <hip/hip_runtime.h>

#include "pxgpu_internal.h"

namespace pxg {

  namespace {
    __device__ __forceinline__ double sq(double a) { return a * a; }

    // ---- elementary functions from IEEE basic operations only -----------------------
    // The fit's own, so that device and host compute the same bits: its numerical
    // Jacobian (float eigenvectors of a perturbed matrix, RiemannFit.h:732-739) and its
    // smallest eigenvalue (a difference of O(1) terms) turn last-bit differences of
    // library functions into visible ones.  Accuracy: a few units in the last place.
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
      __device__ inline double log(double x) {
        if (!(x > 0) || x == HUGE_VAL)
          return x == 0 ? -HUGE_VAL : (x == HUGE_VAL ? x : NAN);
        int k;
        double m = frexp(x, &k);
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
      __device__ inline double atan(double x) {
        double a = fabs(x);
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

      __device__ inline double atan2(double y, double x) {
        if (x > 0)
          return atan(y / x);
        if (x < 0) {
          const double r = atan(y / x);
          return signbit(y) ? r - kPi : r + kPi;
        }
        if (x == 0) {
          if (y > 0)
            return kPio2;
          if (y < 0)
            return -kPio2;
          return signbit(x) ? (signbit(y) ? -kPi : kPi) : y;
        }
        return x + y;  // NaN
      }

      // x = q pi/2 + r with |r| <= pi/4, then the Taylor series of sin and cos to r^21
      __device__ inline void sincos(double x, double& s, double& c) {
        const double q = rint(x * kTwoOverPi);
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
      __device__ inline double sin(double x) {
        double s, c;
        sincos(x, s, c);
        return s;
      }
      __device__ inline double cos(double x) {
        double s, c;
        sincos(x, s, c);
        return c;
      }
    }  // namespace rm


    template <int n>
    __device__ void invertSPDr(const double* a, double* inv) {
      double L[n * n], D[n], Li[n * n];
#pragma unroll
      for (int k = 0; k < n * n; ++k) {
        L[k] = 0;
        Li[k] = 0;
      }
#pragma unroll
      for (int j = 0; j < n; ++j) {
        double d = a[j * n + j];
#pragma unroll
        for (int k = 0; k < j; ++k)
          d -= L[j * n + k] * L[j * n + k] * D[k];
        D[j] = d;
        L[j * n + j] = 1.0;
#pragma unroll
        for (int i = j + 1; i < n; ++i) {
          double v = a[i * n + j];
#pragma unroll
          for (int k = 0; k < j; ++k)
            v -= L[i * n + k] * L[j * n + k] * D[k];
          L[i * n + j] = v / d;
        }
      }
#pragma unroll
      for (int j = 0; j < n; ++j) {
        Li[j * n + j] = 1.0;
#pragma unroll
        for (int i = j + 1; i < n; ++i) {
          double v = 0;
#pragma unroll
          for (int k = j; k < i; ++k)
            v -= L[i * n + k] * Li[k * n + j];
          Li[i * n + j] = v;
        }
      }
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
          double v = 0;
#pragma unroll
          for (int k = i; k < n; ++k)
            v += Li[k * n + i] * Li[k * n + j] / D[k];
          inv[i * n + j] = v;
          inv[j * n + i] = v;
        }
    }

    // ---- closed-form 3x3 symmetric eigen decomposition -----------------------------
    __device__ __forceinline__ double angAtan2(double y, double x) { return rm::atan2(y, x); }
    __device__ __forceinline__ double angCos(double a) { return rm::cos(a); }
    __device__ __forceinline__ double angSin(double a) { return rm::sin(a); }
    __device__ __forceinline__ float angAtan2(float y, float x) { return float(rm::atan2(double(y), double(x))); }
    __device__ __forceinline__ float angCos(float a) { return float(rm::cos(double(a))); }
    __device__ __forceinline__ float angSin(float a) { return float(rm::sin(double(a))); }
    template <typename T>
    __device__ __forceinline__ T epsOf();
    template <>
    __device__ __forceinline__ double epsOf<double>() {
      return 2.220446049250313e-16;
    }
    template <>
    __device__ __forceinline__ float epsOf<float>() {
      return 1.1920929e-07f;
    }

    template <typename T>
    __device__ void eigenRoots(const T m[3][3], T r[3]) {
      const T inv3 = T(1) / T(3);
      const T sqrt3 = sqrt(T(3));
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
      const T rho = sqrt(aOver3);
      const T theta = angAtan2(sqrt(q), halfB) * inv3;
      const T cosT = angCos(theta), sinT = angSin(theta);
      r[0] = c2over3 - rho * (cosT + sqrt3 * sinT);
      r[1] = c2over3 - rho * (cosT - sqrt3 * sinT);
      r[2] = c2over3 + T(2) * rho * cosT;
    }

    template <typename T>
    __device__ __forceinline__ void cross3(const T a[3], const T b[3], T c[3]) {
      c[0] = a[1] * b[2] - a[2] * b[1];
      c[1] = a[2] * b[0] - a[0] * b[2];
      c[2] = a[0] * b[1] - a[1] * b[0];
    }

    // kernel of a rank-2 matrix: the larger cross product of the column with the
    // largest diagonal entry with the other two columns; rep = that column
    template <typename T>
    __device__ void extractKernel(const T m[3][3], T res[3], T rep[3]) {
      int i0 = 0;
      T best = fabs(m[0][0]);
#pragma unroll
      for (int i = 1; i < 3; ++i)
        if (fabs(m[i][i]) > best) {
          best = fabs(m[i][i]);
          i0 = i;
        }
      const int i1 = i0 == 2 ? 0 : i0 + 1, i2 = i0 == 0 ? 2 : i0 - 1;
      T ca[3], cb[3], c1[3], c2[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        rep[r] = i0 == 0 ? m[r][0] : (i0 == 1 ? m[r][1] : m[r][2]);
        ca[r] = i1 == 0 ? m[r][0] : (i1 == 1 ? m[r][1] : m[r][2]);
        cb[r] = i2 == 0 ? m[r][0] : (i2 == 1 ? m[r][1] : m[r][2]);
      }
      cross3(rep, ca, c1);
      cross3(rep, cb, c2);
      const T n0 = c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2];
      const T n1 = c2[0] * c2[0] + c2[1] * c2[1] + c2[2] * c2[2];
      const bool first = n0 > n1;
      const T s = sqrt(first ? n0 : n1);
#pragma unroll
      for (int r = 0; r < 3; ++r)
        res[r] = (first ? c1[r] : c2[r]) / s;
    }

    // eigenvalues ascending, eigenvectors as the columns of V; the lower triangle is read
    template <typename T>
    __device__ void eigen3Direct(const T a[3][3], T ev[3], T V[3][3]) {
      T m[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          m[i][j] = i >= j ? a[i][j] : a[j][i];
      const T shift = (m[0][0] + m[1][1] + m[2][2]) / T(3);
#pragma unroll
      for (int i = 0; i < 3; ++i)
        m[i][i] -= shift;
      T scale = T(0);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          scale = fabs(m[i][j]) > scale ? fabs(m[i][j]) : scale;
      if (scale > T(0))
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            m[i][j] /= scale;
      eigenRoots(m, ev);
      const T eps = epsOf<T>();
      if (ev[2] - ev[0] <= eps) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            V[i][j] = i == j ? T(1) : T(0);
      } else {
        T tmp[3][3];
        T d0 = ev[2] - ev[1];
        const T d1 = ev[1] - ev[0];
        const bool swapKL = d0 > d1;
        if (swapKL)
          d0 = d1;
        const T ek = swapKL ? ev[2] : ev[0], el = swapKL ? ev[0] : ev[2];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            tmp[i][j] = m[i][j] - (i == j ? ek : T(0));
        T vk[3], vl[3];
        extractKernel(tmp, vk, vl);
        if (d0 <= T(2) * eps * d1) {
          // the two other eigenvalues coincide numerically: the saved column, made unit
          // length after the literal update of the restated algorithm
          const T dot = vk[0] * vl[0] + vk[1] * vl[1] + vk[2] * vl[2];
#pragma unroll
          for (int r = 0; r < 3; ++r)
            vl[r] -= dot * vl[r];
          const T n = vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2];
          if (n > T(0)) {
            const T s = sqrt(n);
#pragma unroll
            for (int r = 0; r < 3; ++r)
              vl[r] /= s;
          }
        } else {
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
              tmp[i][j] = m[i][j] - (i == j ? el : T(0));
          T dummy[3];
          extractKernel(tmp, vl, dummy);
        }
        // columns k and l (k = 0, l = 2, or swapped)
        T c0[3], c2[3], c1[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          c0[r] = swapKL ? vl[r] : vk[r];
          c2[r] = swapKL ? vk[r] : vl[r];
        }
        cross3(c2, c0, c1);
        const T n = c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2];
        const T s = n > T(0) ? sqrt(n) : T(1);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          V[r][0] = c0[r];
          V[r][1] = c1[r] / s;
          V[r][2] = c2[r];
        }
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
        ev[i] = ev[i] * scale + shift;
    }

    // RiemannFit.h:304-315 min_eigen3D
    __device__ void minEigen3D(const double A[3][3], double& chi2, double v[3]) {
      double ev[3], V[3][3];
      eigen3Direct<double>(A, ev, V);
      int k = 0;
      if (ev[1] < ev[k])
        k = 1;
      if (ev[2] < ev[k])
        k = 2;
      chi2 = k == 0 ? ev[0] : (k == 1 ? ev[1] : ev[2]);
#pragma unroll
      for (int r = 0; r < 3; ++r)
        v[r] = k == 0 ? V[r][0] : (k == 1 ? V[r][1] : V[r][2]);
    }

    // RiemannFit.h:328-334 min_eigen3D_fast (float)
    __device__ void minEigen3DFast(const double A[3][3], double v[3]) {
      float a[3][3], ev[3], V[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          a[i][j] = float(A[i][j]);
      eigen3Direct<float>(a, ev, V);
      int k = 0;
      if (ev[1] < ev[k])
        k = 1;
      if (ev[2] < ev[k])
        k = 2;
#pragma unroll
      for (int r = 0; r < 3; ++r)
        v[r] = double(k == 0 ? V[r][0] : (k == 1 ? V[r][1] : V[r][2]));
    }

    struct RCircle {
      double par[3];  // (X0, Y0, R), then (phi, Tip, q/R)
      double cov[3][3];
      int q;
      float chi2;  // FitResult.h:35
    };
    struct RLine {
      double par[2];
      double cov[2][2];
      double chi2;
    };

    // RiemannFit.h:30-40
    template <int N>
    __device__ __forceinline__ void radLenUniform(const double len[N], double rl[N]) {
      constexpr double xx0inv = 0.06 / 16.;
      rl[0] = len[0] * xx0inv;
#pragma unroll
      for (int j = 1; j < N; ++j)
        rl[j] = fabs(len[j] - len[j - 1]) * xx0inv;
    }

    // RiemannFit.h:365-417 fastFit
    template <int N>
    __device__ void riemannFastFit(const double h[3][N], double r[4]) {
      constexpr int n = N;
      const double bx0 = h[0][n / 2] - h[0][0], by0 = h[1][n / 2] - h[1][0];
      const double cx0 = h[0][n - 1] - h[0][0], cy0 = h[1][n - 1] - h[1][0];
      const double b2 = bx0 * bx0 + by0 * by0;
      const double c2 = cx0 * cx0 + cy0 * cy0;
      const bool flip = fabs(bx0) < fabs(by0);
      const double bx = flip ? by0 : bx0, by = flip ? bx0 : by0;
      const double cx = flip ? cy0 : cx0, cy = flip ? cx0 : cy0;
      const double div = 2. * (cx * by - bx * cy);
      const double y0 = (cx * b2 - bx * c2) / div;
      const double x0 = (0.5 * b2 - y0 * by) / bx;
      r[0] = h[0][0] + (flip ? y0 : x0);
      r[1] = h[1][0] + (flip ? x0 : y0);
      r[2] = sqrt(sq(x0) + sq(y0));
      const double dx = h[0][0] - r[0], dy = h[1][0] - r[1];
      const double ex = h[0][n - 1] - r[0], ey = h[1][n - 1] - r[1];
      const double dr = r[2] * rm::atan2(dx * ey - dy * ex, dx * ex + dy * ey);
      const double dz = h[2][n - 1] - h[2][0];
      r[3] = dr / dz;
    }

    // RiemannFit.h:120-158 scatter_cov_rad
    template <int N>
    __device__ void scatterCovRad(const double h[3][N], const double ff[4], const double rad[N], double B,
                                  double S[N][N]) {
      const double pt = fmin(20., ff[2] * B);
      const double p2 = pt * pt * (1. + 1. / sq(ff[3]));
      double theta = rm::atan(ff[3]);
      theta = theta < 0. ? theta + M_PI : theta;
      double sf[N], rl[N], sig2[N];
      const double f = sqrt(1. + 1. / sq(ff[3]));
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double px = h[0][i] - ff[0], py = h[1][i] - ff[1];
        const double ox = -ff[0], oy = -ff[1];
        const double cross = ox * py - oy * px;
        const double dot = ox * px + oy * py;
        sf[i] = fabs(rm::atan2(cross, dot) * ff[2]) * f;
      }
      radLenUniform<N>(sf, rl);
      const double k = 0.000225 / (p2 * sq(rm::sin(theta)));
#pragma unroll
      for (int i = 0; i < N; ++i)
        sig2[i] = sq(1. + 0.038 * rm::log(rl[i])) * rl[i] * k;
#pragma unroll
      for (int a = 0; a < N; ++a)
#pragma unroll
        for (int b = a; b < N; ++b) {
          double v = 0;
#pragma unroll
          for (int i = 0; i < (a < b ? a : b); ++i)
            v += (rad[a] - rad[i]) * (rad[b] - rad[i]) * sig2[i];
          S[a][b] = v;
          S[b][a] = v;
        }
    }

    // RiemannFit.h:442-783 circleFit with error propagation; V (2N x 2N, x... then y...)
    // gains the scattering terms
    template <int N>
    __device__ void riemannCircleFit(const double h[3][N], double V[2 * N][2 * N], const double ff[4],
                                     const double rad[N], double B, RCircle& out) {
      constexpr int n = N;
      constexpr int m = 2 * N;
      double covRad[N][N], S[N][N], G[N][N];
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = 0; j < n; ++j)
          covRad[i][j] = 0;
#pragma unroll
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
        // RiemannFit.h:169-194 cov_radtocart, in its assignment order
        double CC[m][m];
#pragma unroll
        for (int i = 0; i < n; ++i)
#pragma unroll
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
#pragma unroll
        for (int a = 0; a < m; ++a)
#pragma unroll
          for (int b = 0; b < m; ++b)
            V[a][b] += CC[a][b];
      }
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = 0; j < n; ++j)
          covRad[i][j] += S[i][j];
      invertSPDr<N>(&covRad[0][0], &G[0][0]);
      double renorm = 0;
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = 0; j < n; ++j)
          renorm += G[i][j];
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = 0; j < n; ++j)
          G[i][j] *= 1. / renorm;
      double w[N];
#pragma unroll
      for (int j = 0; j < n; ++j) {
        w[j] = 0;
#pragma unroll
        for (int i = 0; i < n; ++i)
          w[j] += G[i][j];
      }
      double hc0 = 0, hc1 = 0;
#pragma unroll
      for (int i = 0; i < n; ++i) {
        hc0 += h[0][i];
        hc1 += h[1][i];
      }
      hc0 /= n;
      hc1 /= n;
      double p[3][N], mc[m];
#pragma unroll
      for (int i = 0; i < n; ++i) {
        p[0][i] = h[0][i] - hc0;
        p[1][i] = h[1][i] - hc1;
        mc[i] = p[0][i];
        mc[i + n] = p[1][i];
      }
      double Q = 0;
#pragma unroll
      for (int k = 0; k < m; ++k)
        Q += mc[k] * mc[k];
      const double s = sqrt(n * 1. / Q);
#pragma unroll
      for (int i = 0; i < n; ++i) {
        p[0][i] *= s;
        p[1][i] *= s;
        p[2][i] = p[0][i] * p[0][i] + p[1][i] * p[1][i];
      }
      double r0[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        r0[r] = 0;
#pragma unroll
        for (int i = 0; i < n; ++i)
          r0[r] += p[r][i] * w[i];
      }
      double X[3][N];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int i = 0; i < n; ++i)
          X[r][i] = p[r][i] - r0[r];
      double A[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          double v = 0;
#pragma unroll
          for (int i = 0; i < n; ++i) {
            double t = 0;
#pragma unroll
            for (int j = 0; j < n; ++j)
              t += G[i][j] * X[b][j];
            v += X[a][i] * t;
          }
          A[a][b] = v;
        }
      double chi2, v[3];
      minEigen3D(A, chi2, v);
      if (!(v[2] > 0)) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
          v[r] = -v[r];
      }
      const double c = -(v[0] * r0[0] + v[1] * r0[1] + v[2] * r0[2]);
      const double H = sqrt(1. - sq(v[2]) - 4. * c * v[2]);
      const double v2x2inv = 1. / (2. * v[2]);
      const double sInv = 1. / s;
      const double uvr[3] = {-v[0] * v2x2inv, -v[1] * v2x2inv, H * v2x2inv};
      out.par[0] = uvr[0] * sInv + hc0;
      out.par[1] = uvr[1] * sInv + hc1;
      out.par[2] = uvr[2] * sInv;
      // RiemannFit.h:292-297 charge
      out.q = ((h[0][1] - h[0][0]) * (out.par[1] - h[1][0]) - (h[1][1] - h[1][0]) * (out.par[0] - h[0][0]) > 0) ? -1
                                                                                                                : 1;
      out.chi2 = float(fabs(chi2) * renorm / sq(2 * v[2] * uvr[2] * s));

      // ---- error propagation (RiemannFit.h:592-781) ----
      double c2 = 0, vNorm2 = 0, jq = 0;
#pragma unroll
      for (int a = 0; a < m; ++a) {
        double t = 0;
#pragma unroll
        for (int b = 0; b < m; ++b) {
          t += V[a][b] * mc[b];
          vNorm2 += V[a][b] * V[a][b];
        }
        c2 += mc[a] * t;
      }
#pragma unroll
      for (int a = 0; a < m; ++a) {  // Jq V Jq^T, Jq = mc^T s / n
        double t = 0;
#pragma unroll
        for (int b = 0; b < m; ++b)
          t += V[a][b] * (mc[b] * s * 1. / n);
        jq += (mc[a] * s * 1. / n) * t;
      }
      // covariance of the centred, scaled points (in place of V)
      double (*Vcs)[m] = V;
      {
        const double f = sq(sq(s)) * 1. / (4. * Q * n) * (2. * vNorm2 + 4. * c2);
#pragma unroll
        for (int a = 0; a < m; ++a)
#pragma unroll
          for (int b = a; b < m; ++b) {
            const double x = sq(s) * Vcs[a][b] + f * (mc[a] * mc[b]);
            Vcs[a][b] = x;
            Vcs[b][a] = x;
          }
      }
      // C[a] for (00, 01, 02, 11, 12, 22) from the blocks of Vcs and the points
      auto Cm = [&](int a, int i, int j) -> double {
        const double v00 = Vcs[i][j], v01 = Vcs[i][j + n], v10 = Vcs[j][i + n], v11 = Vcs[i + n][j + n];
        switch (a) {
          case 0:
            return v00;
          case 1:
            return v01;
          case 2:
            return 2. * (v00 * p[0][j] + v01 * p[1][j]);
          case 3:
            return v11;
          case 4:
            return 2. * (v10 * p[0][j] + v11 * p[1][j]);
          default: {
            const int lo = i < j ? i : j, hi = i < j ? j : i;  // upper triangle, mirrored
            const double u00 = Vcs[lo][hi], u01 = Vcs[lo][hi + n], u10 = Vcs[hi][lo + n], u11 = Vcs[lo + n][hi + n];
            const double t00 = p[0][lo] * p[0][hi], t01 = p[0][lo] * p[1][hi], t11 = p[1][lo] * p[1][hi];
            const double t10 = p[0][hi] * p[1][lo];
            return 2. * (u00 * u00 + u00 * u01 + u11 * u10 + u11 * u11) +
                   4. * (u00 * t00 + u01 * t01 + u10 * t10 + u11 * t11);
          }
        }
      };
      constexpr int nu[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
      double C0[3][3];
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        double t = 0;
#pragma unroll
        for (int k = 0; k < n; ++k) {
          double u = 0;
#pragma unroll
          for (int l = 0; l < n; ++l)
            u += Cm(a, k, l) * w[l];
          t += w[k] * u;
        }
        C0[nu[a][0]][nu[a][1]] = t;
        C0[nu[a][1]][nu[a][0]] = t;
      }
      double Hm[N][N], sv[N][3];
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = 0; j < n; ++j)
          Hm[i][j] = (i == j ? 1. : 0.) - w[j];
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          double t = 0;
#pragma unroll
          for (int k = 0; k < n; ++k)
            t += Hm[i][k] * p[r][k];
          sv[i][r] = t;
        }
      double Dm[6][N][N];
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        double T[N][N];
#pragma unroll
        for (int i = 0; i < n; ++i)
#pragma unroll
          for (int l = 0; l < n; ++l) {
            double t = 0;
#pragma unroll
            for (int k = 0; k < n; ++k)
              t += Hm[i][k] * Cm(a, k, l);
            T[i][l] = t;
          }
#pragma unroll
        for (int i = 0; i < n; ++i)
#pragma unroll
          for (int j = 0; j < n; ++j) {
            double t = 0;
#pragma unroll
            for (int l = 0; l < n; ++l)
              t += T[i][l] * Hm[j][l];
            Dm[a][i][j] = t * (w[i] * w[j]);
          }
      }
      // D[i][j] for any i, j (the lower ones transposed); the index of (i, j), i <= j
      auto cIdx = [](int i, int j) { return i == 0 ? j : (i == 1 ? 2 + j : 5); };
      auto D = [&](int i, int j, int r, int q) -> double {
        return i <= j ? Dm[cIdx(i, j)][r][q] : Dm[cIdx(j, i)][q][r];
      };
      double E[6][6];
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        const int i = nu[a][0], j = nu[a][1];
#pragma unroll
        for (int b = a; b < 6; ++b) {
          const int k = nu[b][0], l = nu[b][1];
          double t0[N], t1[N];
#pragma unroll
          for (int r = 0; r < n; ++r) {
            double x = 0, y = 0;
            if (l == k) {
#pragma unroll
              for (int q = 0; q < n; ++q) {
                x += D(j, l, r, q) * sv[q][l];
                y += D(i, l, r, q) * sv[q][l];
              }
              x *= 2.;
              y *= 2.;
            } else {
              double x1 = 0, x2 = 0, y1 = 0, y2 = 0;
#pragma unroll
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
#pragma unroll
            for (int r = 0; r < n; ++r)
              e += sv[r][i] * (t0[r] + t1[r]);
          } else {
            double e0 = 0, e1 = 0;
#pragma unroll
            for (int r = 0; r < n; ++r) {
              e0 += sv[r][i] * t0[r];
              e1 += sv[r][j] * t1[r];
            }
            e = e0 + e1;
          }
          E[a][b] = e;
          E[b][a] = e;
        }
      }
      // J2: numerical Jacobian of the eigenvector with the float solver (FitUtils.h:11)
      constexpr double epsilon = 1.e-4;
      double J2[3][6];
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        const int i = nu[a][0], j = nu[a][1];
        double Ad[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int q = 0; q < 3; ++q)
            Ad[r][q] = A[r][q];
        const double d = fabs(A[i][j] * epsilon);
        Ad[i][j] += d;
        if (i != j)
          Ad[j][i] += d;
        double col[3];
        minEigen3DFast(Ad, col);
        const int sign = col[2] > 0 ? 1 : -1;
#pragma unroll
        for (int r = 0; r < 3; ++r)
          J2[r][a] = (col[r] * sign - v[r]) / d;
      }
      double t0[3][3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          double x = 0;
#pragma unroll
          for (int a = 0; a < 6; ++a) {
            double y = 0;
#pragma unroll
            for (int b = 0; b < 6; ++b)
              y += E[a][b] * J2[q][b];
            x += J2[r][a] * y;
          }
          t0[r][q] = x;
        }
      double Cvc[4][4];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        double t1 = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          Cvc[r][q] = t0[r][q];
          t1 += -t0[r][q] * r0[q];
        }
        Cvc[r][3] = t1;
        Cvc[3][r] = t1;
      }
      {
        double cm1 = 0, cm2 = 0, cm3 = 0;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            cm1 += v[r] * C0[r][q] * v[q];
            cm2 += C0[r][q] * t0[r][q];
            cm3 += r0[r] * t0[r][q] * r0[q];
          }
        Cvc[3][3] = cm1 + cm2 + cm3;
      }
      const double ti = 1. / H;
      const double J3[3][4] = {{-v2x2inv, 0, v[0] * sq(v2x2inv) * 2., 0},
                               {0, -v2x2inv, v[1] * sq(v2x2inv) * 2., 0},
                               {v[0] * v2x2inv * ti, v[1] * v2x2inv * ti,
                                -H * sq(v2x2inv) * 2. - (2. * c + v[2]) * v2x2inv * ti, -ti}};
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          double x = 0;
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            double y = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b)
              y += Cvc[a][b] * J3[q][b];
            x += J3[r][a] * y;
          }
          out.cov[r][q] = x * sq(sInv) + (uvr[r] * uvr[q]) * jq;
        }
    }

    // RiemannFit.h:801-937 lineFit
    template <int N>
    __device__ void riemannLineFit(const double h[3][N], const float ge[6][N], const RCircle& ci, const double ff[4],
                                   double B, RLine& out) {
      constexpr int n = N;
      double theta = -ci.q * rm::atan(ff[3]);
      theta = theta < 0. ? theta + M_PI : theta;
      const double sT = rm::sin(theta), cT = rm::cos(theta);
      double s2[N], covSz11[N];
      const double ox = ci.par[0], oy = ci.par[1];
#pragma unroll
      for (int i = 0; i < n; ++i) {
        const double px = h[0][i] - ox, py = h[1][i] - oy;
        const double cross = (-ox) * py - (-oy) * px;
        const double dot = (-ox) * px + (-oy) * py;
        const double qAtan2 = -ci.q * rm::atan2(cross, dot);
        s2[i] = qAtan2 * ci.par[2];
        const double t0 = -ci.q * ci.par[2] * 1. / (sq(dot) + sq(cross));
        const double J0[6] = {-t0 * ((py + oy) * dot - (px - ox) * cross),
                              t0 * ((px + ox) * dot - (oy - py) * cross),
                              qAtan2,
                              t0 * (oy * dot + ox * cross),
                              t0 * (-ox * dot + oy * cross),
                              0.};
        const double J1[6] = {0., 0., 0., 0., 0., 1.};
        double C[6][6];
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
          for (int b = 0; b < 6; ++b)
            C[a][b] = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b)
            C[a][b] = ci.cov[a][b];
        C[3][3] = ge[0][i];
        C[4][4] = ge[2][i];
        C[5][5] = ge[5][i];
        C[3][4] = C[4][3] = ge[1][i];
        C[3][5] = C[5][3] = ge[3][i];
        C[4][5] = C[5][4] = ge[4][i];
        double M[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            const double* Ja = a == 0 ? J0 : J1;
            const double* Jb = b == 0 ? J0 : J1;
            double x = 0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
              double y = 0;
#pragma unroll
              for (int l = 0; l < 6; ++l)
                y += C[k][l] * Jb[l];
              x += Ja[k] * y;
            }
            M[a][b] = x;
          }
        const double r1[2] = {-cT, sT};
        double x = 0;  // (rot M rot^T)(1, 1)
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int l = 0; l < 2; ++l)
            x += r1[k] * M[k][l] * r1[l];
        covSz11[i] = x;
      }
      // RiemannFit.h:60-110 scatterCovLine (the z' block)
      double covMs[N][N];
      {
        const double pt = fmin(20., ff[2] * B);
        const double p2 = pt * pt * (1. + 1. / sq(ff[3]));
        double sl[N], rl[N], sig2[N];
#pragma unroll
        for (int i = 0; i < n; ++i)
          sl[i] = sqrt(s2[i] * s2[i] + h[2][i] * h[2][i]);
        radLenUniform<N>(sl, rl);
#pragma unroll
        for (int i = 0; i < n; ++i)
          sig2[i] = .000225 / p2 * sq(1. + 0.038 * rm::log(rl[i])) * rl[i];
#pragma unroll
        for (int a = 0; a < n; ++a)
#pragma unroll
          for (int b = a; b < n; ++b) {
            double v = a == b ? covSz11[a] : 0.;
#pragma unroll
            for (int i = 0; i < (a < b ? a : b); ++i)
              v += fabs(sl[a] - sl[i]) * fabs(sl[b] - sl[i]) * sig2[i];
            covMs[a][b] = v;
            covMs[b][a] = v;
          }
      }
      double pr0[N], pr1[N];
#pragma unroll
      for (int i = 0; i < n; ++i) {
        pr0[i] = sT * s2[i] + cT * h[2][i];
        pr1[i] = -cT * s2[i] + sT * h[2][i];
      }
      double Vi[N][N];
      invertSPDr<N>(&covMs[0][0], &Vi[0][0]);
      double AV[2][N];
#pragma unroll
      for (int j = 0; j < n; ++j) {
        double x0 = 0, x1 = 0;
#pragma unroll
        for (int k = 0; k < n; ++k) {
          x0 += 1. * Vi[k][j];
          x1 += pr0[k] * Vi[k][j];
        }
        AV[0][j] = x0;
        AV[1][j] = x1;
      }
      double P[2][2], Pi[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          double x = 0;
#pragma unroll
          for (int j = 0; j < n; ++j)
            x += AV[a][j] * (b == 0 ? 1. : pr0[j]);
          P[a][b] = x;
        }
      invertSPDr<2>(&P[0][0], &Pi[0][0]);
      double sol[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        double x = 0;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          double y = 0;
#pragma unroll
          for (int j = 0; j < n; ++j)
            y += AV[b][j] * pr1[j];
          x += Pi[a][b] * y;
        }
        sol[a] = x;
      }
      const double cf = 1. / (sT - sol[1] * cT);
      const double Jm[2][2] = {{0., cf * cf}, {cf, sol[0] * cT * cf * cf}};
      out.par[0] = cf * (sol[1] * sT + cT);
      out.par[1] = cf * sol[0];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          double x = 0;
#pragma unroll
          for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int l = 0; l < 2; ++l)
              x += Jm[a][k] * Pi[k][l] * Jm[b][l];
          out.cov[a][b] = x;
        }
      double res[N];
#pragma unroll
      for (int i = 0; i < n; ++i)
        res[i] = pr1[i] - (1. * sol[0] + pr0[i] * sol[1]);
      double chi2 = 0;
#pragma unroll
      for (int i = 0; i < n; ++i) {
        double t = 0;
#pragma unroll
        for (int j = 0; j < n; ++j)
          t += Vi[i][j] * res[j];
        chi2 += res[i] * t;
      }
      out.chi2 = chi2;
    }

    // FitUtils.h:159-180 fromCircleToPerigee
    __device__ void toPerigee(RCircle& c) {
      const double t0 = sq(c.par[0]) + sq(c.par[1]);
      const double t1 = sqrt(t0);
      const double pak[3] = {rm::atan2(c.q * c.par[0], -c.q * c.par[1]), c.q * (t1 - c.par[2]), c.q / c.par[2]};
      const double t2 = sq(c.par[0]) * 1. / t0;
      const double t3 = 1. / t1 * c.q;
      const double J[3][3] = {{-c.par[1] * t2 * 1. / sq(c.par[0]), t2 * 1. / c.par[0], 0.},
                              {c.par[0] * t3, c.par[1] * t3, double(-c.q)},
                              {0., 0., -c.q / (c.par[2] * c.par[2])}};
      double r[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          double x = 0;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double y = 0;
#pragma unroll
            for (int l = 0; l < 3; ++l)
              y += c.cov[k][l] * J[b][l];
            x += J[a][k] * y;
          }
          r[a][b] = x;
        }
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        c.par[a] = pak[a];
#pragma unroll
        for (int b = 0; b < 3; ++b)
          c.cov[a][b] = r[a][b];
      }
    }

    template <int N>
    __device__ void riemannTuple(const DevConditions& cond, const Pool& p, const CAPool& ca, uint32_t e, uint32_t tkid,
                                 double B) {
      const size_t tb = ca.tupleBase(e);
      const size_t hb = size_t(e) * kMaxNumClusters;
      const uint32_t M = ca.caps.maxTuples;
      double h[3][N];
      float ge[6][N];
      const uint32_t first = ca.tupleOff[ca.offBase(e) + tkid];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const uint32_t hit = ca.tupleBins[5 * tb + first + i];
        const SOAFrameF& r = cond.cpeDet[p.detInd[hb + hit]].frame;
        // SOAFrame::toGlobal(cxx, 0, cyy) (DataFormats/SOARotation.h:103-111)
        const float cxx = p.xerr[hb + hit], cxy = 0, cyy = p.yerr[hb + hit];
        ge[0][i] = r.R11 * (r.R11 * cxx + r.R21 * cxy) + r.R21 * (r.R11 * cxy + r.R21 * cyy);
        ge[1][i] = r.R11 * (r.R12 * cxx + r.R22 * cxy) + r.R21 * (r.R12 * cxy + r.R22 * cyy);
        ge[2][i] = r.R12 * (r.R12 * cxx + r.R22 * cxy) + r.R22 * (r.R12 * cxy + r.R22 * cyy);
        ge[3][i] = r.R11 * (r.R13 * cxx + r.R23 * cxy) + r.R21 * (r.R13 * cxy + r.R23 * cyy);
        ge[4][i] = r.R12 * (r.R13 * cxx + r.R23 * cxy) + r.R22 * (r.R13 * cxy + r.R23 * cyy);
        ge[5][i] = r.R13 * (r.R13 * cxx + r.R23 * cxy) + r.R23 * (r.R13 * cxy + r.R23 * cyy);
        h[0][i] = p.xg[hb + hit];
        h[1][i] = p.yg[hb + hit];
        h[2][i] = p.zg[hb + hit];
      }
      double ff[4];
      riemannFastFit<N>(h, ff);
      double rad[N];
      double V[2 * N][2 * N];
#pragma unroll
      for (int a = 0; a < 2 * N; ++a)
#pragma unroll
        for (int b = 0; b < 2 * N; ++b)
          V[a][b] = 0;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        rad[i] = sqrt(h[0][i] * h[0][i] + h[1][i] * h[1][i]);
        V[i][i] = ge[0][i];  // FitUtils.h:102-122 loadCovariance2D
        V[i + N][i + N] = ge[2][i];
        V[i][i + N] = ge[1][i];
        V[i + N][i] = ge[1][i];
      }
      RCircle ci;
      riemannCircleFit<N>(h, V, ff, rad, B, ci);
      RLine li;
      riemannLineFit<N>(h, ge, ci, ff, B, li);
      toPerigee(ci);
      // TrajectoryStateSoA::copyFromCircle (CUDADataFormats/TrajectoryStateSoA.h:20-38)
      const float b = 1.f / float(B);
      float* st = ca.state + 5 * tb + tkid;
      st[0 * M] = float(ci.par[0]);
      st[1 * M] = float(ci.par[1]);
      st[2 * M] = b * float(ci.par[2]);
      st[3 * M] = float(li.par[0]);
      st[4 * M] = float(li.par[1]);
      float* cv = ca.cov + 15 * tb + tkid;
      cv[0 * M] = float(ci.cov[0][0]);
      cv[1 * M] = float(ci.cov[0][1]);
      cv[2 * M] = b * float(ci.cov[0][2]);
      cv[3 * M] = 0;
      cv[4 * M] = 0;
      cv[5 * M] = float(ci.cov[1][1]);
      cv[6 * M] = b * float(ci.cov[1][2]);
      cv[7 * M] = 0;
      cv[8 * M] = 0;
      cv[9 * M] = b * b * float(ci.cov[2][2]);
      cv[10 * M] = 0;
      cv[11 * M] = 0;
      cv[12 * M] = float(li.cov[0][0]);
      cv[13 * M] = float(li.cov[0][1]);
      cv[14 * M] = float(li.cov[1][1]);
      // RiemannFitOnGPU.h:173-175
      ca.pt[tb + tkid] = float(B / fabs(ci.par[2]));
      ca.eta[tb + tkid] = asinhf(float(li.par[0]));
      ca.chi2[tb + tkid] = float((double(ci.chi2) + li.chi2) / (2 * N - 5));
    }

    // triplets and quadruplets take different code paths: one launch each, so that
    // a wave runs one of them
    template <int N>
    __global__ void __launch_bounds__(256) riemannKernel(DevConditions cond, Pool p, CAPool ca, double B) {
      const uint32_t e = logicalBlock().y;
      const uint32_t* mo = ca.multOff + size_t(e) * 16;
      const TupleIndex* mb = ca.multBins + size_t(e) * ca.caps.maxTuplesMult;
      // sizes 3 (N = 3), then 4 and 5 on their first 4 hits (N = 4)
      const uint32_t begin = N == 3 ? mo[3] : mo[4], end = N == 3 ? mo[4] : mo[6];
      for (uint32_t k = begin + logicalBlock().x * blockDim.x + threadIdx.x; k < end; k += gridDim.x * blockDim.x)
        riemannTuple<N>(cond, p, ca, e, mb[k], B);
    }
  }  // namespace

  void launchRiemannFit(const CALaunch& l) {
    l.t->b("riemannFit3");
    riemannKernel<3><<<dim3(l.blocks, l.nEvents), 256, 0, l.s>>>(*l.cond, *l.p, *l.ca, l.bField);
    l.t->e();
    l.t->b("riemannFit4");
    riemannKernel<4><<<dim3(l.blocks, l.nEvents), 256, 0, l.s>>>(*l.cond, *l.p, *l.ca, l.bField);
    l.t->e();
  }

}  // namespace pxg

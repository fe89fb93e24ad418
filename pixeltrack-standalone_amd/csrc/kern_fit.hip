// Broken-line helix fit on gfx950 (a18 kernelBLFastFit + a19 kernelBLFit,
// BrokenLineFitOnGPU.h:24-182, BrokenLine.h:46-490): one tuple per lane, in
// double precision, over the entries of the tuple multiplicity container
// (triplets with N = 3, quadruplets and pentuplets on their first 4 hits).
// The small symmetric positive-definite systems ((N + 1) x (N + 1) at most)
// are solved through an LDL^T factorisation held in registers.
//
// Triplets and quadruplets run as two kernels, each with its own register
// footprint, and the double-precision atan2, sin, cos and log are called, not
// inlined: inlined, the unrolled loops interleaved N of them at once and the
// fit needed about 370 registers per lane (one wave per SIMD).  Called, the
// triplet fit takes 146 and the quadruplet fit 168 (three waves per SIMD).
#include <hip/hip_runtime.h>

#include "pxgpu_internal.h"

namespace pxg {

  namespace {
    __device__ __forceinline__ double sq(double a) { return a * a; }

    // the math library's double functions behind a call each (same results)
    __device__ __noinline__ double atan2Call(double y, double x) { return atan2(y, x); }
    __device__ __noinline__ double cosCall(double x) { return cos(x); }
    __device__ __noinline__ double sinCall(double x) { return sin(x); }
    __device__ __noinline__ double logCall(double x) { return log(x); }

    template <int n>
    __device__ void invertSPD(const double* a, double* inv) {
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

    // BrokenLine.h:46-59
    __device__ __forceinline__ double multScatt(double length, double B, double R, double slope) {
      double pt2 = fmin(20., B * R);
      pt2 *= pt2;
      constexpr double XXI_0 = 0.06 / 16.;
      constexpr double fact = 0.7 * (13.6 / 1000.) * (13.6 / 1000.);
      return fact / (pt2 * (1. + sq(slope))) * (fabs(length) * XXI_0) * sq(1. + 0.038 * logCall(fabs(length) * XXI_0));
    }

    struct Rot2 {
      double r00, r01, r10, r11;
    };
    __device__ __forceinline__ Rot2 rotation(double slope) {
      Rot2 r;
      r.r00 = 1. / sqrt(1. + sq(slope));
      r.r01 = slope * r.r00;
      r.r10 = -r.r01;
      r.r11 = r.r00;
      return r;
    }

    __device__ __forceinline__ double cross2(double ax, double ay, double bx, double by) { return ax * by - ay * bx; }

    __device__ __forceinline__ void sandwich3(const double* J, double* cov) {
      double t[9], r[9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          t[i * 3 + j] = J[i * 3 + 0] * cov[0 * 3 + j] + J[i * 3 + 1] * cov[1 * 3 + j] + J[i * 3 + 2] * cov[2 * 3 + j];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          r[i * 3 + j] = t[i * 3 + 0] * J[j * 3 + 0] + t[i * 3 + 1] * J[j * 3 + 1] + t[i * 3 + 2] * J[j * 3 + 2];
#pragma unroll
      for (int k = 0; k < 9; ++k)
        cov[k] = r[k];
    }

    __device__ __forceinline__ void sandwich2(const double* J, double* cov) {
      double t[4], r[4];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          t[i * 2 + j] = J[i * 2 + 0] * cov[0 * 2 + j] + J[i * 2 + 1] * cov[1 * 2 + j];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          r[i * 2 + j] = t[i * 2 + 0] * J[j * 2 + 0] + t[i * 2 + 1] * J[j * 2 + 1];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        cov[k] = r[k];
    }

    struct CircleFit {
      double par[3];
      double cov[9];
      int q;
      float chi2;
    };
    struct LineFit {
      double par[2];
      double cov[4];
      double chi2;
    };

    // BrokenLine.h:85-108
    __device__ void translateKarimaki(CircleFit& c, double x0, double y0) {
      const double p0 = c.par[0], p1 = c.par[1], p2 = c.par[2];
      const double DP = x0 * cosCall(p0) + y0 * sinCall(p0);
      const double DO = x0 * sinCall(p0) - y0 * cosCall(p0) + p1;
      const double uu = 1 + p2 * p1;
      const double C = -p2 * y0 + uu * cosCall(p0);
      const double BB = p2 * x0 + uu * sinCall(p0);
      const double A = 2. * DO + p2 * (sq(DO) + sq(DP));
      const double U = sqrt(1. + p2 * A);
      const double xi = 1. / (sq(BB) + sq(C));
      const double v = 1. + p2 * DO;
      const double lambda = (0.5 * A) / (U * sq(1. + U));
      const double mu = 1. / (U * (1. + U)) + p2 * lambda;
      const double zeta = sq(DO) + sq(DP);
      const double J[9] = {xi * uu * v, -xi * sq(p2) * DP, xi * DP, 2. * mu * uu * DP, 2. * mu * v,
                           mu * zeta - lambda * A, 0, 0, 1.};
      c.par[0] = atan2Call(BB, C);
      c.par[1] = A / (1 + U);
      sandwich3(J, c.cov);
    }

    // BrokenLine.h:173-203
    template <int n>
    __device__ void matrixCu(const double* w, const double* S, const double* VB, double* out) {
      double C[n * n];
#pragma unroll
      for (int k = 0; k < n * n; ++k)
        C[k] = 0;
#pragma unroll
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
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = 0; j < n; ++j)
          out[i * n + j] = C[i * n + j] + C[j * n + i];
    }

    template <int N>
    struct BLData {
      int q;
      double radii[2][N];
      double s[N], S[N], Z[N], VarBeta[N];
    };

    // BrokenLine.h:216-237
    template <int N>
    __device__ void fastFit(const double (&h)[3][N], double* r) {
      constexpr int n = N;
      const double ax = h[0][n / 2] - h[0][0], ay = h[1][n / 2] - h[1][0];
      const double bx = h[0][n - 1] - h[0][n / 2], by = h[1][n - 1] - h[1][n / 2];
      const double cx = h[0][0] - h[0][n - 1], cy = h[1][0] - h[1][n - 1];
      const double a2 = ax * ax + ay * ay, b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
      const double tmp = 0.5 / cross2(cx, cy, ax, ay);
      r[0] = h[0][0] - (ay * c2 + cy * a2) * tmp;
      r[1] = h[1][0] + (ax * c2 + cx * a2) * tmp;
      r[2] = sqrt(a2 * b2 * c2) / (2. * fabs(cross2(bx, by, ax, ay)));
      const double dx = h[0][0] - r[0], dy = h[1][0] - r[1];
      const double ex = h[0][n - 1] - r[0], ey = h[1][n - 1] - r[1];
      r[3] = r[2] * atan2Call(cross2(dx, dy, ex, ey), dx * ex + dy * ey) / (h[2][n - 1] - h[2][0]);
    }

    // BrokenLine.h:119-162
    template <int N>
    __device__ void prepare(const double (&h)[3][N], const double* ff, double B, BLData<N>& d) {
      constexpr int n = N;
      double dx = h[0][1] - h[0][0], dy = h[1][1] - h[1][0];
      double ex = h[0][n - 1] - h[0][n - 2], ey = h[1][n - 1] - h[1][n - 2];
      d.q = cross2(dx, dy, ex, ey) > 0 ? -1 : 1;
      const double slope = -d.q / ff[3];
      const Rot2 R = rotation(slope);
#pragma unroll
      for (int i = 0; i < n; ++i) {
        d.radii[0][i] = h[0][i] - ff[0];
        d.radii[1][i] = h[1][i] - ff[1];
      }
      const double nrm = sqrt(ff[0] * ff[0] + ff[1] * ff[1]);
      ex = -ff[2] * ff[0] / nrm;
      ey = -ff[2] * ff[1] / nrm;
#pragma unroll
      for (int i = 0; i < n; ++i) {
        dx = d.radii[0][i];
        dy = d.radii[1][i];
        d.s[i] = d.q * ff[2] * atan2Call(cross2(dx, dy, ex, ey), dx * ex + dy * ey);
      }
#pragma unroll
      for (int i = 0; i < n; ++i) {
        const double si = d.s[i], zi = h[2][i];
        d.S[i] = R.r00 * si + R.r01 * zi;
        d.Z[i] = R.r10 * si + R.r11 * zi;
      }
      d.VarBeta[0] = d.VarBeta[n - 1] = 0;
#pragma unroll
      for (int i = 1; i < n - 1; ++i)
        d.VarBeta[i] = multScatt(d.S[i + 1] - d.S[i], B, ff[2], slope) + multScatt(d.S[i] - d.S[i - 1], B, ff[2], slope);
    }

    // BrokenLine.h:401-490
    template <int N>
    __device__ void lineFit(const float (&ge)[6][N], const double* ff, double B, const BLData<N>& d, LineFit& L) {
      constexpr int n = N;
      const double slope = -d.q / ff[3];
      const Rot2 R = rotation(slope);
      double w[N], ru[N];
#pragma unroll
      for (int i = 0; i < n; ++i) {
        const double V[3][3] = {{ge[0][i], ge[1][i], ge[3][i]}, {ge[1][i], ge[2][i], ge[4][i]},
                                {ge[3][i], ge[4][i], ge[5][i]}};
        const double tmp = 1. / sqrt(d.radii[0][i] * d.radii[0][i] + d.radii[1][i] * d.radii[1][i]);
        const double J[2][3] = {{d.radii[1][i] * tmp, -d.radii[0][i] * tmp, 0}, {0, 0, 1.}};
        double RJ[2][3];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b)
            RJ[a][b] = (a == 0 ? R.r00 : R.r10) * J[0][b] + (a == 0 ? R.r01 : R.r11) * J[1][b];
        double RJV[3];
#pragma unroll
        for (int b = 0; b < 3; ++b)
          RJV[b] = RJ[1][0] * V[0][b] + RJ[1][1] * V[1][b] + RJ[1][2] * V[2][b];
        double RJVJt[2];
#pragma unroll
        for (int b = 0; b < 2; ++b)
          RJVJt[b] = RJV[0] * J[b][0] + RJV[1] * J[b][1] + RJV[2] * J[b][2];
        const double m11 = RJVJt[0] * R.r10 + RJVJt[1] * R.r11;
        w[i] = 1. / m11;
      }
#pragma unroll
      for (int i = 0; i < n; ++i)
        ru[i] = w[i] * d.Z[i];
      double Cu[N * N], I[N * N];
      matrixCu<N>(w, d.S, d.VarBeta, Cu);
      invertSPD<N>(Cu, I);
      double u[N];
#pragma unroll
      for (int i = 0; i < n; ++i) {
        double v = 0;
#pragma unroll
        for (int k = 0; k < n; ++k)
          v += I[i * n + k] * ru[k];
        u[i] = v;
      }
      L.par[0] = (u[1] - u[0]) / (d.S[1] - d.S[0]);
      L.par[1] = u[0];
      const double idiff = 1. / (d.S[1] - d.S[0]);
      L.cov[0] = (I[0] - 2 * I[1] + I[n + 1]) * sq(idiff) + multScatt(d.S[1] - d.S[0], B, ff[2], slope);
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
#pragma unroll
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
    __device__ void circleFit(const double (&h)[3][N], const float (&ge)[6][N], const double* ff, double B,
                              BLData<N>& d, CircleFit& c) {
      constexpr int n = N;
      c.q = d.q;
      const double slope = -c.q / ff[3];
#pragma unroll
      for (int i = 0; i < n; ++i)
        d.VarBeta[i] *= 1. + sq(slope);
#pragma unroll
      for (int i = 0; i < n; ++i)
        d.Z[i] = sqrt(d.radii[0][i] * d.radii[0][i] + d.radii[1][i] * d.radii[1][i]) - ff[2];
      double w[N];
#pragma unroll
      for (int i = 0; i < n; ++i) {
        const double V00 = ge[0][i], V01 = ge[1][i], V11 = ge[2][i];
        const Rot2 RR = rotation(-d.radii[0][i] / d.radii[1][i]);
        const double a0 = RR.r10 * V00 + RR.r11 * V01;
        const double a1 = RR.r10 * V01 + RR.r11 * V11;
        w[i] = 1. / (a0 * RR.r10 + a1 * RR.r11);
      }
      constexpr int m = N + 1;
      double ru[m];
      ru[n] = 0;
#pragma unroll
      for (int i = 0; i < n; ++i)
        ru[i] = w[i] * d.Z[i];
      double Cn[N * N], C[m * m], I[m * m];
#pragma unroll
      for (int k = 0; k < m * m; ++k)
        C[k] = 0;
      matrixCu<N>(w, d.s, d.VarBeta, Cn);
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = 0; j < n; ++j)
          C[i * m + j] = Cn[i * n + j];
      const double* s = d.s;
      const double* VB = d.VarBeta;
      C[n * m + n] = 0;
#pragma unroll
      for (int i = 0; i < n; ++i) {
        double cin = 0;
        if (i > 0 && i < n - 1)
          cin += -(s[i + 1] - s[i - 1]) * (s[i + 1] - s[i - 1]) / (2. * VB[i] * (s[i + 1] - s[i]) * (s[i] - s[i - 1]));
        if (i > 1)
          cin += (s[i] - s[i - 2]) / (2. * VB[i - 1] * (s[i] - s[i - 1]));
        if (i < n - 2)
          cin += (s[i + 2] - s[i]) / (2. * VB[i + 1] * (s[i + 1] - s[i]));
        C[i * m + n] = C[n * m + i] = cin;
        if (i > 0 && i < n - 1)
          C[n * m + n] += sq(s[i + 1] - s[i - 1]) / (4. * VB[i]);
      }
      invertSPD<m>(C, I);
      double u[m];
#pragma unroll
      for (int i = 0; i < m; ++i) {
        double v = 0;
#pragma unroll
        for (int k = 0; k < m; ++k)
          v += I[i * m + k] * ru[k];
        u[i] = v;
      }
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const double nr = sqrt(d.radii[0][k] * d.radii[0][k] + d.radii[1][k] * d.radii[1][k]);
        d.radii[0][k] /= nr;
        d.radii[1][k] /= nr;
      }
      const double dx = h[0][0] + (-d.Z[0] + u[0]) * d.radii[0][0];
      const double dy = h[1][0] + (-d.Z[0] + u[0]) * d.radii[1][0];
      const double ex = h[0][1] + (-d.Z[1] + u[1]) * d.radii[0][1];
      const double ey = h[1][1] + (-d.Z[1] + u[1]) * d.radii[1][1];
      const double emdx = ex - dx, emdy = ey - dy;
      const double tmp1 = emdx * emdx + emdy * emdy;
      c.par[0] = atan2Call(emdy, emdx);
      c.par[1] = -c.q * (ff[2] - sqrt(sq(ff[2]) - 0.25 * tmp1));
      c.par[2] = c.q * (1. / ff[2] + u[n]);
      const int qHalf = c.q / 2;  // integer division as in the reference: 0 for q = +-1
      const double J[9] = {(d.radii[1][0] * emdx - emdy * d.radii[0][0]) / tmp1,
                           (d.radii[1][1] * emdx - emdy * d.radii[0][1]) / tmp1,
                           0,
                           qHalf * (emdx * d.radii[0][0] + emdy * d.radii[1][0]) / sqrt(sq(2 * ff[2]) - tmp1),
                           qHalf * (emdx * d.radii[0][1] + emdy * d.radii[1][1]) / sqrt(sq(2 * ff[2]) - tmp1),
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
      translateKarimaki(c, 0.5 * emdx, 0.5 * emdy);
      c.cov[0] += (1 + sq(slope)) * multScatt(d.S[1] - d.S[0], B, ff[2], slope);
      translateKarimaki(c, dx, dy);
      c.chi2 = 0;
#pragma unroll
      for (int i = 0; i < n; ++i) {
        c.chi2 += w[i] * sq(d.Z[i] - u[i]);
        if (i > 0 && i < n - 1)
          c.chi2 += sq(u[i - 1] / (s[i] - s[i - 1]) - u[i] * (s[i + 1] - s[i - 1]) / ((s[i + 1] - s[i]) * (s[i] - s[i - 1])) +
                       u[i + 1] / (s[i + 1] - s[i]) + (s[i + 1] - s[i - 1]) * u[n] / 2) /
                    VB[i];
      }
    }

    template <int N>
    __device__ void fitTuple(const DevConditions& cond, const Pool& p, const CAPool& ca, uint32_t e, uint32_t tkid,
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
      fastFit<N>(h, ff);
      BLData<N> d;
      prepare<N>(h, ff, B, d);
      LineFit L;
      lineFit<N>(ge, ff, B, d, L);
      CircleFit Ci;
      circleFit<N>(h, ge, ff, B, d, Ci);
      // TrajectoryStateSoA::copyFromCircle (CUDADataFormats/TrajectoryStateSoA.h:20-38)
      const float b = 1.f / float(B);
      float* st = ca.state + 5 * tb + tkid;
      st[0 * M] = float(Ci.par[0]);
      st[1 * M] = float(Ci.par[1]);
      st[2 * M] = float(Ci.par[2]) * b;
      st[3 * M] = float(L.par[0]);
      st[4 * M] = float(L.par[1]);
      float* cv = ca.cov + 15 * tb + tkid;
      cv[0 * M] = float(Ci.cov[0]);
      cv[1 * M] = float(Ci.cov[1]);
      cv[2 * M] = b * float(Ci.cov[2]);
      cv[3 * M] = 0;
      cv[4 * M] = 0;
      cv[5 * M] = float(Ci.cov[4]);
      cv[6 * M] = b * float(Ci.cov[5]);
      cv[7 * M] = 0;
      cv[8 * M] = 0;
      cv[9 * M] = b * b * float(Ci.cov[8]);
      cv[10 * M] = 0;
      cv[11 * M] = 0;
      cv[12 * M] = float(L.cov[0]);
      cv[13 * M] = float(L.cov[1]);
      cv[14 * M] = float(L.cov[3]);
      ca.pt[tb + tkid] = float(B) / float(fabs(Ci.par[2]));
      ca.eta[tb + tkid] = asinhf(float(L.par[0]));
      ca.chi2[tb + tkid] = float((Ci.chi2 + L.chi2) / (2 * N - 5));
    }

    // the fits of the tuples with N hits: sizes 3, then 4 and 5 (fitted on 4 hits)
    template <int N>
    __global__ void __launch_bounds__(256, 3) fitKernel(DevConditions cond, Pool p, CAPool ca, double B) {
      const uint32_t e = logicalBlock().y;
      const uint32_t* mo = ca.multOff + size_t(e) * 16;
      const TupleIndex* mb = ca.multBins + size_t(e) * ca.caps.maxTuplesMult;
      const uint32_t begin = N == 3 ? mo[3] : mo[4], end = N == 3 ? mo[4] : mo[6];
      for (uint32_t k = begin + logicalBlock().x * blockDim.x + threadIdx.x; k < end; k += gridDim.x * blockDim.x)
        fitTuple<N>(cond, p, ca, e, mb[k], B);
    }
  }  // namespace

  void launchFit(const CALaunch& l) {
    l.t->b("fit");
    fitKernel<3><<<dim3(l.blocks, l.nEvents), 256, 0, l.s>>>(*l.cond, *l.p, *l.ca, l.bField);
    fitKernel<4><<<dim3(l.blocks, l.nEvents), 256, 0, l.s>>>(*l.cond, *l.p, *l.ca, l.bField);
    l.t->e();
  }

}  // namespace pxg

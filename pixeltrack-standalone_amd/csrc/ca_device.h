// Device helpers shared by the ntuplet kernels: the doublet search tables and
// cut parameters of the serial program, the CA geometric cuts, and
// workgroup-wide scans for wave64.
#pragma once

#include <hip/hip_runtime.h>

#include "primitives.h"
#include "pxgpu_internal.h"

namespace pxg {

  // gpuPixelDoublets.h:14-54: the 19 layer pairs of the doublet search with
  // their phi windows (16-bit phi units) and z / r windows of the inner hit
  __constant__ constexpr uint8_t kPairInner[kNumPairs] = {0, 0, 0, 1, 1, 1, 4, 7, 2, 2, 2, 5, 8, 0, 1, 0, 0, 4, 7};
  __constant__ constexpr uint8_t kPairOuter[kNumPairs] = {1, 4, 7, 2, 4, 7, 5, 8, 3, 4, 7, 6, 9, 2, 3, 5, 8, 6, 9};
  __constant__ constexpr int16_t kPairPhiCut[kNumPairs] = {522, 730, 730, 522, 626, 626, 522, 522, 626, 626,
                                                           626, 522, 522, 522, 522, 522, 522, 522, 522};
  __constant__ constexpr float kPairMinZ[kNumPairs] = {-20., 0.,   -30., -22., 10.,  -30., -70., -70., -22., 15.,
                                                       -30,  -70., -70., -20., -22., 0,    -30., -70., -70.};
  __constant__ constexpr float kPairMaxZ[kNumPairs] = {20., 30., 0.,  22., 30., -10., 70., 70., 22., 30.,
                                                       -15., 70., 70., 20., 22., 30.,  0.,  70., 70.};
  __constant__ constexpr float kPairMaxR[kNumPairs] = {20., 9., 9., 20., 7., 7., 5., 5., 20., 6.,
                                                       6.,  5., 5., 20., 20., 9., 9., 9., 9.};
  // pairs whose inner layer is l, in pair order (connect walks them in this order)
  __constant__ constexpr int8_t kPairsFromLayer[10][6] = {
      {0, 1, 2, 13, 15, 16}, {3, 4, 5, 14, -1, -1}, {8, 9, 10, -1, -1, -1}, {-1, -1, -1, -1, -1, -1},
      {6, 17, -1, -1, -1, -1}, {11, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1}, {7, 18, -1, -1, -1, -1},
      {12, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1}};

  // CAHitNtupletGeneratorOnGPU.cc:45-65: CA parameters of the serial program
  constexpr float kCaPtMin = 0.899999976158f;
  constexpr float kCaThetaCutBarrel = 0.00200000009499f;
  constexpr float kCaThetaCutForward = 0.00300000002608f;
  constexpr float kCaHardCurvCut = 0.0328407224959f;
  constexpr float kCaDcaCutInner = 0.15000000596f;
  constexpr float kCaDcaCutOuter = 0.25f;

  __device__ __forceinline__ int phiBinOf(int16_t t) { return (int(t) >> 9) & 127; }

  // GPUCACell.h:175-185
  __device__ __forceinline__ bool alignedRZ(float r1, float z1, float ri, float zi, float ro, float zo, float ptmin,
                                            float thetaCut) {
    const float rd = fabsf(r1 - ro);
    const float d13 = rd * rd + (z1 - zo) * (z1 - zo);
    const float pMin = ptmin * sqrtf(d13);
    const float t = fabsf(z1 * (ri - ro) + zi * (ro - r1) + zo * (r1 - ri));
    return t * pMin <= thetaCut * d13 * rd;
  }

  // CircleEq.h:69-105 (float) and GPUCACell.h:187-206: the circle through the
  // three hits must not bend more than maxCurv and must pass within `cut` of
  // the origin (|dca0| < cut |c|)
  __device__ __forceinline__ bool dcaCut(float x1, float y1, float x2, float y2, float x3, float y3, float cut,
                                         float maxCurv) {
    const bool noflip = fabsf(x3 - x1) < fabsf(y3 - y1);
    const float x1p = noflip ? x1 - x2 : y1 - y2;
    const float y1p = noflip ? y1 - y2 : x1 - x2;
    const float d12 = x1p * x1p + y1p * y1p;
    const float x3p = noflip ? x3 - x2 : y3 - y2;
    const float y3p = noflip ? y3 - y2 : x3 - x2;
    const float d32 = x3p * x3p + y3p * y3p;
    const float num = x1p * y3p - y1p * x3p;
    const float det = d12 * y3p - d32 * y1p;
    const float st2 = d12 * x3p - d32 * x1p;
    const float seq = det * det + st2 * st2;
    float al2 = 1.f / sqrtf(seq);
    const float be2 = -st2 * al2;
    const float ct = 2.f * num * al2;
    al2 *= det;
    const float c = noflip ? ct : -ct;
    const float alpha = noflip ? al2 : -be2;
    const float beta = noflip ? be2 : -al2;
    if (c > maxCurv)
      return false;
    const float x = c * x2 + alpha;
    const float y = c * y2 + beta;
    const float dca0 = sqrtf(x * x + y * y) - 1.f;
    return fabsf(dca0) < cut * fabsf(c);
  }

  // layer starts and the cumulative inner-hit counts of the pairs of event e
  struct PairTable {
    uint32_t layer[11];
    uint32_t cum[kNumPairs];  // cum[p] = inner hits of pairs 0..p
  };

  __device__ __forceinline__ void loadPairTable(const uint32_t* hitsLayerStart, PairTable& t) {
    for (int l = 0; l < 11; ++l)
      t.layer[l] = hitsLayerStart[l];
    uint32_t c = 0;
    for (int p = 0; p < kNumPairs; ++p) {
      c += t.layer[kPairInner[p] + 1] - t.layer[kPairInner[p]];
      t.cum[p] = c;
    }
  }

}  // namespace pxg

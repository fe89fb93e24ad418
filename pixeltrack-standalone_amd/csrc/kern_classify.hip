// Track classification and duplicate cleaning on gfx950: a20 classifyTracks,
// a21 fastDuplicateRemover, a22 hit -> tuple association and tripletCleaner
// (CAHitNtupletGeneratorKernelsImpl.h:151-189, :340-406, :421-534).
//
// The serial fast duplicate remover visits cells in index order and each
// visit depends on the qualities left by earlier visits of cells that share a
// tuple.  Here a cell is processed once every lower-index cell sharing one of
// its tuples has been processed (rounds inside one workgroup per event);
// cells that share no tuple commute, so the result is the serial one.  The
// triplet cleaner and the hit -> tuple fill only mark duplicates from
// quantities that do not change meanwhile, so they run per hit in parallel.
#include <hip/hip_runtime.h>

#include "ca_device.h"

namespace pxg {

  namespace {
    __device__ __forceinline__ uint32_t gridStart() { return logicalBlock().x * blockDim.x + threadIdx.x; }
    __device__ __forceinline__ uint32_t gridStride() { return gridDim.x * blockDim.x; }

    // a20 kernel_classifyTracks
    __global__ void __launch_bounds__(kCaBlock) classifyKernel(CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const uint32_t nT = ca.cnt[e].nTuples;
      const size_t tb = ca.tupleBase(e);
      const uint32_t M = ca.caps.maxTuples;
      const uint32_t* off = ca.tupleOff + ca.offBase(e);
      const float c0 = float(0.68177776), c1 = float(0.74609577), c2 = float(-0.08035491), c3 = float(0.00315399);
      for (uint32_t t = gridStart(); t < nT; t += gridStride()) {
        const uint32_t nh = off[t + 1] - off[t];
        // the serial loop stops at the first empty tuple; only the last stored
        // tuple can be empty (when the container overflowed)
        if (nh == 0 || ca.quality[tb + t] == kQualityDup || nh < 3)
          continue;
        const float* st = ca.state + 5 * tb + t;
        bool isNaN = false;
        for (int i = 0; i < 5; ++i)
          isNaN |= __builtin_isnan(st[i * M]);
        if (isNaN)
          continue;
        const float ptt = ca.pt[tb + t];
        const float pt = fminf(ptt, 10.f);
        const float chi2Cut = 30.f * (c0 + pt * (c1 + pt * (c2 + pt * c3)));
        if (3.f * ca.chi2[tb + t] >= chi2Cut)
          continue;
        const bool quad = nh > 3;
        const float maxTip = quad ? 0.5f : 0.3f, minPt = quad ? 0.3f : 0.5f, maxZip = 12.f;
        if (fabsf(st[1 * M]) < maxTip && ptt > minPt && fabsf(st[4 * M]) < maxZip)
          ca.quality[tb + t] = kQualityLoose;
      }
    }

    // a21 kernel_fastDuplicateRemover, one workgroup per event
    __global__ void __launch_bounds__(kEventBlock) fastDuplicateKernel(CAPool ca) {
      const uint32_t e = blockIdx.x;
      const uint32_t nMulti = ca.cnt[e].nMulti;
      const size_t cb = ca.cellBase(e);
      const size_t tb = ca.tupleBase(e);
      const uint32_t M = ca.caps.maxTuples;
      const uint32_t* multi = ca.multiCells + cb;
      const uint16_t* pool = ca.trPool + size_t(e) * ca.caps.maxCellLists * kCaMaxCellTracks;
      uint8_t* q = ca.quality + tb;
      const float* tip = ca.state + 5 * tb + 1 * size_t(M);
      // bit2 of the cell flags: processed.  The lowest pending cell is always
      // ready, so every round makes progress; nMulti + 1 rounds bound the loop.
      for (uint32_t round = 0; round <= nMulti; ++round) {
        int pending = 0;
        for (uint32_t k = threadIdx.x; k < nMulti; k += kEventBlock) {
          const uint32_t c = multi[k];
          if (ca.cellFlags[cb + c] & 4)
            continue;
          const uint16_t* list = pool + size_t(ca.trSlot[cb + c]) * kCaMaxCellTracks;
          const uint32_t n = ca.trCount[cb + c];
          bool ready = true;
          for (uint32_t a = 0; a < n && ready; ++a) {
            const uint32_t* cells = ca.tupleCells + 4 * tb + 4 * size_t(list[a]);
            for (int k2 = 0; k2 < 4; ++k2) {
              const uint32_t o = cells[k2];
              if (o == 0xffffffffu)
                break;
              if (o < c && ca.trCount[cb + o] >= 2 && !(ca.cellFlags[cb + o] & 4)) {
                ready = false;
                break;
              }
            }
          }
          if (!ready) {
            pending = 1;
            continue;
          }
          float mc = 10000.f;
          uint16_t im = 60000;
          for (uint32_t a = 0; a < n; ++a) {
            const uint16_t it = list[a];
            const float score = fabsf(tip[it]);
            if (q[it] == kQualityLoose && score < mc) {
              mc = score;
              im = it;
            }
          }
          for (uint32_t a = 0; a < n; ++a) {
            const uint16_t it = list[a];
            if (q[it] != kQualityBad && it != im)
              q[it] = kQualityDup;
          }
          ca.cellFlags[cb + c] |= 8;  // done in this round
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < nMulti; k += kEventBlock) {
          const uint32_t c = multi[k];
          const uint8_t f = ca.cellFlags[cb + c];
          if (f & 8)
            ca.cellFlags[cb + c] = uint8_t((f & ~8) | 4);
        }
        if (!__syncthreads_or(pending))
          break;
      }
    }

    // a22 hit -> tuple association of the loose tuples
    __global__ void __launch_bounds__(kCaBlock) hitToTupleCountKernel(CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const uint32_t nT = ca.cnt[e].nTuples;
      const size_t tb = ca.tupleBase(e);
      const uint32_t* off = ca.tupleOff + ca.offBase(e);
      uint32_t* cnt = ca.h2tCount + size_t(e) * kHitStride;
      for (uint32_t t = gridStart(); t < nT; t += gridStride()) {
        if (ca.quality[tb + t] != kQualityLoose)
          continue;
        for (uint32_t k = off[t]; k < off[t + 1]; ++k)
          atomicAdd(&cnt[ca.tupleBins[5 * tb + k]], 1u);
      }
    }

    __global__ void __launch_bounds__(kEventBlock) hitToTupleScanKernel(Pool p, CAPool ca) {
      const uint32_t e = blockIdx.x;
      __shared__ uint32_t lds[kEventBlock / 64];
      const uint32_t nHits = p.hitsLayerStart[size_t(e) * kLayerStartStride + 10];
      uint32_t* cnt = ca.h2tCount + size_t(e) * kHitStride;
      uint32_t* off = ca.h2tOff + size_t(e) * kHitStride;
      const uint32_t cap = ca.caps.maxHitToTuple;
      uint32_t carry = 0;
      for (uint32_t base = 0; base < nHits; base += kEventBlock) {
        const uint32_t h = base + threadIdx.x;
        const uint32_t v = h < nHits ? cnt[h] : 0;
        uint32_t tot;
        const uint32_t ex = wgExclusiveScan<kEventBlock>(v, tot, lds);
        if (h < nHits) {
          off[h] = min(carry + ex, cap);
          cnt[h] = 0;
        }
        carry += tot;
      }
      if (threadIdx.x == 0) {
        off[nHits] = min(carry, cap);
        if (carry > cap)
          atomicOr(&ca.cnt[e].status, kCaHitToTupleOverflow);
      }
    }

    __global__ void __launch_bounds__(kCaBlock) hitToTupleFillKernel(CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const uint32_t nT = ca.cnt[e].nTuples;
      const size_t tb = ca.tupleBase(e);
      const uint32_t* off = ca.tupleOff + ca.offBase(e);
      uint32_t* cur = ca.h2tCount + size_t(e) * kHitStride;
      const uint32_t* hoff = ca.h2tOff + size_t(e) * kHitStride;
      uint16_t* bins = ca.h2tBins + size_t(e) * ca.caps.maxHitToTuple;
      for (uint32_t t = gridStart(); t < nT; t += gridStride()) {
        if (ca.quality[tb + t] != kQualityLoose)
          continue;
        for (uint32_t k = off[t]; k < off[t + 1]; ++k) {
          const uint32_t h = ca.tupleBins[5 * tb + k];
          const uint32_t pos = hoff[h] + atomicAdd(&cur[h], 1u);
          if (pos < hoff[h + 1])
            bins[pos] = uint16_t(t);
        }
      }
    }

    // sorts each hit's tuples (descending, the serial fill order) and runs the
    // triplet cleaner (CAHitNtupletGeneratorKernelsImpl.h:479-534) on them
    __global__ void __launch_bounds__(kCaBlock) tripletCleanerKernel(Pool p, CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const uint32_t nHits = p.hitsLayerStart[size_t(e) * kLayerStartStride + 10];
      const size_t tb = ca.tupleBase(e);
      const uint32_t M = ca.caps.maxTuples;
      const uint32_t* off = ca.tupleOff + ca.offBase(e);
      const uint32_t* hoff = ca.h2tOff + size_t(e) * kHitStride;
      uint16_t* bins = ca.h2tBins + size_t(e) * ca.caps.maxHitToTuple;
      uint8_t* q = ca.quality + tb;
      const float* tip = ca.state + 5 * tb + 1 * size_t(M);
      for (uint32_t h = gridStart(); h < nHits; h += gridStride()) {
        const uint32_t b = hoff[h], n = hoff[h + 1] - b;
        for (uint32_t k = 1; k < n; ++k) {
          const uint16_t v = bins[b + k];
          uint32_t m = k;
          while (m > 0 && bins[b + m - 1] < v) {
            bins[b + m] = bins[b + m - 1];
            --m;
          }
          bins[b + m] = v;
        }
        if (n < 2)
          continue;
        uint32_t maxNh = 0;
        for (uint32_t k = 0; k < n; ++k) {
          const uint16_t it = bins[b + k];
          maxNh = max(maxNh, off[it + 1] - off[it]);
        }
        for (uint32_t k = 0; k < n; ++k) {
          const uint16_t it = bins[b + k];
          if (off[it + 1] - off[it] != maxNh)
            q[it] = kQualityDup;
        }
        if (maxNh > 3)
          continue;
        float mc = 10000.f;
        uint16_t im = 60000;
        for (uint32_t k = 0; k < n; ++k) {
          const uint16_t it = bins[b + k];
          if (q[it] != kQualityBad && fabsf(tip[it]) < mc) {
            mc = fabsf(tip[it]);
            im = it;
          }
        }
        for (uint32_t k = 0; k < n; ++k) {
          const uint16_t it = bins[b + k];
          if (q[it] != kQualityBad && it != im)
            q[it] = kQualityDup;
        }
      }
    }
  }  // namespace

  void launchClassify(const CALaunch& l) {
    l.t->b("classify");
    classifyKernel<<<dim3(l.blocks, l.nEvents), kCaBlock, 0, l.s>>>(*l.ca);
    l.t->e();
    l.t->b("fastDuplicate");
    fastDuplicateKernel<<<l.nEvents, kEventBlock, 0, l.s>>>(*l.ca);
    l.t->e();
  }

  void launchHitToTuple(const CALaunch& l) {
    const dim3 g(l.blocks, l.nEvents);
    l.t->b("hitToTupleCount");
    hitToTupleCountKernel<<<g, kCaBlock, 0, l.s>>>(*l.ca);
    l.t->e();
    l.t->b("hitToTupleScan");
    hitToTupleScanKernel<<<l.nEvents, kEventBlock, 0, l.s>>>(*l.p, *l.ca);
    l.t->e();
    l.t->b("hitToTupleFill");
    hitToTupleFillKernel<<<g, kCaBlock, 0, l.s>>>(*l.ca);
    l.t->e();
    l.t->b("tripletCleaner");
    tripletCleanerKernel<<<g, kCaBlock, 0, l.s>>>(*l.p, *l.ca);
    l.t->e();
  }

}  // namespace pxg

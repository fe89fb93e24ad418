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

    // kernel_fishboneCleaner (CAHitNtupletGeneratorKernelsImpl.h:113-125): the tuples of
    // a cell the late fishbone killed are bad.  Visits the cells that are in tuples
    // through their track list slots.
    __global__ void __launch_bounds__(kCaBlock) fishboneCleanerKernel(CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const size_t cb = ca.cellBase(e);
      const size_t tb = ca.tupleBase(e);
      const TupleIndex* pool = ca.trPool + size_t(e) * ca.caps.maxCellLists * kCaMaxCellTracks;
      const uint32_t* trCells = ca.dfsT + cb;
      const uint32_t nLists = min(ca.cnt[e].trLists, ca.caps.maxCellLists - 1);
      for (uint32_t slot = gridStart(); slot < nLists; slot += gridStride()) {
        const uint32_t c = trCells[slot];
        if (!(ca.cellFlags[cb + c] & 2))
          continue;
        const TupleIndex* list = pool + size_t(slot) * kCaMaxCellTracks;
        const uint32_t n = ca.trCount[cb + c];
        for (uint32_t k = 0; k < n; ++k)
          ca.quality[tb + list[k]] = kQualityBad;
      }
    }

    // a21 kernel_fastDuplicateRemover.  The serial loop visits the cells with two
    // or more tuples in ascending order, keeps the loose tuple with the smallest
    // |tip| of each and marks the other (not bad) tuples duplicates.  A cell's
    // outcome depends only on the lower cells sharing one of its tuples (its
    // predecessors), so cells run in rounds: a cell is processed once all its
    // predecessors have been, and the cells of one round share no tuple.

    // the predecessors of cell c: lower cells with two or more tuples in its tuples
    template <typename F>
    __device__ __forceinline__ void forPreds(const CAPool& ca, size_t cb, size_t tb, const TupleIndex* pool, uint32_t c,
                                             F&& f) {
      const TupleIndex* list = pool + size_t(ca.trSlot[cb + c]) * kCaMaxCellTracks;
      const uint32_t n = ca.trCount[cb + c];
      for (uint32_t a = 0; a < n; ++a) {
        const uint32_t* cells = ca.tupleCells + 4 * tb + 4 * size_t(list[a]);
        for (int k2 = 0; k2 < 4; ++k2) {
          const uint32_t o = cells[k2];
          if (o == 0xffffffffu)
            break;
          if (o < c && ca.trCount[cb + o] >= 2 && f(o))
            return;
        }
      }
    }

    // keep the loose tuple of cell c with the smallest |tip|, the others are duplicates
    __device__ __forceinline__ void dupProcess(const CAPool& ca, size_t cb, const TupleIndex* pool, uint8_t* q,
                                               const float* tip, uint32_t c) {
      const TupleIndex* list = pool + size_t(ca.trSlot[cb + c]) * kCaMaxCellTracks;
      const uint32_t n = ca.trCount[cb + c];
      float mc = 10000.f;
      TupleIndex im = kNoTuple;
      for (uint32_t a = 0; a < n; ++a) {
        const TupleIndex it = list[a];
        const float score = fabsf(tip[it]);
        if (q[it] == kQualityLoose && score < mc) {
          mc = score;
          im = it;
        }
      }
      for (uint32_t a = 0; a < n; ++a) {
        const TupleIndex it = list[a];
        if (q[it] != kQualityBad && it != im)
          q[it] = kQualityDup;
      }
    }

    // the first kDupPredCap predecessors of every cell with two or more tuples
    // (all slots written, unused ones ~0), on the whole grid
    __global__ void __launch_bounds__(kCaBlock) dupPredsKernel(CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const uint32_t nMulti = ca.cnt[e].nMulti;
      const size_t cb = ca.cellBase(e);
      const size_t tb = ca.tupleBase(e);
      const size_t lb = size_t(e) * ca.caps.maxCellLists;
      const uint32_t* multi = ca.multiCells + cb;
      const TupleIndex* pool = ca.trPool + lb * kCaMaxCellTracks;
      uint32_t* preds = ca.dupPreds + lb * kDupPredCap;
      for (uint32_t k = gridStart(); k < nMulti; k += gridStride()) {
        uint32_t m = 0;
        uint32_t keep[kDupPredCap];
#pragma unroll
        for (uint32_t j = 0; j < kDupPredCap; ++j)
          keep[j] = 0xffffffffu;
        bool over = false;
        forPreds(ca, cb, tb, pool, multi[k], [&](uint32_t o) {
          bool seen = false;
#pragma unroll
          for (uint32_t j = 0; j < kDupPredCap; ++j)
            seen |= keep[j] == o;
          if (seen)
            return false;
          if (m == kDupPredCap) {
            over = true;  // more distinct predecessors than kept: walked when needed
            return true;
          }
#pragma unroll
          for (uint32_t j = 0; j < kDupPredCap; ++j)
            if (j == m)
              keep[j] = o;
          ++m;
          return false;
        });
#pragma unroll
        for (uint32_t j = 0; j < kDupPredCap; ++j)
          preds[size_t(k) * kDupPredCap + j] = keep[j];
        ca.dupCount[lb + k] = over ? uint8_t(255) : uint8_t(m);
      }
    }

    constexpr uint32_t kDupBitmapCells = kCapsReference.maxDoublets;  // done bits of the cells in LDS

    __global__ void __launch_bounds__(kEventBlock) fastDuplicateKernel(CAPool ca) {
      const uint32_t e = blockIdx.x;
      const uint32_t nMulti = ca.cnt[e].nMulti;
      const uint32_t nCells = ca.cnt[e].nCells;
      const size_t cb = ca.cellBase(e);
      const size_t tb = ca.tupleBase(e);
      const size_t lb = size_t(e) * ca.caps.maxCellLists;
      const uint32_t M = ca.caps.maxTuples;
      const uint32_t* multi = ca.multiCells + cb;
      const TupleIndex* pool = ca.trPool + lb * kCaMaxCellTracks;
      const uint32_t* preds = ca.dupPreds + lb * kDupPredCap;
      const uint8_t* count = ca.dupCount + lb;
      uint8_t* q = ca.quality + tb;
      const float* tip = ca.state + 5 * tb + 1 * size_t(M);

      // ---- done bits of all cells in LDS; each thread keeps its (at most kDupPer)
      // cells with their predecessors from dupPredsKernel in registers
      constexpr uint32_t kDupPer = 6;
      __shared__ uint32_t sDone[kDupBitmapCells / 32];
      bool lds = nCells <= kDupBitmapCells && nMulti <= kDupPer * kEventBlock;
      uint32_t myC[kDupPer], myN[kDupPer], myP[kDupPer][kDupPredCap];
      if (lds) {
#pragma unroll
        for (uint32_t i = 0; i < kDupPer; ++i) {
          const uint32_t k = threadIdx.x + i * kEventBlock;
          myN[i] = 0;
          myC[i] = 0;
          if (k < nMulti) {
            myC[i] = multi[k];
            myN[i] = count[k];
#pragma unroll
            for (uint32_t j = 0; j < kDupPredCap; ++j)
              myP[i][j] = preds[size_t(k) * kDupPredCap + j];
          }
        }
        for (uint32_t w = threadIdx.x; w < (nCells + 31) / 32; w += kEventBlock)
          sDone[w] = 0;
        __syncthreads();
      }
      if (lds) {
        auto done = [&](uint32_t c) { return (sDone[c >> 5] >> (c & 31)) & 1; };
        uint32_t mine = 0;  // bit i: cell i of this thread still pending
#pragma unroll
        for (uint32_t i = 0; i < kDupPer; ++i)
          mine |= uint32_t(threadIdx.x + i * kEventBlock < nMulti) << i;
        // the lowest pending cell is always ready: nMulti + 1 rounds bound the loop
        for (uint32_t round = 0; round <= nMulti; ++round) {
          uint32_t now = 0;  // processed in this round
#pragma unroll
          for (uint32_t i = 0; i < kDupPer; ++i) {
            if (!((mine >> i) & 1))
              continue;
            bool ready = true;
            if (myN[i] <= kDupPredCap) {
#pragma unroll
              for (uint32_t j = 0; j < kDupPredCap; ++j)
                ready = ready && (j >= myN[i] || done(myP[i][j]));
            } else {  // too many predecessors to keep: walk them
              forPreds(ca, cb, tb, pool, myC[i], [&](uint32_t o) {
                ready = done(o);
                return !ready;
              });
            }
            if (!ready)
              continue;
            dupProcess(ca, cb, pool, q, tip, myC[i]);
            now |= 1u << i;
          }
          __syncthreads();
#pragma unroll
          for (uint32_t i = 0; i < kDupPer; ++i)
            if ((now >> i) & 1)
              atomicOr(&sDone[myC[i] >> 5], 1u << (myC[i] & 31));
          mine &= ~now;
          if (!__syncthreads_or(mine != 0))
            break;
        }
        return;
      }

      // ---- otherwise the same rounds over per-cell flags in memory (bit2
      // processed, bit3 processed in this round) and the full predecessor walk
      for (uint32_t round = 0; round <= nMulti; ++round) {
        int pending = 0;
        for (uint32_t k = threadIdx.x; k < nMulti; k += kEventBlock) {
          const uint32_t c = multi[k];
          if (ca.cellFlags[cb + c] & 4)
            continue;
          bool ready = true;
          forPreds(ca, cb, tb, pool, c, [&](uint32_t o) {
            ready = (ca.cellFlags[cb + o] & 4) != 0;
            return !ready;
          });
          if (!ready) {
            pending = 1;
            continue;
          }
          dupProcess(ca, cb, pool, q, tip, c);
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

    // doStats counters (CAHitNtupletGeneratorKernels.h:14-26): kernel_checkOverflows
    // per-event and per-cell counts (CAHitNtupletGeneratorKernelsImpl.h:30-104; it runs
    // while the hit -> tuple map is still empty, so every cell is a zero-track cell),
    // kernel_doStatsForTracks (:410-421) and kernel_doStatsForHitInTracks (:468-478).
    // Counters accumulate over the stream's events; each wave adds its partial sums.
    __global__ void __launch_bounds__(kCaBlock) statsKernel(Pool p, CAPool ca, unsigned long long* counters) {
      const uint32_t e = logicalBlock().y;
      const CACounts cnt = ca.cnt[e];
      const uint32_t nHits = p.hitsLayerStart[size_t(e) * kLayerStartStride + 10];
      const size_t cb = ca.cellBase(e);
      const size_t tb = ca.tupleBase(e);
      const uint8_t* used = ca.cellUsed + cb;
      const uint32_t* toff = ca.tupleOff + ca.offBase(e);
      const uint32_t* hoff = ca.h2tOff + size_t(e) * kHitStride;
      uint32_t good = 0, usedHits = 0, dupHits = 0, killed = 0, empty = 0;
      for (uint32_t c = gridStart(); c < cnt.nCells; c += gridStride()) {
        killed += (ca.cellFlags[cb + c] & 2) != 0;
        empty += used[c] == 0;
      }
      // the serial loop stops at the first empty tuple: only the last stored one can be
      for (uint32_t t = gridStart(); t < cnt.nTuples; t += gridStride())
        good += toff[t + 1] > toff[t] && ca.quality[tb + t] == kQualityLoose;
      for (uint32_t h = gridStart(); h < nHits; h += gridStride()) {
        const uint32_t n = hoff[h + 1] - hoff[h];
        usedHits += n > 0;
        dupHits += n > 1;
      }
      const uint32_t v[5] = {good, usedHits, dupHits, killed, empty};
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        uint32_t x = v[k];
        for (int o = 32; o > 0; o >>= 1)
          x += __shfl_xor(x, o);
        if ((threadIdx.x & 63) == 0 && x)
          atomicAdd(&counters[5 + k], (unsigned long long)x);
      }
      if (logicalBlock().x == 0 && threadIdx.x == 0) {
        atomicAdd(&counters[0], 1ull);
        atomicAdd(&counters[1], (unsigned long long)nHits);
        atomicAdd(&counters[2], (unsigned long long)cnt.nCells);
        atomicAdd(&counters[3], (unsigned long long)cnt.apcM);
        atomicAdd(&counters[4], (unsigned long long)cnt.nMult);
        atomicAdd(&counters[10], (unsigned long long)cnt.nCells);
      }
    }

    // a22 hit -> tuple association of the loose tuples (HitToTuple, CAConstants.h:65), a
    // one-to-many container over the event's hits with the reference's capacity, built in
    // LDS: workgroup (s, e) owns the hits [s * kH2tSlice, (s + 1) * kH2tSlice) of event e
    // and passes twice over the event's loose tuples.  The first pass counts its hits'
    // tuples with LDS atomics and the entries of lower hits in a register, so that the
    // offsets (clamped to the capacity, as the container's) need no other workgroup; the
    // second writes each entry at its hit's offset plus the rank an LDS atomic gives (the
    // triplet cleaner sorts each hit's list).  An entry past the capacity is dropped.
    constexpr uint32_t kH2tSlice = 16384;  // hits per workgroup: 64 KB of counts

    __global__ void __launch_bounds__(kEventBlock) hitToTupleBuildKernel(Pool p, CAPool ca) {
      const uint32_t s = blockIdx.x, e = blockIdx.y;
      __shared__ uint32_t cnt[kH2tSlice];
      __shared__ uint32_t lds[kEventBlock / 64];
      const uint32_t nHits = p.hitsLayerStart[size_t(e) * kLayerStartStride + 10];
      const uint32_t h0 = s * kH2tSlice;
      if (h0 >= nHits && s > 0)
        return;  // uniform per workgroup; workgroup 0 always writes the end offset
      const uint32_t n = h0 < nHits ? min(kH2tSlice, nHits - h0) : 0;
      const uint32_t nT = ca.cnt[e].nTuples;
      const size_t tb = ca.tupleBase(e);
      const uint32_t* toff = ca.tupleOff + ca.offBase(e);
      const uint16_t* tbins = ca.tupleBins + 5 * tb;
      const uint8_t* q = ca.quality + tb;
      const uint32_t cap = ca.caps.maxHitToTuple;
      uint32_t* hoff = ca.h2tOff + size_t(e) * kHitStride;
      TupleIndex* bins = ca.h2tBins + size_t(e) * cap;
      for (uint32_t i = threadIdx.x; i < n; i += kEventBlock)
        cnt[i] = 0;
      __syncthreads();
      uint32_t below = 0;
      for (uint32_t t = threadIdx.x; t < nT; t += kEventBlock) {
        if (q[t] != kQualityLoose)
          continue;
        for (uint32_t k = toff[t]; k < toff[t + 1]; ++k) {
          const uint32_t h = tbins[k];
          if (h < h0)
            ++below;
          else if (h - h0 < n)
            atomicAdd(&cnt[h - h0], 1u);
        }
      }
      uint32_t base;
      wgExclusiveScan<kEventBlock>(below, base, lds);  // base: the entries of the lower hits
      // offsets: each thread takes a run of consecutive hits, the runs in order
      const uint32_t per = (n + kEventBlock - 1) / kEventBlock;
      const uint32_t i0 = min(threadIdx.x * per, n), i1 = min(i0 + per, n);
      uint32_t mine = 0;
      for (uint32_t i = i0; i < i1; ++i)
        mine += cnt[i];
      uint32_t total;
      uint32_t run = base + wgExclusiveScan<kEventBlock>(mine, total, lds);
      for (uint32_t i = i0; i < i1; ++i) {
        hoff[h0 + i] = min(run, cap);
        run += cnt[i];
        cnt[i] = 0;  // the counts become ranks
      }
      if (h0 + n == nHits && threadIdx.x == 0) {
        hoff[nHits] = min(base + total, cap);
        if (base + total > cap)
          atomicOr(&ca.cnt[e].status, kCaHitToTupleOverflow);
      }
      __syncthreads();
      for (uint32_t t = threadIdx.x; t < nT; t += kEventBlock) {
        if (q[t] != kQualityLoose)
          continue;
        for (uint32_t k = toff[t]; k < toff[t + 1]; ++k) {
          const uint32_t h = tbins[k];
          if (h - h0 >= n)
            continue;
          // hoff[h] is the unclamped offset unless it reached the capacity, where every
          // entry of h is dropped: the entry is kept iff its place is below the capacity
          const uint32_t pos = hoff[h] + atomicAdd(&cnt[h - h0], 1u);
          if (pos < cap)
            bins[pos] = TupleIndex(t);
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
      TupleIndex* bins = ca.h2tBins + size_t(e) * ca.caps.maxHitToTuple;
      uint8_t* q = ca.quality + tb;
      const float* tip = ca.state + 5 * tb + 1 * size_t(M);
      for (uint32_t h = gridStart(); h < nHits; h += gridStride()) {
        const uint32_t b = hoff[h], n = hoff[h + 1] - b;
        for (uint32_t k = 1; k < n; ++k) {
          const TupleIndex v = bins[b + k];
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
          const TupleIndex it = bins[b + k];
          maxNh = max(maxNh, off[it + 1] - off[it]);
        }
        for (uint32_t k = 0; k < n; ++k) {
          const TupleIndex it = bins[b + k];
          if (off[it + 1] - off[it] != maxNh)
            q[it] = kQualityDup;
        }
        if (maxNh > 3)
          continue;
        float mc = 10000.f;
        TupleIndex im = kNoTuple;
        for (uint32_t k = 0; k < n; ++k) {
          const TupleIndex it = bins[b + k];
          if (q[it] != kQualityBad && fabsf(tip[it]) < mc) {
            mc = fabsf(tip[it]);
            im = it;
          }
        }
        for (uint32_t k = 0; k < n; ++k) {
          const TupleIndex it = bins[b + k];
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
    if (l.lateFishbone) {
      l.t->b("fishboneCleaner");
      fishboneCleanerKernel<<<dim3(l.blocks, l.nEvents), kCaBlock, 0, l.s>>>(*l.ca);
      l.t->e();
    }
    l.t->b("dupPreds");
    dupPredsKernel<<<dim3(l.blocks, l.nEvents), kCaBlock, 0, l.s>>>(*l.ca);
    l.t->e();
    l.t->b("fastDuplicate");
    fastDuplicateKernel<<<l.nEvents, kEventBlock, 0, l.s>>>(*l.ca);
    l.t->e();
  }

  void launchHitToTuple(const CALaunch& l) {
    const dim3 g(l.blocks, l.nEvents);
    l.t->b("hitToTupleFill");
    hitToTupleBuildKernel<<<dim3((kMaxNumClusters + kH2tSlice - 1) / kH2tSlice, l.nEvents), kEventBlock, 0, l.s>>>(
        *l.p, *l.ca);
    l.t->e();
    l.t->b("tripletCleaner");
    tripletCleanerKernel<<<g, kCaBlock, 0, l.s>>>(*l.p, *l.ca);
    l.t->e();
    if (l.counters) {
      l.t->b("stats");
      statsKernel<<<g, kCaBlock, 0, l.s>>>(*l.p, *l.ca, l.counters);
      l.t->e();
    }
  }

}  // namespace pxg

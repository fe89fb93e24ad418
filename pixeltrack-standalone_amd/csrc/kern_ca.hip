// Ntuplet stage on gfx950: doublets, outer-hit cell lists, CA connection,
// fishbone, depth-first ntuplet search, per-cell track lists and the early
// duplicate remover, tuple multiplicity and detector indices.
//
// Every product is the serial program's, element by element: where the serial
// code appends to a container in loop order, the kernels here count first,
// scan, and write each element at its serial position (doublets, tuples), or
// append with atomics and then sort each short list (outer-hit lists, track
// lists); kills and flags are written by exactly one thread.  Grid-stride
// kernels run on a (workgroups per event, events) grid, taken in the
// XCD-aware logical order (logicalBlock), and stride over that event's items
// with gridDim.x workgroups; per-event scans run one workgroup per event.
#include <hip/hip_runtime.h>

#include "ca_device.h"

namespace pxg {

  namespace {
    struct HitView {
      const float *xg, *yg, *zg, *rg;
      const int16_t *iphi, *ysize;
      const uint16_t* detInd;
      const uint32_t* layerStart;
      const uint32_t* phiOff;
      const uint16_t* phiBins;
      const float4* pos;
      const HitAttr* attr;
      const PhiCand* candPhi;   // hit phiBins[k] at k
    };

    __device__ __forceinline__ HitView hitsOf(const Pool& p, uint32_t e) {
      const size_t h = size_t(e) * kMaxNumClusters;
      return {p.xg + h,
              p.yg + h,
              p.zg + h,
              p.rg + h,
              p.iphi + h,
              p.ysize + h,
              p.detInd + h,
              p.hitsLayerStart + size_t(e) * kLayerStartStride,
              p.phiOff + size_t(e) * kPhiOffStride,
              p.phiBins + h,
              p.pos + h,
              p.attr + h,
              p.candPhi + h};
    }

    __device__ __forceinline__ uint32_t gridStart() { return logicalBlock().x * blockDim.x + threadIdx.x; }
    __device__ __forceinline__ uint32_t gridStride() { return gridDim.x * blockDim.x; }

    // wave index within the event's workgroups, uniform over the wave
    __device__ __forceinline__ uint32_t waveStart() {
      return __builtin_amdgcn_readfirstlane(logicalBlock().x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    }
    __device__ __forceinline__ uint32_t waveStride() { return gridDim.x * (blockDim.x >> 6); }

    // ---- a11/a12 doublets: gpuPixelDoubletsAlgos.h:25-238 ----------------------------
    // One wave per inner slot j = (layer pair, inner hit); the lanes test the
    // outer-layer hits of the phi window in parallel.  The window's hits are
    // visited in the serial order (phi bins from the low edge upwards, each
    // bin in histogram order), which is one or (across phi = pi) two
    // contiguous ranges of the histogram contents; ballots compact the
    // accepted ones in that order.
    struct SlotScan {
      bool live;            // the inner hit passes its own cuts
      int pairId, inner, outer;
      uint32_t i;           // inner hit
      uint32_t s1, n1, s2, n;  // candidate ranges: [s1, s1+n1) then [s2, s2+n-n1)
      float mez, mer;
      int16_t mes, mep, iphicut;
      float maxr;
    };

    __device__ __forceinline__ SlotScan slotOf(const HitView& H, const PairTable& T, uint32_t j) {
      SlotScan q{};
      int pairId = 0;
      while (j >= T.cum[pairId])
        ++pairId;
      q.pairId = pairId;
      q.inner = kPairInner[pairId];
      q.outer = kPairOuter[pairId];
      q.i = j - (pairId ? T.cum[pairId - 1] : 0) + T.layer[q.inner];
      q.live = false;
      const HitAttr a = H.attr[q.i];
      if (a.detInd > 2000)
        return q;
      const float4 pi = H.pos[q.i];
      q.mez = pi.z;
      if (q.mez < kPairMinZ[pairId] || q.mez > kPairMaxZ[pairId])
        return q;
      constexpr int minYsizeB1 = 36, minYsizeB2 = 28;
      q.mes = a.ysize;  // ideal conditions: every ladder counts as an outer ladder
      if (q.inner == 0 && q.outer > 3 && q.mes > 0 && q.mes < minYsizeB1)
        return q;
      if (q.inner == 1 && q.outer > 3 && q.mes > 0 && q.mes < minYsizeB2)
        return q;
      q.live = true;
      q.mep = a.iphi;
      q.mer = pi.w;
      q.iphicut = kPairPhiCut[pairId];
      q.maxr = kPairMaxR[pairId];
      const uint32_t hoff = kPhiBins * q.outer;
      const int kl = phiBinOf(int16_t(q.mep - q.iphicut));
      const int kh = phiBinOf(int16_t(q.mep + q.iphicut));
      if (kl <= kh) {
        q.s1 = H.phiOff[hoff + kl];
        q.n1 = H.phiOff[hoff + kh + 1] - q.s1;
        q.s2 = 0;
        q.n = q.n1;
      } else {  // the window wraps through phi = pi
        q.s1 = H.phiOff[hoff + kl];
        q.n1 = H.phiOff[hoff + kPhiBins] - q.s1;
        q.s2 = H.phiOff[hoff];
        q.n = q.n1 + (H.phiOff[hoff + kh + 1] - q.s2);
      }
      return q;
    }

    // the cuts of one (inner, outer) hit pair: gpuPixelDoubletsAlgos.h:141-215; the
    // outer hit is the one at position k of the phi histogram's contents
    __device__ __forceinline__ bool acceptDoublet(const HitView& H, const SlotScan& q, uint32_t k) {
      constexpr int maxDYsize12 = 28, maxDYsize = 20, maxDYPred = 20;
      constexpr float dzdrFact = 8 * 0.0285 / 0.015;
      constexpr float z0cut = 12.f;
      constexpr float minRadius = 0.5f * 87.78f;
      constexpr float minRadius2T4 = 4.f * minRadius * minRadius;
      constexpr float i2p = 3.14159265358979323846 / 32768;
      const PhiCand a = H.candPhi[k];
      if (a.detInd > 2000)
        return false;
      const float zo = a.z, ro = a.r;
      const float dr = ro - q.mer;
      if (dr > q.maxr || dr < 0 || fabsf(q.mez * ro - q.mer * zo) > z0cut * dr)
        return false;
      const int16_t mop = a.iphi;
      const int a1 = abs(int(int16_t(mop - q.mep))), a2 = abs(int(int16_t(q.mep - mop)));
      const uint16_t idphi = uint16_t(a1 < a2 ? a1 : a2);
      if (idphi > q.iphicut)
        return false;
      const int16_t so = a.ysize;
      const int dy = q.inner == 0 ? maxDYsize12 : maxDYsize;
      const bool zsize = q.outer < 4
                             ? (q.mes > 0 && so > 0 && abs(so - q.mes) > dy)
                             : (q.inner < 4) && q.mes > 0 &&
                                   abs(q.mes - int(fabsf((q.mez - zo) / (q.mer - ro)) * dzdrFact + 0.5f)) > maxDYPred;
      if (zsize)
        return false;
      const float dphi = float(int16_t(idphi)) * i2p;
      return !(dphi * dphi * (minRadius2T4 - q.mer * ro) > (ro - q.mer) * (ro - q.mer));
    }

    // position in the phi histogram's contents of the slot's k-th candidate
    __device__ __forceinline__ uint32_t windowPos(const SlotScan& q, uint32_t k) {
      return k < q.n1 ? q.s1 + k : q.s2 + (k - q.n1);
    }
    __device__ __forceinline__ uint32_t candidate(const HitView& H, const SlotScan& q, uint32_t k) {
      return H.phiBins[windowPos(q, k)];
    }

    // position of the r-th set bit of w (r < popcount(w))
    __device__ __forceinline__ uint32_t nthBit(uint64_t w, uint32_t r) {
      uint32_t pos = 0;
#pragma unroll
      for (int width = 32; width >= 1; width >>= 1) {
        const uint32_t c = __popcll(w & ((1ull << width) - 1));
        if (r >= c) {
          r -= c;
          w >>= width;
          pos += width;
        }
      }
      return pos;
    }

    // G lanes per slot (G divides 64): 64 / G slots per wave
    template <int G>
    __device__ __forceinline__ uint32_t groupSlotStart() {
      return (logicalBlock().x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 / G) + laneId() / G;
    }
    template <int G>
    __device__ __forceinline__ uint32_t groupSlotStride() {
      return gridDim.x * (blockDim.x >> 6) * (64 / G);
    }
    template <int G>
    __device__ __forceinline__ uint64_t groupBits(uint64_t m) {
      const uint32_t g = laneId() / G;
      return G == 64 ? m : (m >> (g * G)) & ((1ull << G) - 1);
    }

    template <int G>
    __global__ void __launch_bounds__(kCaBlock) doubletCountKernel(Pool p, CAPool ca) {
      const uint32_t e = logicalBlock().y;
      __shared__ PairTable T;
      const HitView H = hitsOf(p, e);
      if (threadIdx.x == 0)
        loadPairTable(H.layerStart, T);
      __syncthreads();
      uint32_t* dOff = ca.dOff + size_t(e) * (kDoubletSlots + 64);
      uint64_t* dMask = ca.dMask + size_t(e) * 2 * kDoubletSlots;
      const uint32_t ntot = T.cum[kNumPairs - 1];
      const uint32_t sub = laneId() % G;
      for (uint32_t j = groupSlotStart<G>(); j < ntot; j += groupSlotStride<G>()) {
        const SlotScan q = slotOf(H, T, j);
        uint32_t n = 0;
        uint64_t m0 = 0, m1 = 0;  // accepted candidates among the first 128 of the window
        if (q.live) {
          for (uint32_t k0 = 0; k0 < q.n; k0 += G) {
            const uint32_t k = k0 + sub;
            const bool ok = k < q.n && acceptDoublet(H, q, windowPos(q, k));
            const uint64_t bits = groupBits<G>(__ballot(ok));
            n += __popcll(bits);
            if (k0 < 64)
              m0 |= bits << k0;
            else if (k0 < 128)
              m1 |= bits << (k0 - 64);
          }
        }
        if (sub == 0) {
          dOff[j] = n;
          dMask[2 * j] = m0;
          dMask[2 * j + 1] = m1;
        }
      }
    }

    // one workgroup per event: offsets of the doublets, cell count, counter reset
    __global__ void __launch_bounds__(kEventBlock) doubletScanKernel(Pool p, CAPool ca) {
      const uint32_t e = blockIdx.x;
      __shared__ uint32_t lds[kEventBlock / 64];
      __shared__ PairTable T;
      const uint32_t* ls = p.hitsLayerStart + size_t(e) * kLayerStartStride;
      if (threadIdx.x == 0)
        loadPairTable(ls, T);
      __syncthreads();
      uint32_t* dOff = ca.dOff + size_t(e) * (kDoubletSlots + 64);
      const uint32_t ntot = T.cum[kNumPairs - 1];
      const uint32_t total = wgScanArray<kEventBlock>(dOff, ntot, lds);
      const uint32_t nCells = min(total, ca.caps.maxDoublets);
      const uint32_t nHits = T.layer[10];
      // cleared for the following kernels
      uint32_t* ohc = ca.ohCount + size_t(e) * kHitStride;
      for (uint32_t h = threadIdx.x; h < nHits + 1; h += kEventBlock)
        ohc[h] = 0;
      if (threadIdx.x == 0) {
        CACounts c{};
        c.nCells = nCells;
        c.nDoublets = total;
        c.nSlots = ntot;
        c.status = total > ca.caps.maxDoublets ? kCaCellOverflow : 0;
        ca.cnt[e] = c;
      }
    }

    // kRank: each cell takes its place in its outer hit's list with an atomic on the
    // hit's counter (the lists are then built by ohScanKernel and ohFillKernel); without
    // it, ohBuildKernel builds the lists per event in LDS
    template <int G, bool kRank>
    __global__ void __launch_bounds__(kCaBlock) doubletFillKernel(Pool p, CAPool ca) {
      const uint32_t e = logicalBlock().y;
      __shared__ PairTable T;
      const HitView H = hitsOf(p, e);
      if (threadIdx.x == 0)
        loadPairTable(H.layerStart, T);
      __syncthreads();
      const uint32_t* dOff = ca.dOff + size_t(e) * (kDoubletSlots + 64);
      const uint32_t ntot = T.cum[kNumPairs - 1];
      const uint32_t nCells = ca.cnt[e].nCells;
      const size_t cb = ca.cellBase(e);
      uint32_t* ohc = ca.ohCount + size_t(e) * kHitStride;
      const uint64_t* dMask = ca.dMask + size_t(e) * 2 * kDoubletSlots;
      const uint32_t sub = laneId() % G;
      const uint64_t below = (1ull << sub) - 1;
      if (!kRank) {
        // the cells' state for connect and the fishbone, cleared in cell order on the
        // whole grid, four cells per thread with one store per array (the per-event list
        // build would do it on one CU); the cell stride of an event is a multiple of four,
        // so the stores stay inside the event's arrays
        static_assert(kCapsReference.maxDoublets % 4 == 0 && kCapsHighOccupancy.maxDoublets % 4 == 0, "cell stride");
        for (uint32_t c = 4 * gridStart(); c < nCells; c += 4 * gridStride()) {
          *reinterpret_cast<uint32_t*>(ca.cellFlags + cb + c) = 0;
          *reinterpret_cast<uint32_t*>(ca.cellUsed + cb + c) = 0;
          *reinterpret_cast<uint32_t*>(ca.nbCount + cb + c) = 0;
          *reinterpret_cast<int4*>(ca.nbSlot + cb + c) = make_int4(-1, -1, -1, -1);
        }
      }
      for (uint32_t j = groupSlotStart<G>(); j < ntot; j += groupSlotStride<G>()) {
        const uint32_t base = dOff[j];
        if (base >= nCells || dOff[j + 1] == base)
          continue;
        const SlotScan q = slotOf(H, T, j);
        auto write = [&](uint32_t c, uint32_t oi) {
          ca.cellInner[cb + c] = uint16_t(q.i);
          ca.cellOuter[cb + c] = uint16_t(oi);
          ca.cellPair[cb + c] = uint8_t(q.pairId);
          if (kRank)
            ca.ohRank[cb + c] = atomicAdd(&ohc[oi], 1u);
        };
        // the first 128 candidates: the count pass recorded their verdicts; the
        // r-th accepted one is the slot's r-th cell, written by lane r % G
        const uint64_t m0 = dMask[2 * j], m1 = dMask[2 * j + 1];
        const uint32_t c0 = __popcll(m0), c01 = c0 + __popcll(m1);
        for (uint32_t r = sub; r < c01; r += G) {
          const uint32_t k = r < c0 ? nthBit(m0, r) : 64 + nthBit(m1, r - c0);
          if (base + r < nCells)
            write(base + r, candidate(H, q, k));
        }
        // candidates past the first 128 are tested here, G at a time
        uint32_t run = c01;
        for (uint32_t k0 = 128; k0 < q.n; k0 += G) {
          const uint32_t k = k0 + sub;
          uint32_t oi = 0;
          bool ok = false;
          if (k < q.n) {
            const uint32_t w = windowPos(q, k);
            oi = H.phiBins[w];
            ok = acceptDoublet(H, q, w);
          }
          const uint64_t m = groupBits<G>(__ballot(ok));
          const uint32_t c = base + run + __popcll(m & below);
          if (ok && c < nCells)
            write(c, oi);
          run += __popcll(m);
        }
      }
    }

    // ---- outer-hit cell lists (GPUCACell::OuterHitOfCell, first maxCellsPerHit cells ascending) ----
    __global__ void __launch_bounds__(kEventBlock) ohScanKernel(Pool p, CAPool ca) {
      const uint32_t e = blockIdx.x;
      __shared__ uint32_t lds[kEventBlock / 64];
      const uint32_t nHits = p.hitsLayerStart[size_t(e) * kLayerStartStride + 10];
      uint32_t* cnt = ca.ohCount + size_t(e) * kHitStride;
      uint32_t* off = ca.ohOff + size_t(e) * kHitStride;
      uint32_t carry = 0;
      for (uint32_t base = 0; base < nHits; base += kEventBlock) {
        const uint32_t h = base + threadIdx.x;
        const uint32_t v = h < nHits ? cnt[h] : 0;
        uint32_t tot;
        const uint32_t ex = wgExclusiveScan<kEventBlock>(v, tot, lds);
        if (h < nHits)
          off[h] = carry + ex;
        carry += tot;
      }
      if (threadIdx.x == 0)
        off[nHits] = carry;
    }

    // each cell at its outer hit's offset plus the position the doublet fill's
    // count gave it
    __global__ void __launch_bounds__(kCaBlock) ohFillKernel(CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const uint32_t nCells = ca.cnt[e].nCells;
      const size_t cb = ca.cellBase(e);
      const uint32_t* off = ca.ohOff + size_t(e) * kHitStride;
      // four cells per iteration: their loads are issued together
      const uint32_t stride = gridStride();
      for (uint32_t c0 = gridStart(); c0 < nCells; c0 += 4 * stride) {
        uint32_t h[4], hi[4], pos[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t c = c0 + u * stride;
          if (c < nCells) {
            h[u] = ca.cellOuter[cb + c];
            hi[u] = ca.cellInner[cb + c];
            pos[u] = ca.ohRank[cb + c];
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t c = c0 + u * stride;
          if (c < nCells) {
            ca.ohCells[cb + off[h[u]] + pos[u]] = make_uint2(c, hi[u] | h[u] << 16);
            ca.cellFlags[cb + c] = 0;
            ca.cellUsed[cb + c] = 0;
            ca.nbSlot[cb + c] = -1;
            ca.nbCount[cb + c] = 0;
            ca.trSlot[cb + c] = -1;
            ca.trCount[cb + c] = 0;
          }
        }
      }
    }

    // The outer-hit lists of one event by one workgroup, with the counts in LDS: a global
    // atomic per cell costs ~48 ns of device throughput at scattered addresses on this
    // chip, an LDS atomic a few cycles of one CU.  Counts per hit are 16 bits, two to a
    // word (a hit's cells come from the inner hits of at most the 48k of the event, so
    // no count reaches 65536): the counts, a workgroup scan into the offsets, then each
    // cell at its hit's offset plus its rank, taken with a second round of LDS atomics
    // (the consumers do not depend on the order within a list).  The doublet fill has
    // cleared the cells' state for the later stages.
    __global__ void __launch_bounds__(kEventBlock) ohBuildKernel(Pool p, CAPool ca) {
      const uint32_t e = blockIdx.x;
      __shared__ uint32_t hist[kMaxNumClusters / 2];
      __shared__ uint32_t lds[kEventBlock / 64];
      const uint32_t nHits = p.hitsLayerStart[size_t(e) * kLayerStartStride + 10];
      const uint32_t nCells = ca.cnt[e].nCells;
      const size_t cb = ca.cellBase(e);
      uint32_t* off = ca.ohOff + size_t(e) * kHitStride;
      const uint32_t nWords = (nHits + 1) / 2;
      for (uint32_t w = threadIdx.x; w < nWords; w += kEventBlock)
        hist[w] = 0;
      __syncthreads();
      constexpr uint32_t kU = 16;  // cells per thread and iteration: their loads are issued together
      for (uint32_t c0 = threadIdx.x; c0 < nCells; c0 += kU * kEventBlock) {
        uint32_t h[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
          const uint32_t c = c0 + u * kEventBlock;
          h[u] = c < nCells ? ca.cellOuter[cb + c] : 0xffffffffu;
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u)
          if (h[u] != 0xffffffffu)
            atomicAdd(&hist[h[u] >> 1], 1u << ((h[u] & 1) * 16));
      }
      __syncthreads();
      // offsets: each thread scans an even-sized run of hits (whole words), the runs in order
      const uint32_t per = 2 * ((nWords + kEventBlock - 1) / kEventBlock);
      const uint32_t h0 = min(threadIdx.x * per, nHits), h1 = min(h0 + per, nHits);
      uint32_t sum = 0;
      for (uint32_t h = h0; h < h1; ++h)
        sum += (hist[h >> 1] >> ((h & 1) * 16)) & 0xffffu;
      uint32_t total;
      uint32_t run = wgExclusiveScan<kEventBlock>(sum, total, lds);
      for (uint32_t h = h0; h < h1; ++h) {
        off[h] = run;
        run += (hist[h >> 1] >> ((h & 1) * 16)) & 0xffffu;
      }
      if (threadIdx.x == 0)
        off[nHits] = total;
      __syncthreads();
      for (uint32_t w = h0 / 2; w < (h1 + 1) / 2; ++w)
        hist[w] = 0;  // the counts become ranks
      __syncthreads();
      for (uint32_t c0 = threadIdx.x; c0 < nCells; c0 += kU * kEventBlock) {
        uint32_t h[kU], hi[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
          const uint32_t c = c0 + u * kEventBlock;
          h[u] = hi[u] = 0;
          if (c < nCells) {
            h[u] = ca.cellOuter[cb + c];
            hi[u] = ca.cellInner[cb + c];
          }
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
          const uint32_t c = c0 + u * kEventBlock;
          if (c < nCells) {
            const uint32_t sh = (h[u] & 1) * 16;
            const uint32_t rank = (atomicAdd(&hist[h[u] >> 1], 1u << sh) >> sh) & 0xffffu;
            ca.ohCells[cb + off[h[u]] + rank] = make_uint2(c, hi[u] | h[u] << 16);
          }
        }
      }
    }

    // rank of cell c in the outer-hit list of h (ascending cell order); only
    // needed when the list is longer than the `cap` cells the serial container keeps
    __device__ __forceinline__ bool listedInOuterHit(const uint2* cells, uint32_t b, uint32_t n, uint32_t c,
                                                     uint32_t cap) {
      if (n <= cap)
        return true;
      uint32_t rank = 0;
      for (uint32_t k = 0; k < n; ++k)
        rank += cells[b + k].x < c;
      return rank < cap;
    }

    // ---- a13 kernel_connect: CAHitNtupletGeneratorKernelsImpl.h:191-262 ----------------
    // Gathered per outer cell: the neighbours of cell `oc` are the cells `me`
    // whose inner hit is oc's outer hit and which pass the cuts, ascending in
    // `me` (the serial program's append order); `oc` must be among the first
    // maxCellsPerHit cells of its outer hit's list (GPUCACell::OuterHitOfCell).
    //
    // Thread per cell in outer-hit list order, so the lanes of a wave form a few
    // groups sharing an outer hit and hence a candidate list.  The wave first
    // stages the candidates of all its groups in LDS (cell and hit position,
    // loaded by all lanes at once), then each lane runs the cuts of its cell
    // over its group's staged candidates.  Groups that do not fit the staging
    // area read their candidates from memory.
    __device__ __forceinline__ void waveLdsSync() {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    // kStage: staged candidate cells per wave (a launch-shape knob: 128, 256 or 512).  At
    // 256 the LDS it takes (about 22 bytes per entry and wave) allows seven waves per SIMD,
    // whose 72 registers per lane hold the kernel without spills; 128 keeps the same
    // bound (eight waves would leave 64 registers and spill), 512 allows three
    constexpr uint32_t kSegs = 64;  // staged candidate ranges per wave
    template <uint32_t kStage>
    constexpr int connectWaves() {
      return kStage <= 256 ? 7 : 3;
    }

    template <uint32_t kStage>
    __global__ void __launch_bounds__(kCaBlock, connectWaves<kStage>()) connectKernel(Pool p, CAPool ca) {
      const uint32_t e = logicalBlock().y;
      __shared__ PairTable T;
      __shared__ uint32_t sMe[kCaBlock / 64][kStage];
      __shared__ float4 sPos[kCaBlock / 64][kStage];
      __shared__ uint32_t sSegBeg[kCaBlock / 64][kSegs];
      __shared__ uint32_t sSegBase[kCaBlock / 64][kSegs];
      const HitView H = hitsOf(p, e);
      if (threadIdx.x == 0)
        loadPairTable(H.layerStart, T);
      __syncthreads();
      const uint32_t nCells = ca.cnt[e].nCells;
      const size_t cb = ca.cellBase(e);
      const uint32_t* dOff = ca.dOff + size_t(e) * (kDoubletSlots + 64);
      const uint32_t* off = ca.ohOff + size_t(e) * kHitStride;
      uint8_t* used = ca.cellUsed + cb;
      uint32_t* pool = ca.nbPool + size_t(e) * ca.caps.maxCellLists * kCaMaxNeighbors;
      const uint32_t lane = laneId(), w = threadIdx.x >> 6;
      uint32_t status = 0, stored = 0;
      for (uint32_t pos0 = gridStart() - lane; pos0 < nCells; pos0 += gridStride()) {  // uniform per wave
        const uint32_t pos = pos0 + lane;
        const bool valid = pos < nCells;
        uint32_t oc = 0, hi = 0, hm = 0xffffffffu;
        bool live = false;
        float4 p1{0.f, 0.f, 0.f, 0.f}, p2{0.f, 0.f, 0.f, 0.f};
        float thetaCut = 0.f, dcut = 0.f;
        // candidate ranges: the cells of each pair from oc's outer layer starting at hm
        uint32_t beg[6], len[6];
        uint32_t total = 0, nseg = 0;
        if (valid) {
          const uint2 rec = ca.ohCells[cb + pos];
          oc = rec.x;
          hi = rec.y & 0xffff;
          hm = rec.y >> 16;
          const uint32_t lb = off[hm], ln = off[hm + 1] - lb;
          if (ln > ca.caps.maxCellsPerHit)
            status |= kCaOuterHitOverflow;
          live = listedInOuterHit(ca.ohCells + cb, lb, ln, oc, ca.caps.maxCellsPerHit);
          p1 = H.pos[hi];
          p2 = H.pos[hm];
          thetaCut = H.detInd[hm] < 1184 ? kCaThetaCutBarrel : kCaThetaCutForward;
          dcut = H.detInd[hi] < 96 ? kCaDcaCutInner : kCaDcaCutOuter;
        }
        // the outer layer of the cell's pair: the layer of its outer hit
        int lo = 0;
        for (int l = 1; l < 10; ++l)
          lo += valid && hm >= T.layer[l];
        bool more = valid;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          beg[q] = 0;
          len[q] = 0;
          const int pr = more ? kPairsFromLayer[lo][q] : -1;
          more = pr >= 0;
          if (more) {
            const uint32_t j = (pr ? T.cum[pr - 1] : 0) + (hm - T.layer[lo]);
            const uint32_t b = dOff[j], en = min(dOff[j + 1], nCells);
            beg[q] = b;
            len[q] = en > b ? en - b : 0;
            total += len[q];
            nseg += len[q] > 0;
          }
        }
        // groups: consecutive lanes with the same outer hit; the first lane leads
        const uint32_t hmPrev = __shfl_up(hm, 1);
        const bool leader = valid && (lane == 0 || hmPrev != hm);
        const uint32_t segOff = waveExclusiveScan(leader ? nseg : 0);
        const uint32_t candOff = waveExclusiveScan(leader ? total : 0);
        const bool fits = leader && segOff + nseg <= kSegs && candOff + total <= kStage;
        if (fits) {
          uint32_t sg = segOff, base = candOff;
#pragma unroll
          for (int q = 0; q < 6; ++q) {
            if (len[q]) {
              sSegBeg[w][sg] = beg[q];
              sSegBase[w][sg] = base;
              ++sg;
              base += len[q];
            }
          }
        }
        // the groups that fit are the first ones of the wave (offsets only grow)
        const uint32_t nSegs = waveMax(fits ? segOff + nseg : 0);
        const uint32_t nStaged = waveMax(fits ? candOff + total : 0);
        const uint64_t leaders = __ballot(leader);
        const uint64_t upTo = lane == 63 ? ~0ull : (2ull << lane) - 1;
        const uint64_t mine = leaders & upTo;
        const uint32_t lead = mine ? 63 - __clzll(mine) : 0;
        const bool staged = __shfl(fits ? 1u : 0u, lead) != 0;
        const uint32_t gBase = __shfl(candOff, lead);
        waveLdsSync();
        // all lanes load the staged candidates: slot s belongs to the last
        // range whose base is <= s
        for (uint32_t s = lane; s < nStaged; s += 64) {
          uint32_t lo2 = 0, hi2 = nSegs;  // sSegBase[lo2] <= s < sSegBase[hi2] (hi2 = nSegs: end)
          while (hi2 - lo2 > 1) {
            const uint32_t mid = (lo2 + hi2) >> 1;
            if (sSegBase[w][mid] <= s)
              lo2 = mid;
            else
              hi2 = mid;
          }
          const uint32_t me = sSegBeg[w][lo2] + (s - sSegBase[w][lo2]);
          sMe[w][s] = me;
          sPos[w][s] = H.pos[ca.cellOuter[cb + me]];
        }
        waveLdsSync();
        if (valid && live) {
          int32_t slot = -1;
          uint32_t n = 0;
          // the accepted candidates, in ascending order within each range
          auto append = [&](uint32_t me) {
            // a byte per cell: every writer stores the same value, no atomic needed
            used[me] = 1;
            if (n == 0)
              used[oc] = 1;
            if (n < kCaMaxNeighbors) {
              if (slot == -1) {
                const uint32_t sl = waveAlloc(&ca.cnt[e].nbLists, true);
                if (sl + 1 < ca.caps.maxCellLists) {
                  slot = int32_t(sl);
                } else {
                  slot = -2;
                  status |= kCaNeighborStoreOverflow;
                }
              }
              if (slot >= 0)
                pool[size_t(slot) * kCaMaxNeighbors + n] = me;
            } else {
              status |= kCaNeighborOverflow;
            }
            ++n;
          };
          // The two cuts in two passes: the r-z alignment over every candidate, then the
          // circle cut over the aligned ones only.  About 2 % of the candidates are
          // aligned, so the circle cut (three square roots and a division) runs a couple
          // of times per chunk instead of on every candidate some lane of the wave is
          // still testing.  A candidate passes when both cuts pass, as in the single test.
          auto aligned = [&](const float4& p3) {
            return alignedRZ(p1.w, p1.z, p2.w, p2.z, p3.w, p3.z, kCaPtMin, thetaCut);
          };
          auto curved = [&](const float4& p3) {
            return dcaCut(p1.x, p1.y, p2.x, p2.y, p3.x, p3.y, dcut, kCaHardCurvCut);
          };
          if (staged) {
            for (uint32_t b0 = 0; b0 < total; b0 += 64) {
              const uint32_t m = min(64u, total - b0);
              const float4* cand = sPos[w] + gBase + b0;
              uint64_t al = 0;
              for (uint32_t k = 0; k < m; ++k)
                if (aligned(cand[k]))
                  al |= 1ull << k;
              uint64_t acc = 0;
              for (uint64_t rest = al; rest; rest &= rest - 1) {
                const uint32_t k = __ffsll((unsigned long long)rest) - 1;
                if (curved(cand[k]))
                  acc |= 1ull << k;
              }
              for (; acc; acc &= acc - 1)
                append(sMe[w][gBase + b0 + __ffsll((unsigned long long)acc) - 1]);
            }
          } else {
            // the ranges again (read anew rather than held in registers through the
            // staging: a group that does not fit the staging area is rare)
            for (int q = 0; q < 6; ++q) {
              const int pr = kPairsFromLayer[lo][q];
              if (pr < 0)
                break;
              const uint32_t j = (pr ? T.cum[pr - 1] : 0) + (hm - T.layer[lo]);
              const uint32_t rb = dOff[j], re = min(dOff[j + 1], nCells);
              const uint32_t rl = re > rb ? re - rb : 0;
              // chunks of up to 64 candidates: the alignment first, in a loop without
              // stores (so the candidates' loads are issued together), then the circle
              // cut on the aligned ones, then the accepted ones in ascending order
              for (uint32_t b0 = 0; b0 < rl; b0 += 64) {
                const uint32_t m = min(64u, rl - b0), base = rb + b0;
                uint64_t al = 0;
                for (uint32_t k = 0; k < m; ++k)
                  if (aligned(H.pos[ca.cellOuter[cb + base + k]]))
                    al |= 1ull << k;
                uint64_t acc = 0;
                for (uint64_t rest = al; rest; rest &= rest - 1) {
                  const uint32_t k = __ffsll((unsigned long long)rest) - 1;
                  if (curved(H.pos[ca.cellOuter[cb + base + k]]))
                    acc |= 1ull << k;
                }
                for (; acc; acc &= acc - 1)
                  append(base + __ffsll((unsigned long long)acc) - 1);
              }
            }
          }
          if (slot >= 0) {
            ca.nbSlot[cb + oc] = slot;
            ca.nbCount[cb + oc] = uint8_t(min(n, kCaMaxNeighbors));
            stored += min(n, kCaMaxNeighbors);
          }
        }
        waveLdsSync();
      }
      if (status)
        atomicOr(&ca.cnt[e].status, status);
      waveAdd(&ca.cnt[e].nNeighbors, stored);
    }

    // ---- a14 fishbone: gpuFishbone.h:21-90 --------------------------------------------
    // The serial loop visits the used cells of each outer-hit list (first maxCellsPerHit,
    // ascending) in rows: row ci kills the later aligned cells until the first
    // aligned later cell that is closer, which kills ci and ends the row.  The
    // rows do not read the kills, so each row runs on its own (one thread per
    // cell, in list order), finding its first killer as the smallest such cell.
    // The late pass (checkTrack) takes part only with the cells that are in tuples.
    // A cell seen from the list's hit: its inner hit relative to that hit.
    struct RowCell {
      uint32_t c;
      uint32_t det;      // detInd of the inner hit; bit 16: the cell is used
      float x, y, z, n;  // inner hit minus the list's hit, squared norm
    };

    // lists that the wave's 64 positions touch, staged whole (up to kFbWindow entries, a
    // launch-shape knob: 128, 256 or 384; at 256 six workgroups fit a CU); a wave whose
    // lists span more reads them from memory
    template <bool kCheckTrack, uint32_t kFbWindow>
    __global__ void __launch_bounds__(kCaBlock) fishboneKernel(Pool p, CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const HitView H = hitsOf(p, e);
      if (H.layerStart[10] < 2)
        return;
      // per wave: every cell of the lists its positions belong to, as seen from its
      // list's hit (direction, squared norm; cell, detInd | used << 16)
      __shared__ float4 sDir[kCaBlock / 64][kFbWindow];
      __shared__ uint2 sCell[kCaBlock / 64][kFbWindow];
      const uint32_t nCells = ca.cnt[e].nCells;
      const size_t cb = ca.cellBase(e);
      const uint32_t* off = ca.ohOff + size_t(e) * kHitStride;
      const uint2* cells = ca.ohCells + cb;
      const uint8_t* used = ca.cellUsed + cb;
      const uint16_t* trCount = ca.trCount + cb;
      auto isUsed = [&](uint32_t c) {
        return used[c] && (!kCheckTrack || trCount[c] > 0);
      };
      // a list entry seen from its own list's hit
      auto entry = [&](uint32_t k) {
        RowCell d;
        const uint2 rec = cells[k];
        d.c = rec.x;
        const uint32_t hj = rec.y & 0xffff;
        const float4 ph = H.pos[rec.y >> 16];
        const float4 pj = H.pos[hj];
        d.x = pj.x - ph.x;
        d.y = pj.y - ph.y;
        d.z = pj.z - ph.z;
        d.n = d.x * d.x + d.y * d.y + d.z * d.z;
        d.det = uint32_t(H.detInd[hj]) | (isUsed(d.c) << 16);
        return d;
      };
      const uint32_t lane = laneId(), wave = threadIdx.x / 64;
      float4* wDir = sDir[wave];
      uint2* wCell = sCell[wave];
      for (uint32_t pos0 = gridStart() - lane; pos0 < nCells; pos0 += gridStride()) {  // uniform per wave
        const uint32_t pos = pos0 + lane;
        uint32_t b = 0, n = 0;
        if (pos < nCells) {
          const uint32_t h = cells[pos].y >> 16;
          b = off[h];
          n = off[h + 1] - b;
        }
        // the window: from the first list's start to the last list's end
        const uint32_t last = min(63u, nCells - 1 - pos0);
        const uint32_t wb = __shfl(b, 0);
        const uint32_t we = __shfl(b + n, int(last));
        const bool staged = we - wb <= kFbWindow;
        if (staged) {
          for (uint32_t k = lane; k < we - wb; k += 64) {
            const RowCell d = entry(wb + k);
            wDir[k] = make_float4(d.x, d.y, d.z, d.n);
            wCell[k] = make_uint2(d.c, d.det);
          }
        }
        waveLdsSync();
        if (pos < nCells && n >= 2) {
          const RowCell me = staged ? RowCell{wCell[pos - wb].x, wCell[pos - wb].y, wDir[pos - wb].x, wDir[pos - wb].y,
                                              wDir[pos - wb].z, wDir[pos - wb].w}
                                    : entry(pos);
          const uint32_t ci = me.c;
          if (me.det >> 16) {
            // with more than maxCellsPerHit cells only that many lowest indices take part
            const uint32_t cap = ca.caps.maxCellsPerHit;
            uint32_t limit = 0xffffffffu;
            bool listed = true;
            if (n > cap) {
              listed = listedInOuterHit(cells, b, n, ci, cap);
              for (uint32_t k = 0; listed && k < n; ++k) {
                const uint32_t c = cells[b + k].x;
                uint32_t rank = 0;
                for (uint32_t m = 0; m < n; ++m)
                  rank += cells[b + m].x < c;
                if (rank == cap)
                  limit = c;
              }
            }
            auto fromMemory = [&](uint32_t k) { return entry(b + k); };
            const uint32_t lo = b - wb;
            auto fromLds = [&](uint32_t k) {
              const float4 v = wDir[lo + k];
              const uint2 c = wCell[lo + k];
              return RowCell{c.x, c.y, v.x, v.y, v.z, v.w};
            };
            // one pass: the aligned later cells, and the first closer one among them
            auto row = [&](auto get) {
              uint32_t killer = limit;
              uint64_t alignedMask = 0;  // list positions (lists of up to 64 cells)
              for (uint32_t k = 0; k < n; ++k) {
                const RowCell d = get(k);
                if (d.c <= ci || d.c >= limit || !(d.det >> 16))
                  continue;
                const float cos12 = me.x * d.x + me.y * d.y + me.z * d.z;
                if (!((me.det & 0xffff) != (d.det & 0xffff) && cos12 * cos12 >= 0.99999f * me.n * d.n))
                  continue;
                if (me.n > d.n && d.c < killer)
                  killer = d.c;
                if (k < 64)
                  alignedMask |= 1ull << k;
              }
              // the aligned cells before the killer are killed (several rows may
              // kill a cell: they store the same byte)
              for (uint64_t m = alignedMask; m; m &= m - 1) {
                const uint32_t cj = get(__ffsll((unsigned long long)m) - 1).c;
                if (cj < killer)
                  ca.cellFlags[cb + cj] |= 2;
              }
              for (uint32_t k = 64; k < n; ++k) {  // positions past 64: second pass
                const RowCell d = get(k);
                if (d.c <= ci || d.c >= killer || !(d.det >> 16))
                  continue;
                const float cos12 = me.x * d.x + me.y * d.y + me.z * d.z;
                if ((me.det & 0xffff) != (d.det & 0xffff) && cos12 * cos12 >= 0.99999f * me.n * d.n)
                  ca.cellFlags[cb + d.c] |= 2;
              }
              if (killer != limit)
                ca.cellFlags[cb + ci] |= 2;
            };
            if (listed) {
              if (staged)
                row(fromLds);
              else
                row(fromMemory);
            }
          }
        }
        waveLdsSync();
      }
    }

    // ---- a15 kernel_find_ntuplets: GPUCACell.h:278-331 ----------------------------------
    // Iterative depth-first walk from a start cell, visiting neighbours in list
    // order: tuples come out in the serial program's order.  Count mode counts
    // tuples and hits and records the paths of the start cell's first kPaths tuples
    // (the neighbour-list position taken at each level, the tuple's size and its place
    // among the cell's tuples and hits: 27 bits); fill mode writes the tuples from the
    // cell's scanned offsets, from the skip-th on.
    constexpr uint32_t kPaths = 4;
    template <bool kFill>
    __device__ void walk(const CAPool& ca, uint32_t e, uint32_t c, uint32_t tBase, uint32_t hBase, uint32_t& nT,
                         uint32_t& nH, uint32_t& status, uint32_t skip, uint32_t (&paths)[kPaths]) {
      const size_t cb = ca.cellBase(e);
      const uint32_t* pool = ca.nbPool + size_t(e) * ca.caps.maxCellLists * kCaMaxNeighbors;
      const size_t tb = ca.tupleBase(e);
      // the stack (at most 4 cells) in registers: the cells as four values read through
      // selects, the list positions as bytes of one word, the "any child" flags as bits
      uint32_t st0 = c, st1 = 0, st2 = 0, st3 = 0;
      auto cellAt = [&](int d) { return d == 0 ? st0 : d == 1 ? st1 : d == 2 ? st2 : st3; };
      auto setCell = [&](int d, uint32_t v) {
        st1 = d == 1 ? v : st1;
        st2 = d == 2 ? v : st2;
        st3 = d == 3 ? v : st3;
      };
      uint32_t posBytes = 0, anyBits = 0;
      int depth = 1;
      nT = nH = 0;
      while (depth > 0) {
        const int top = depth - 1;
        const uint32_t cell = cellAt(top);
        const int32_t slot = ca.nbSlot[cb + cell];
        const uint32_t cnt = slot >= 0 ? ca.nbCount[cb + cell] : 0;
        bool pushed = false;
        uint32_t pos = (posBytes >> (8 * top)) & 0xff;
        while (pos < cnt) {
          const uint32_t other = pool[size_t(slot) * kCaMaxNeighbors + pos];
          ++pos;
          if (ca.cellFlags[cb + other] & 2)
            continue;
          anyBits |= 1u << top;
          if (depth >= 4) {
            status |= kCaDepthOverflow;
            continue;
          }
          setCell(depth, other);
          posBytes &= ~(0xffu << (8 * depth));  // the child's position starts at 0
          anyBits &= ~(1u << depth);
          ++depth;
          pushed = true;
          break;
        }
        posBytes = (posBytes & ~(0xffu << (8 * top))) | (pos << (8 * top));
        if (pushed)
          continue;
        if (!((anyBits >> top) & 1) && depth >= 2) {
          const uint32_t nh = uint32_t(depth) + 1;
          if (!kFill && nT < kPaths) {
            const uint32_t code = (((posBytes >> 0) & 0xff) - 1) | ((((posBytes >> 8) & 0xff) - 1) & 63) << 6 |
                                  ((((posBytes >> 16) & 0xff) - 1) & 63) << 12 | uint32_t(depth - 2) << 18 |
                                  nT << 20 | nH << 22;
#pragma unroll
            for (uint32_t i = 0; i < kPaths; ++i)
              paths[i] = i == nT ? code : paths[i];
          }
          if (kFill && nT >= skip) {
            const uint32_t m = tBase + nT;
            if (m < ca.caps.maxTuples) {
              const uint32_t first = hBase + nH;
              ca.tupleOff[ca.offBase(e) + m] = first;
              uint16_t* bins = ca.tupleBins + 5 * tb + first;
              for (int k = 0; k < depth; ++k) {
                const uint32_t ck = cellAt(k);
                bins[k] = ca.cellInner[cb + ck];
                ca.tupleCells[4 * tb + 4 * size_t(m) + k] = ck;
              }
              bins[depth] = ca.cellOuter[cb + cell];
              for (int k = depth; k < 4; ++k)
                ca.tupleCells[4 * tb + 4 * size_t(m) + k] = 0xffffffffu;
              ca.quality[tb + m] = kQualityBad;
            }
          }
          ++nT;
          nH += nh;
        }
        --depth;
      }
    }

    // The walk starts from every live cell of a starting layer pair that has
    // neighbours (a cell without neighbours makes no tuple).  These start cells
    // are compacted in ascending order first: a bit per cell, a scan of the
    // words' popcounts, and each word's cells written at its offset.
    __global__ void __launch_bounds__(kCaBlock) startMaskKernel(CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const uint32_t nCells = ca.cnt[e].nCells;
      const size_t cb = ca.cellBase(e);
      uint64_t* mask = ca.startMask + size_t(e) * (ca.caps.maxDoublets / 64 + 64);
      const uint32_t words = (nCells + 63) / 64;
      for (uint32_t c = gridStart(); c < words * 64; c += gridStride()) {
        bool start = false;
        if (c < nCells) {
          const int pid = ca.cellPair[cb + c];
          // a cell has a neighbour list exactly when its count is positive (one byte
          // read instead of the four of its slot)
          start = !(ca.cellFlags[cb + c] & 2) && (pid < 8 || pid > 12) && ca.nbCount[cb + c] > 0;
        }
        const uint64_t m = __ballot(start);
        if ((threadIdx.x & 63) == 0)
          mask[c / 64] = m;
      }
    }

    __global__ void __launch_bounds__(kEventBlock) startCompactKernel(CAPool ca) {
      const uint32_t e = blockIdx.x;
      __shared__ uint32_t lds[kEventBlock / 64];
      const uint32_t nCells = ca.cnt[e].nCells;
      const size_t cb = ca.cellBase(e);
      const uint64_t* mask = ca.startMask + size_t(e) * (ca.caps.maxDoublets / 64 + 64);
      uint32_t* list = ca.startCells + cb;
      const uint32_t words = (nCells + 63) / 64;
      uint32_t carry = 0;
      for (uint32_t base = 0; base < words; base += kEventBlock) {
        const uint32_t w = base + threadIdx.x;
        uint64_t m = w < words ? mask[w] : 0;
        uint32_t tot;
        uint32_t at = carry + wgExclusiveScan<kEventBlock>(uint32_t(__popcll(m)), tot, lds);
        while (m) {
          const uint32_t bit = __ffsll((unsigned long long)m) - 1;
          list[at++] = w * 64 + bit;
          m &= m - 1;
        }
        carry += tot;
      }
      if (threadIdx.x == 0)
        ca.cnt[e].nStart = carry;
    }

    __global__ void __launch_bounds__(kCaBlock) ntupletCountKernel(CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const uint32_t nStart = ca.cnt[e].nStart;
      const size_t cb = ca.cellBase(e);
      const size_t tb = ca.tupleBase(e);
      const uint32_t* list = ca.startCells + cb;
      uint32_t status = 0;
      const uint32_t lane = laneId();
      for (uint32_t k0 = gridStart() - lane; k0 < nStart; k0 += gridStride()) {  // uniform per wave
        const uint32_t k = k0 + lane;
        uint32_t nT = 0, nH = 0;
        uint32_t paths[kPaths] = {0, 0, 0, 0};
        if (k < nStart) {
          walk<false>(ca, e, list[k], 0, 0, nT, nH, status, 0, paths);
          ca.dfsT[cb + k] = nT;
          ca.dfsH[cb + k] = nH;
        }
        // the recorded paths, stored compactly: one counter update per wave
        const uint32_t n = min(nT, kPaths);
        const uint32_t at = waveExclusiveScan(n);
        const uint32_t total = __shfl(at + n, 63);
        uint32_t base = 0;
        if (lane == 0 && total)
          base = atomicAdd(&ca.cnt[e].nPaths, total);
        base = __shfl(base, 0);
        for (uint32_t i = 0; i < n; ++i)
          if (base + at + i < ca.caps.maxTuples)
            ca.tuplePaths[tb + base + at + i] = make_uint2(k, paths[i]);
      }
      if (status)
        atomicOr(&ca.cnt[e].status, status);
    }

    // tuples and hits per start cell scanned together (the two counts packed in 64 bits,
    // four start cells per thread and pass)
    __global__ void __launch_bounds__(kEventBlock) ntupletScanKernel(CAPool ca) {
      const uint32_t e = blockIdx.x;
      __shared__ uint64_t lds[kEventBlock / 64];
      const uint32_t nStart = ca.cnt[e].nStart;
      const size_t cb = ca.cellBase(e);
      uint32_t* t = ca.dfsT + cb;
      uint32_t* h = ca.dfsH + cb;
      constexpr uint32_t kV = 4;
      uint64_t carry = 0;
      for (uint32_t base = 0; base < nStart; base += kEventBlock * kV) {
        const uint32_t i0 = base + threadIdx.x * kV;
        uint64_t v[kV], sum = 0;
#pragma unroll
        for (uint32_t u = 0; u < kV; ++u) {
          v[u] = i0 + u < nStart ? uint64_t(t[i0 + u]) | uint64_t(h[i0 + u]) << 32 : 0;
          sum += v[u];
        }
        uint64_t tot;
        uint64_t ex = carry + wgExclusiveScan<kEventBlock>(sum, tot, lds);
#pragma unroll
        for (uint32_t u = 0; u < kV; ++u) {
          if (i0 + u < nStart) {
            t[i0 + u] = uint32_t(ex);
            h[i0 + u] = uint32_t(ex >> 32);
          }
          ex += v[u];
        }
        carry += tot;
      }
      const uint32_t carryT = uint32_t(carry), carryH = uint32_t(carry >> 32);
      if (threadIdx.x == 0) {
        CACounts& c = ca.cnt[e];
        c.apcM = carryT;
        c.apcN = carryH;
        c.nTuples = min(carryT, ca.caps.maxTuples);
        if (carryT > ca.caps.maxTuples)
          c.status |= kCaTupleOverflow;
      }
    }

    // The tuples whose paths the counting walk recorded are written one per thread,
    // following the recorded neighbour-list positions; only start cells with more than
    // kPaths tuples are walked again, for the rest of their tuples (all of them if the
    // path store overflowed: the tuples then written twice are written identically).
    __global__ void __launch_bounds__(kCaBlock) ntupletFillKernel(CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const CACounts cnt = ca.cnt[e];
      const size_t cb = ca.cellBase(e);
      const size_t tb = ca.tupleBase(e);
      const uint32_t* list = ca.startCells + cb;
      const uint32_t M = ca.caps.maxTuples;
      const uint32_t* pool = ca.nbPool + size_t(e) * ca.caps.maxCellLists * kCaMaxNeighbors;
      const uint32_t nPaths = min(cnt.nPaths, M);
      for (uint32_t r = gridStart(); r < nPaths; r += gridStride()) {
        const uint2 rec = ca.tuplePaths[tb + r];
        const uint32_t k = rec.x, code = rec.y;
        const uint32_t m = ca.dfsT[cb + k] + ((code >> 20) & 3);
        if (m >= M)
          continue;
        const uint32_t depth = ((code >> 18) & 3) + 2;
        const uint32_t first = ca.dfsH[cb + k] + (code >> 22);
        uint32_t cell = list[k];
        ca.tupleOff[ca.offBase(e) + m] = first;
        uint16_t* bins = ca.tupleBins + 5 * tb + first;
        uint32_t* cells = ca.tupleCells + 4 * tb + 4 * size_t(m);
        for (uint32_t d = 0; d < depth; ++d) {
          if (d > 0) {
            const int32_t slot = ca.nbSlot[cb + cell];
            cell = pool[size_t(slot) * kCaMaxNeighbors + ((code >> (6 * (d - 1))) & 63)];
          }
          bins[d] = ca.cellInner[cb + cell];
          cells[d] = cell;
        }
        bins[depth] = ca.cellOuter[cb + cell];
        for (uint32_t d = depth; d < 4; ++d)
          cells[d] = 0xffffffffu;
        ca.quality[tb + m] = kQualityBad;
      }
      const bool rewalkAll = cnt.nPaths > M;
      uint32_t status = 0;
      uint32_t paths[kPaths];
      for (uint32_t k = gridStart(); k < cnt.nStart; k += gridStride()) {
        const uint32_t tBase = ca.dfsT[cb + k];
        const uint32_t next = k + 1 < cnt.nStart ? ca.dfsT[cb + k + 1] : cnt.apcM;
        if (next - tBase <= (rewalkAll ? 0 : kPaths) || tBase >= M)
          continue;  // every tuple of this cell was recorded (or none is stored)
        uint32_t nT, nH;
        walk<true>(ca, e, list[k], tBase, ca.dfsH[cb + k], nT, nH, status, rewalkAll ? 0 : kPaths, paths);
      }
    }

    // Per-cell track lists (the tuples each cell is in; CAHitNtupletGeneratorKernelsImpl.h:
    // 118-149 fills them during the walk), built per event in LDS: workgroup (s, e) takes the
    // slices of kTrSlice cells s, s + gridDim.x, ... of event e and passes twice over the
    // event's tuples per slice (about 16 bytes each, 10k at PU50): counting each of its
    // cells' tuples with LDS atomics, then, once every cell with tuples has a slot of the
    // list store (one counter update per workgroup, slots in cell order within it),
    // writing each tuple into its cells' lists at the rank an LDS atomic gives (the sort
    // below orders each list).  Between the passes a cell's LDS word becomes its list's
    // rank in the workgroup (high 14 bits) plus its tuples placed so far (low 18 bits; a
    // cell has fewer than 2^18 tuples), so the second pass reads no slot from memory.
    // Slot s records its cell and the cell's tuple count (the walk's scratch is free now),
    // so the later passes visit only the cells that are in tuples.  Workgroup 0 of each
    // event also finalises the tuple offsets (HistoContainer.h:193-204 bulkFinalizeFill).
    constexpr uint32_t kTrSlice = 16384;       // cells per slice: 64 KB of counts
    constexpr uint32_t kTrGroupCells = 40960;  // capacity cells per workgroup: 12 per event at
                                               // 458752 cells, one slice each up to 196608 cells
    static_assert(kTrSlice <= (1u << 14) && kCapsReference.maxTuples < (1u << 18) &&
                      kCapsHighOccupancy.maxTuples < (1u << 18),
                  "rank and position share a 32-bit LDS word");

    __global__ void __launch_bounds__(kEventBlock) trackListBuildKernel(Pool p, CAPool ca) {
      const uint32_t s = blockIdx.x, e = blockIdx.y;
      __shared__ uint32_t cnt[kTrSlice];
      __shared__ uint32_t lds[kEventBlock / 64];
      __shared__ uint32_t sBase;
      const CACounts cn = ca.cnt[e];
      const size_t cb = ca.cellBase(e);
      const size_t tb = ca.tupleBase(e);
      const uint32_t M = ca.caps.maxTuples;
      if (s == 0) {
        uint32_t* off = ca.tupleOff + ca.offBase(e);
        if (cn.apcM >= M) {
          if (threadIdx.x == 0)
            off[M] = off[M - 1];
        } else {
          for (uint32_t i = cn.apcM + threadIdx.x; i <= M; i += kEventBlock)
            off[i] = cn.apcN;
        }
      }
      // a17 kernel_fillHitDetIndices, and the fit outputs of every stored tuple cleared
      // (the fits write those of the tuples they fit), spread over the event's workgroups
      {
        const uint16_t* detInd = p.detInd + size_t(e) * kMaxNumClusters;
        const uint32_t nEntries = cn.apcM >= M ? ca.tupleOff[ca.offBase(e) + M - 1] : cn.apcN;
        const uint32_t stride = gridDim.x * kEventBlock;
        for (uint32_t i = s * kEventBlock + threadIdx.x; i < nEntries; i += stride)
          ca.detIndices[5 * tb + i] = detInd[ca.tupleBins[5 * tb + i]];
        for (uint32_t t = s * kEventBlock + threadIdx.x; t < cn.nTuples; t += stride) {
          for (int k = 0; k < 5; ++k)
            ca.state[5 * tb + size_t(k) * M + t] = 0;
          for (int k = 0; k < 15; ++k)
            ca.cov[15 * tb + size_t(k) * M + t] = 0;
          ca.pt[tb + t] = 0;
          ca.eta[tb + t] = 0;
          ca.chi2[tb + t] = 0;
        }
      }
      const uint32_t nT = cn.nTuples;
      const uint4* cells4 = reinterpret_cast<const uint4*>(ca.tupleCells + 4 * tb);
      uint32_t* trCells = ca.dfsT + cb;
      uint32_t* trTotal = ca.dfsH + cb;
      TupleIndex* pool = ca.trPool + size_t(e) * ca.caps.maxCellLists * kCaMaxCellTracks;
      uint32_t status = 0;
      for (uint32_t c0 = s * kTrSlice; c0 < cn.nCells; c0 += gridDim.x * kTrSlice) {
        const uint32_t n = min(kTrSlice, cn.nCells - c0);
        __syncthreads();  // the previous slice's ranks are no longer read
        for (uint32_t i = threadIdx.x; i < n; i += kEventBlock)
          cnt[i] = 0;
        __syncthreads();
        auto inSlice = [&](uint32_t c) { return c != 0xffffffffu && c - c0 < n; };
        for (uint32_t t = threadIdx.x; t < nT; t += kEventBlock) {
          const uint4 c = cells4[t];
          if (inSlice(c.x))
            atomicAdd(&cnt[c.x - c0], 1u);
          if (inSlice(c.y))
            atomicAdd(&cnt[c.y - c0], 1u);
          if (inSlice(c.z))
            atomicAdd(&cnt[c.z - c0], 1u);
          if (inSlice(c.w))
            atomicAdd(&cnt[c.w - c0], 1u);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += kEventBlock)
          if (cnt[i] == 0) {
            ca.trSlot[cb + c0 + i] = -1;
            ca.trCount[cb + c0 + i] = 0;
          }
        const uint32_t per = (n + kEventBlock - 1) / kEventBlock;
        const uint32_t i0 = min(threadIdx.x * per, n), i1 = min(i0 + per, n);
        uint32_t mine = 0;
        for (uint32_t i = i0; i < i1; ++i)
          mine += cnt[i] > 0;
        uint32_t total;
        uint32_t rank = wgExclusiveScan<kEventBlock>(mine, total, lds);
        if (threadIdx.x == 0)
          sBase = total ? atomicAdd(&ca.cnt[e].trLists, total) : 0;
        __syncthreads();
        for (uint32_t i = i0; i < i1; ++i) {
          const uint32_t k = cnt[i];
          if (k == 0)
            continue;
          const uint32_t slot = sBase + rank++;
          const uint32_t c = c0 + i;
          if (slot + 1 < ca.caps.maxCellLists) {
            ca.trSlot[cb + c] = int32_t(slot);
            trCells[slot] = c;
            trTotal[slot] = k;
          } else {
            ca.trSlot[cb + c] = -2;
            status |= kCaTrackStoreOverflow;
          }
          cnt[i] = (rank - 1) << 18;  // the list's rank in the workgroup above, tuples placed below
        }
        __syncthreads();
        auto place = [&](uint32_t c, uint32_t t) {
          if (!inSlice(c))
            return;
          const uint32_t v = atomicAdd(&cnt[c - c0], 1u);
          const uint32_t slot = sBase + (v >> 18), pos = v & 0x3ffffu;
          if (slot + 1 < ca.caps.maxCellLists && pos < kCaMaxCellTracks)
            pool[size_t(slot) * kCaMaxCellTracks + pos] = t;
        };
        for (uint32_t t = threadIdx.x; t < nT; t += kEventBlock) {
          const uint4 c = cells4[t];
          place(c.x, t);
          place(c.y, t);
          place(c.z, t);
          place(c.w, t);
        }
      }
      if (status)
        atomicOr(&ca.cnt[e].status, status);
    }

    // sorts each cell's track list (ascending tuple index, first 48 kept) and
    // runs a16 kernel_earlyDuplicateRemover (CAHitNtupletGeneratorKernelsImpl.h:126-149)
    // on it: order independent, it only marks shorter tuples as duplicates.
    __global__ void __launch_bounds__(kCaBlock) trackListSortKernel(CAPool ca) {
      const uint32_t e = logicalBlock().y;
      const uint32_t nT = ca.cnt[e].nTuples;
      const size_t cb = ca.cellBase(e);
      const size_t tb = ca.tupleBase(e);
      TupleIndex* pool = ca.trPool + size_t(e) * ca.caps.maxCellLists * kCaMaxCellTracks;
      const uint32_t* off = ca.tupleOff + ca.offBase(e);
      const uint32_t* trCells = ca.dfsT + cb;
      const uint32_t* trTotal = ca.dfsH + cb;
      const uint32_t nLists = min(ca.cnt[e].trLists, ca.caps.maxCellLists - 1);
      uint32_t status = 0;
      for (uint32_t slot = gridStart(); slot < nLists; slot += gridStride()) {
        const uint32_t c = trCells[slot], total = trTotal[slot];
        TupleIndex* list = pool + size_t(slot) * kCaMaxCellTracks;
        uint32_t n = total;
        if (total > kCaMaxCellTracks) {
          // keep the 48 smallest tuple indices: rebuild from the tuples
          status |= kCaTrackOverflow;
          n = 0;
          for (uint32_t t = 0; t < nT && n < kCaMaxCellTracks; ++t)
            for (int k = 0; k < 4; ++k) {
              const uint32_t x = ca.tupleCells[4 * tb + 4 * size_t(t) + k];
              if (x == 0xffffffffu)
                break;
              if (x == c) {
                list[n++] = t;
                break;
              }
            }
        } else {
          for (uint32_t k = 1; k < n; ++k) {
            const TupleIndex v = list[k];
            uint32_t m = k;
            while (m > 0 && list[m - 1] > v) {
              list[m] = list[m - 1];
              --m;
            }
            list[m] = v;
          }
        }
        ca.trCount[cb + c] = uint16_t(n);
        if (n < 2)
          continue;
        uint32_t maxNh = 0;
        for (uint32_t k = 0; k < n; ++k)
          maxNh = max(maxNh, off[list[k] + 1] - off[list[k]]);
        for (uint32_t k = 0; k < n; ++k)
          if (off[list[k] + 1] - off[list[k]] != maxNh)
            ca.quality[tb + list[k]] = kQualityDup;
        const uint32_t at = waveAlloc(&ca.cnt[e].nMulti, true);
        ca.multiCells[cb + at] = c;
      }
      if (status)
        atomicOr(&ca.cnt[e].status, status);
    }

    // ---- a16 tuple multiplicity (kernel_countMultiplicity / fillMultiplicity) ----------
    // One workgroup per event.  Within a size the serial fill leaves tuples in
    // descending index order; each tuple's place is computed from its rank.
    __global__ void __launch_bounds__(kEventBlock) multiplicityKernel(Pool p, CAPool ca) {
      const uint32_t e = blockIdx.x;
      __shared__ uint32_t binCount[8], binEnd[8], carry[8];
      __shared__ uint32_t waveCnt[8][kEventBlock / 64];
      const uint32_t nT = ca.cnt[e].nTuples;
      const size_t tb = ca.tupleBase(e);
      const uint32_t* off = ca.tupleOff + ca.offBase(e);
      const uint8_t* q = ca.quality + tb;
      const uint32_t cap = ca.caps.maxTuplesMult;
      const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
      if (threadIdx.x < 8) {
        binCount[threadIdx.x] = 0;
        carry[threadIdx.x] = 0;
      }
      __syncthreads();
      auto binOf = [&](uint32_t t) -> int {
        const uint32_t nh = off[t + 1] - off[t];
        return (nh < 3 || nh > 7 || q[t] == kQualityDup) ? -1 : int(nh);
      };
      for (uint32_t t = threadIdx.x; t < nT; t += kEventBlock) {
        const int b = binOf(t);
        if (b >= 0)
          atomicAdd(&binCount[b], 1u);
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int b = 0; b < 8; ++b) {
          s += binCount[b];
          binEnd[b] = s;
        }
        uint32_t* mo = ca.multOff + size_t(e) * 16;
        for (int b = 0; b < 8; ++b)
          mo[b] = min(binEnd[b] - binCount[b], cap);
        mo[8] = min(s, cap);
        ca.cnt[e].nMult = min(s, cap);
        if (s > cap)
          atomicOr(&ca.cnt[e].status, kCaMultiplicityOverflow);
      }
      __syncthreads();
      TupleIndex* mb = ca.multBins + size_t(e) * cap;
      for (uint32_t base = 0; base < nT; base += kEventBlock) {
        const uint32_t t = base + threadIdx.x;
        const int b = t < nT ? binOf(t) : -1;
        uint32_t below = 0;
        for (int bb = 3; bb < 8; ++bb) {
          const uint64_t m = __ballot(b == bb);
          if (lane == 0)
            waveCnt[bb][wid] = __popcll(m);
          if (b == bb)
            below = __popcll(m & ((1ull << lane) - 1));
        }
        __syncthreads();
        if (b >= 0) {
          uint32_t rank = carry[b] + below;
          for (int w = 0; w < wid; ++w)
            rank += waveCnt[b][w];
          const uint32_t pos = binEnd[b] - 1 - rank;
          if (pos < cap)
            mb[pos] = t;
        }
        __syncthreads();
        if (threadIdx.x < 8 && threadIdx.x >= 3) {
          uint32_t s = 0;
          for (int w = 0; w < kEventBlock / 64; ++w)
            s += waveCnt[threadIdx.x][w];
          carry[threadIdx.x] += s;
        }
        __syncthreads();
      }
    }

  }  // namespace

#define PXG_K(NAME, KERNEL, GRID, BLOCK, ...) \
  do {                                           \
    l.t->b(NAME);                                \
    KERNEL<<<GRID, BLOCK, 0, l.s>>>(__VA_ARGS__); \
    l.t->e();                                    \
  } while (0)

  // the doublets and the outer-hit lists: per event in LDS for batches of 8 or more
  // events (one workgroup per event fills the device), with per-cell atomics across
  // the grid for smaller ones
  template <int G>
  void doubletKernels(const CALaunch& l) {
    const dim3 g(l.blocks, l.nEvents);
    PXG_K("doubletCount", doubletCountKernel<G>, g, kCaBlock, *l.p, *l.ca);
    PXG_K("doubletScan", doubletScanKernel, l.nEvents, kEventBlock, *l.p, *l.ca);
    if (l.nEvents >= kListsInLdsMinEvents) {
      PXG_K("doubletFill", (doubletFillKernel<G, false>), g, kCaBlock, *l.p, *l.ca);
      PXG_K("outerHitFill", ohBuildKernel, l.nEvents, kEventBlock, *l.p, *l.ca);
    } else {
      PXG_K("doubletFill", (doubletFillKernel<G, true>), g, kCaBlock, *l.p, *l.ca);
      PXG_K("outerHitScan", ohScanKernel, l.nEvents, kEventBlock, *l.p, *l.ca);
      PXG_K("outerHitFill", ohFillKernel, g, kCaBlock, *l.ca);
    }
  }

  void launchDoublets(const CALaunch& l) {
    switch (l.doubletLanes) {
      case 1: doubletKernels<1>(l); break;
      case 2: doubletKernels<2>(l); break;
      case 4: doubletKernels<4>(l); break;
      case 8: doubletKernels<8>(l); break;
      case 16: doubletKernels<16>(l); break;
      case 32: doubletKernels<32>(l); break;
      default: doubletKernels<64>(l); break;
    }
  }

  void launchConnect(const CALaunch& l) {
    const dim3 g(l.blocks, l.nEvents);
    if (l.connectStage <= 128)
      PXG_K("connect", connectKernel<128>, g, kCaBlock, *l.p, *l.ca);
    else if (l.connectStage <= 256)
      PXG_K("connect", connectKernel<256>, g, kCaBlock, *l.p, *l.ca);
    else
      PXG_K("connect", connectKernel<512>, g, kCaBlock, *l.p, *l.ca);
  }

  template <bool kCheckTrack>
  void fishboneLaunch(const CALaunch& l, const char* name) {
    const dim3 g(l.blocks, l.nEvents);
    l.t->b(name);
    if (l.fishboneWindow <= 128)
      fishboneKernel<kCheckTrack, 128><<<g, kCaBlock, 0, l.s>>>(*l.p, *l.ca);
    else if (l.fishboneWindow <= 256)
      fishboneKernel<kCheckTrack, 256><<<g, kCaBlock, 0, l.s>>>(*l.p, *l.ca);
    else
      fishboneKernel<kCheckTrack, 384><<<g, kCaBlock, 0, l.s>>>(*l.p, *l.ca);
    l.t->e();
  }

  void launchFishbone(const CALaunch& l, bool checkTrack) {
    if (checkTrack)
      fishboneLaunch<true>(l, "lateFishbone");
    else
      fishboneLaunch<false>(l, "fishbone");
  }

  void launchNtuplets(const CALaunch& l) {
    const dim3 g(l.blocks, l.nEvents);
    PXG_K("startMask", startMaskKernel, g, kCaBlock, *l.ca);
    PXG_K("startCompact", startCompactKernel, l.nEvents, kEventBlock, *l.ca);
    PXG_K("ntupletCount", ntupletCountKernel, g, kCaBlock, *l.ca);
    PXG_K("ntupletScan", ntupletScanKernel, l.nEvents, kEventBlock, *l.ca);
    PXG_K("ntupletFill", ntupletFillKernel, g, kCaBlock, *l.ca);
    PXG_K("trackListFill", trackListBuildKernel, dim3((l.ca->caps.maxDoublets + kTrGroupCells - 1) / kTrGroupCells, l.nEvents),
          kEventBlock, *l.p, *l.ca);
    PXG_K("trackListSort", trackListSortKernel, g, kCaBlock, *l.ca);
  }

  void launchMultiplicity(const CALaunch& l) {
    PXG_K("multiplicity", multiplicityKernel, l.nEvents, kEventBlock, *l.p, *l.ca);
  }

}  // namespace pxg

// Stage B: clusters -> rec-hits, layer offsets and the per-layer phi histogram.
//
//   hitsKernel      G lanes per (module run, event): 16 for batches (a module
//                   holds ~30 digis and ~7 clusters at PU50, so a wavefront per
//                   run would leave most lanes idle and make four times as many
//                   latency chains), 64 for one event (more workgroups to fill
//                   the device), one cluster per lane for the cluster position
//                   estimate and the local -> global transform.  A run of at
//                   most 256 digis (all but a few at PU50) takes its clusters'
//                   extents and integer charges from the records the clustering
//                   wrote (ClusterExt); a larger run accumulates them from its
//                   digis with LDS atomics (exact in any order), G clusters per
//                   pass.  The float
//                   arithmetic keeps the serial program's operation order and
//                   the library is compiled without FMA contraction, so the
//                   positions are bit-identical to the serial ones.
//   phiBinnerKernel one workgroup per event, one wavefront per layer: layer
//                   offsets, per-layer counts of the 128 phi bins, a workgroup
//                   scan, then the fill.  The serial fill decrements the bin
//                   counters, which lists every bin in decreasing hit index;
//                   each wavefront reproduces that order deterministically by
//                   ranking equal bins inside a 64-hit chunk with 7 ballots.
//                   Beside each hit index it writes what the doublet search
//                   tests of the hit (candPhi: z, r, iphi, ysize, detInd in 16
//                   bytes), so that search reads a phi window's candidates
//                   from contiguous memory, one load each.
//
// Reference behaviour: plugin-SiPixelRecHits/gpuPixelRecHits.h:16-207,
// CondFormats/pixelCPEforGPU.h:106-340, DataFormats/SOARotation.h:55-94,
// DataFormats/approx_atan2.h:230-263, PixelRecHits.cc:15-48,
// CUDACore/HistoContainer.h:18-73.
#include "primitives.h"
#include "pxgpu_internal.h"

namespace pxg {

  namespace {
    constexpr int kBlock = 256;

    // orders the wavefront's LDS accesses (lanes read what other lanes wrote)
    __device__ __forceinline__ void waveSync() {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    __device__ __forceinline__ uint32_t div52(uint32_t n) { return n / 52; }
    __device__ __forceinline__ bool bigX(uint32_t px) { return px == 79 || px == 80; }
    __device__ __forceinline__ bool bigY(uint32_t py) {
      const uint32_t ly = py - 52 * div52(py);
      return ly == 0 || ly == 51;
    }
    __device__ __forceinline__ uint32_t unitX(uint32_t px) { return px + (px > 79 ? 1 : 0) + (px > 80 ? 1 : 0); }
    __device__ __forceinline__ uint32_t unitY(uint32_t py) {
      const uint32_t roc = div52(py);
      return py + 2 * roc + ((py - 52 * roc) > 0 ? 1 : 0);
    }

    __device__ __forceinline__ float edgeCorrection(int sizeM1, int qf, int ql, uint32_t upperFirst, uint32_t lowerLast,
                                                    float lorentz, float thickness, float cotAngle, float pitch,
                                                    bool firstBig, bool lastBig) {
      if (sizeM1 == 0)
        return 0.f;
      float weff = 0.f;
      bool simple = true;
      if (sizeM1 == 1) {
        const float winner = pitch * float(int(lowerLast) - int(upperFirst));
        const float wpred = thickness * cotAngle - lorentz;
        weff = fabsf(wpred) - winner;
        simple = (weff < 0.0f) | (weff > pitch);
      }
      if (simple) {
        float edges = 2.0f;
        if (firstBig)
          edges += 1.0f;
        if (lastBig)
          edges += 1.0f;
        weff = pitch * 0.5f * edges;
      }
      const float qdiff = float(ql - qf);
      float qsum = float(ql + qf);
      if (qsum == 0.f)
        qsum = 1.0f;
      return 0.5f * (qdiff / qsum) * weff;
    }

    __device__ __forceinline__ int16_t atan2s7(float y, float x) {
      float r = (fabsf(x) - fabsf(y)) / (fabsf(x) + fabsf(y));
      if (x < 0)
        r = -r;
      const float z = r * r;
      const float poly =
          r * ((-10422.177734375f) + z * (3349.97412109375f + z * ((-1525.589599609375f) + z * 406.64190673828125f)));
      int angle = (x >= 0) ? 8192 : 24576;
      angle = int16_t(angle + int(int16_t(poly)));
      return int16_t(y < 0 ? -angle : angle);
    }
  }  // namespace

  // eight waves per SIMD (64 registers, no spills): the kernel waits on loads far more
  // than it computes (0.167 → 0.155 ms per 167-event launch against five)
  template <int G>
  __global__ void __launch_bounds__(kBlock, 8) hitsKernel(DevConditions c, Pool p) {
    constexpr int kGroup = G;              // lanes per module run
    constexpr int kRuns = kBlock / kGroup;  // module runs per workgroup
    constexpr int kChunk = kGroup;          // clusters per pass of a run
    const uint32_t lane = threadIdx.x % kGroup, run = threadIdx.x / kGroup;
    const LogicalBlock lb = logicalBlock();
    const uint32_t k = lb.x * kRuns + run;
    const uint32_t e = lb.y;
    const EventDesc& ev = p.events[e];
    const BeamSpotPOD bs = ev.bs;
    const uint32_t nModules = p.counts[e].nModules;

    if (k == 0) {  // average geometry corrected for the beam spot
      const AverageGeometry& ag = *c.cpeAverage;
      AverageGeometry& agc = p.avg[e];
      for (uint32_t il = lane; il < kLaddersInBarrel; il += kGroup) {
        const float lx = ag.ladderX[il] - bs.x, ly = ag.ladderY[il] - bs.y;
        agc.ladderZ[il] = ag.ladderZ[il] - bs.z;
        agc.ladderX[il] = lx;
        agc.ladderY[il] = ly;
        agc.ladderR[il] = sqrtf(lx * lx + ly * ly);
        agc.ladderMinZ[il] = ag.ladderMinZ[il] - bs.z;
        agc.ladderMaxZ[il] = ag.ladderMaxZ[il] - bs.z;
      }
      if (lane == 0) {
        agc.endCapZ[0] = ag.endCapZ[0] - bs.z;
        agc.endCapZ[1] = ag.endCapZ[1] - bs.z;
      }
    }
    if (k >= nModules)
      return;

    const size_t mo = size_t(e) * kMaxModuleRanges;
    const uint32_t me = p.moduleId[mo + k];
    if (me >= kNumModules)
      return;
    const uint32_t nclus = p.clusInModule[mo + me];
    if (nclus == 0)
      return;
    const uint32_t first = p.moduleStart[mo + 1 + k];
    const uint32_t end = (k + 1 < nModules) ? p.moduleStart[mo + 2 + k] : ev.nWords;
    const uint32_t hitBase = p.clusModuleStart[mo + me];
    const size_t ho = size_t(e) * kMaxNumClusters;
    const CPEDetParams& det = c.cpeDet[me];
    const CPECommonParams com = *c.cpeCommon;

    // the hit of the run's cluster ic from its extent and charges
    auto emit = [&](uint32_t ic, uint32_t r0, uint32_t r1, uint32_t c0, uint32_t c1, int32_t fx, int32_t lx,
                    int32_t fy, int32_t ly, int32_t qs) {
      const uint32_t h = hitBase + ic;
      if (h >= kMaxNumClusters)
        return;
      // position (pixelCPEforGPU.h:176-259)
      const uint32_t llxl = unitX(r0 + 1), llyl = unitY(c0 + 1), urxl = unitX(r1), uryl = unitY(c1);
      const int mx = int(llxl + urxl), my = int(llyl + uryl);
      int xsize = int(urxl) + 2 - int(llxl);
      int ysize = int(uryl) + 2 - int(llyl);
      if (bigX(r0))
        ++xsize;
      if (bigX(r1))
        ++xsize;
      if (bigY(c0))
        ++ysize;
      if (bigY(c1))
        ++ysize;
      const int unbX = int(8. * double(fabsf(float(fx - lx))) / double(float(fx + lx)));
      const int unbY = int(8. * double(fabsf(float(fy - ly))) / double(float(fy + ly)));
      xsize = 8 * xsize - unbX;
      ysize = 8 * ysize - unbY;
      int16_t sx = int16_t(min(xsize, 1023));
      int16_t sy = int16_t(min(ysize, 1023));
      if (r0 == 0 || r1 == kLastRow)
        sx = int16_t(-sx);
      if (c0 == 0 || c1 == kLastCol)
        sy = int16_t(-sy);
      const float xPos = det.shiftX + com.thePitchX * (0.5f * float(mx) + float(kXOffset));
      const float yPos = det.shiftY + com.thePitchY * (0.5f * float(my) + float(kYOffset));
      const float gvz = -1.f / det.z0;
      const float cotalpha = (xPos - det.x0) * gvz;
      const float cotbeta = (yPos - det.y0) * gvz;
      const float thickness = det.isBarrel ? com.theThicknessB : com.theThicknessE;
      const float xcorr = edgeCorrection(int(r1) - int(r0), fx, lx, llxl, urxl, det.chargeWidthX, thickness, cotalpha,
                                         com.thePitchX, bigX(r0), bigX(r1));
      const float ycorr = edgeCorrection(int(c1) - int(c0), fy, ly, llyl, uryl, det.chargeWidthY, thickness, cotbeta,
                                         com.thePitchY, bigY(c0), bigY(c1));
      const float xl = xPos + xcorr;
      const float yl = yPos + ycorr;
      // errors (pixelCPEforGPU.h:316-340)
      const uint32_t dsx = r1 - r0, dsy = c1 - c0;
      float xerr = 0.0050f, yerr = 0.0085f;
      const uint32_t ix = (dsx == 0) + ((dsx == 0) && bigX(r0));
      const uint32_t iy = (dsy == 0) + ((dsy == 0) && bigY(c0));
      if (!(r0 == 0 || r1 == kLastRow))
        xerr = det.sx[ix];
      if (!(c0 == 0 || c1 == kLastCol))
        yerr = det.sy[iy];
      // local -> global, beam spot subtracted
      const SOAFrameF& f = det.frame;
      float xg = f.R11 * xl + f.R21 * yl;
      float yg = f.R12 * xl + f.R22 * yl;
      float zg = f.R13 * xl + f.R23 * yl;
      xg += f.px;
      yg += f.py;
      zg += f.pz;
      xg -= bs.x;
      yg -= bs.y;
      zg -= bs.z;
      const size_t o = ho + h;
      p.charge[o] = qs;
      p.detInd[o] = uint16_t(me);
      p.xl[o] = xl;
      p.yl[o] = yl;
      p.xsize[o] = sx;
      p.ysize[o] = sy;
      p.xerr[o] = xerr * xerr;
      p.yerr[o] = yerr * yerr;
      p.xg[o] = xg;
      p.yg[o] = yg;
      p.zg[o] = zg;
      const float rg = sqrtf(xg * xg + yg * yg);
      p.rg[o] = rg;
      p.pos[o] = make_float4(xg, yg, zg, rg);
      const int16_t iphi = atan2s7(yg, xg);
      p.iphi[o] = iphi;
      p.attr[o] = HitAttr{iphi, int16_t(sy), uint16_t(me), 0};
    };

    const uint32_t wo = ev.wordOffset;
    if (end - first <= 256 && p.clExt) {
      // a run of at most 256 digis: the clustering recorded its clusters' extents
      const ClusterExt* ext = p.clExt + wo + first;
      for (uint32_t ic = lane; ic < nclus; ic += kGroup) {
        const ClusterExt x = ext[ic];
        emit(ic, x.minRow, x.maxRow, x.minCol, x.maxCol, x.qfx, x.qlx, x.qfy, x.qly, x.qsum);
      }
      return;
    }
    // a larger run: cluster extents and charges from the digis, G clusters per pass
    __shared__ uint32_t sMinRow[kRuns][kChunk], sMaxRow[kRuns][kChunk], sMinCol[kRuns][kChunk],
        sMaxCol[kRuns][kChunk];
    __shared__ int32_t sQsum[kRuns][kChunk], sQfx[kRuns][kChunk], sQlx[kRuns][kChunk], sQfy[kRuns][kChunk],
        sQly[kRuns][kChunk];
    uint32_t *minRow = sMinRow[run], *maxRow = sMaxRow[run], *minCol = sMinCol[run], *maxCol = sMaxCol[run];
    int32_t *qsum = sQsum[run], *qfx = sQfx[run], *qlx = sQlx[run], *qfy = sQfy[run], *qly = sQly[run];

    for (uint32_t start = 0; start < nclus; start += kChunk) {
      const uint32_t nIn = min(uint32_t(kChunk), nclus - start);
      if (lane < nIn) {
        minRow[lane] = minCol[lane] = 0xffffffffu;
        maxRow[lane] = maxCol[lane] = 0;
        qsum[lane] = qfx[lane] = qlx[lane] = qfy[lane] = qly[lane] = 0;
      }
      waveSync();
      for (uint32_t i = first + lane; i < end; i += kGroup) {
        const uint32_t g = wo + i;
        if (p.moduleInd[g] != me)
          continue;  // invalid (or dropped) pixel
        const int32_t cl = p.clus[g];
        if (cl < int32_t(start) || cl >= int32_t(start + nIn))
          continue;
        const uint32_t j = uint32_t(cl) - start;
        const uint32_t x = p.xx[g], y = p.yy[g];
        atomicMin(&minRow[j], x);
        atomicMax(&maxRow[j], x);
        atomicMin(&minCol[j], y);
        atomicMax(&maxCol[j], y);
      }
      waveSync();
      for (uint32_t i = first + lane; i < end; i += kGroup) {
        const uint32_t g = wo + i;
        if (p.moduleInd[g] != me)
          continue;
        const int32_t cl = p.clus[g];
        if (cl < int32_t(start) || cl >= int32_t(start + nIn))
          continue;
        const uint32_t j = uint32_t(cl) - start;
        const uint32_t x = p.xx[g], y = p.yy[g];
        const int32_t ch = p.adc[g];  // the 65535 clamp of the reference is a no-op on uint16
        atomicAdd(&qsum[j], ch);
        if (minRow[j] == x)
          atomicAdd(&qfx[j], ch);
        if (maxRow[j] == x)
          atomicAdd(&qlx[j], ch);
        if (minCol[j] == y)
          atomicAdd(&qfy[j], ch);
        if (maxCol[j] == y)
          atomicAdd(&qly[j], ch);
      }
      waveSync();
      for (uint32_t ic = lane; ic < nIn; ic += kGroup) {
        if (hitBase + start + ic >= kMaxNumClusters)
          break;
        emit(start + ic, minRow[ic], maxRow[ic], minCol[ic], maxCol[ic], qfx[ic], qlx[ic], qfy[ic], qly[ic],
             qsum[ic]);
      }
      waveSync();
    }
  }

  // ---------------------------------------------------------------------------
  namespace {
    constexpr int kPhiWaves = kNumLayers;  // one wavefront per layer
  }

  __global__ void __launch_bounds__(64 * kPhiWaves) phiBinnerKernel(DevConditions c, Pool p) {
    const uint32_t e = blockIdx.x;
    const uint32_t nHits = p.counts[e].nClusters;
    const uint32_t* cms = p.clusModuleStart + size_t(e) * kMaxModuleRanges;
    uint32_t* hls = p.hitsLayerStart + size_t(e) * kLayerStartStride;
    uint32_t* off = p.phiOff + size_t(e) * kPhiOffStride;
    uint16_t* bins = p.phiBins + size_t(e) * kMaxNumClusters;
    const int16_t* iphi = p.iphi + size_t(e) * kMaxNumClusters;
    const float4* pos = p.pos + size_t(e) * kMaxNumClusters;
    const HitAttr* attr = p.attr + size_t(e) * kMaxNumClusters;
    PhiCand* candPhi = p.candPhi + size_t(e) * kMaxNumClusters;

    __shared__ uint32_t ls[kNumLayers + 1];
    __shared__ uint32_t cnt[kPhiHistTotBins + 1];
    __shared__ uint32_t wsum[kPhiWaves];
    if (threadIdx.x <= kNumLayers)
      ls[threadIdx.x] = nHits ? cms[c.cpeLayer->layerStart[threadIdx.x]] : 0;
    for (uint32_t i = threadIdx.x; i <= kPhiHistTotBins; i += blockDim.x)
      cnt[i] = 0;
    __syncthreads();
    if (threadIdx.x <= kNumLayers)
      hls[threadIdx.x] = ls[threadIdx.x];
    if (nHits == 0) {
      for (uint32_t i = threadIdx.x; i < kPhiHistTotBins; i += blockDim.x)
        off[i] = 0;
      return;
    }
    const uint32_t lane = threadIdx.x & 63, layer = threadIdx.x >> 6;
    const uint32_t h0 = ls[layer], h1 = ls[layer + 1];
    for (uint32_t i = h0 + lane; i < h1; i += 64)
      atomicAdd(&cnt[layer * kPhiBins + ((int(iphi[i]) >> 9) & 127)], 1u);
    __syncthreads();
    // inclusive scan of the 1281 counters: 640 threads x 2 consecutive bins (+ the last)
    {
      const uint32_t b0 = 2 * threadIdx.x;
      const uint32_t v0 = cnt[b0], v1 = cnt[b0 + 1];
      const uint32_t s = v0 + v1;
      uint32_t tot;
      const uint32_t excl = wgExclusiveScan<64 * kPhiWaves>(s, tot, wsum);
      cnt[b0] = excl + v0;       // inclusive: end of bin b0
      cnt[b0 + 1] = excl + s;    // end of bin b0 + 1
      // bin 1280 has no entries: its end is the total
    }
    __syncthreads();
    if (threadIdx.x == 0)
      cnt[kPhiHistTotBins - 1] = cnt[kPhiHistTotBins - 2];
    __syncthreads();
    // per wave: fill its layer in 64-hit chunks, ranking equal bins by lane
    __shared__ uint32_t taken[kPhiHistTotBins];
    for (uint32_t b = lane; b < kPhiBins; b += 64)
      taken[layer * kPhiBins + b] = 0;
    __builtin_amdgcn_wave_barrier();
    // the chunk's inputs are loaded one chunk ahead, so the loads overlap the
    // ranking of the previous chunk
    int16_t ipNext = 0;
    float4 posNext{};
    HitAttr attrNext{};
    if (h0 + lane < h1) {
      ipNext = iphi[h0 + lane];
      posNext = pos[h0 + lane];
      attrNext = attr[h0 + lane];
    }
    for (uint32_t c0 = h0; c0 < h1; c0 += 64) {
      const uint32_t i = c0 + lane;
      const bool act = i < h1;
      const int16_t ip = ipNext;
      const float4 pv = posNext;
      const HitAttr av = attrNext;
      if (i + 64 < h1) {
        ipNext = iphi[i + 64];
        posNext = pos[i + 64];
        attrNext = attr[i + 64];
      }
      const uint32_t key = act ? ((int(ip) >> 9) & 127) : 0;
      uint64_t peers = __ballot(act);
#pragma unroll
      for (int bit = 0; bit < 7; ++bit) {
        const bool set = (key >> bit) & 1;
        const uint64_t m = __ballot(set);
        peers &= set ? m : ~m;
      }
      const uint32_t rank = __popcll(peers & ((1ull << lane) - 1ull));
      const uint32_t group = __popcll(peers);
      const uint32_t bin = layer * kPhiBins + key;
      uint32_t before = 0;
      if (act)
        before = taken[bin];
      __builtin_amdgcn_wave_barrier();
      if (act) {
        const uint32_t k = cnt[bin] - 1 - (before + rank);
        bins[k] = uint16_t(i);
        candPhi[k] = PhiCand{pv.z, pv.w, av.iphi, av.ysize, av.detInd, 0};
        if (rank == group - 1)
          taken[bin] = before + group;
      }
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // final offsets = bin starts (what the serial decrementing fill leaves)
    for (uint32_t b = threadIdx.x; b < kPhiHistTotBins; b += blockDim.x)
      off[b] = b == 0 ? 0 : cnt[b - 1];
  }

  void launchHits(const DevConditions& c, const Pool& p, uint32_t nEvents, hipStream_t s) {
    if (nEvents < 8)
      hitsKernel<64><<<dim3((kMaxNumModules + kBlock / 64 - 1) / (kBlock / 64), nEvents), kBlock, 0, s>>>(c, p);
    else
      hitsKernel<16><<<dim3((kMaxNumModules + kBlock / 16 - 1) / (kBlock / 16), nEvents), kBlock, 0, s>>>(c, p);
  }

  void launchPhiBinner(const DevConditions& c, const Pool& p, uint32_t nEvents, hipStream_t s) {
    phiBinnerKernel<<<nEvents, 64 * kPhiWaves, 0, s>>>(c, p);
  }

}  // namespace pxg

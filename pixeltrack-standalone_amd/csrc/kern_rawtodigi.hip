// Stage A, part 1: raw FED words -> calibrated digis, module boundaries and
// ordered error lists.
//
//   decodeKernel       one lane per 32-bit FED word (coalesced loads): field
//                      extraction, cabling lookup, error classification, ROC ->
//                      module frame conversion, gain/pedestal calibration, and
//                      the "new module starts here" flag of every digi.  Flags
//                      and error lanes leave the kernel as one 64-bit ballot
//                      mask per wavefront, so no atomics order anything.
//   moduleStartKernel  one workgroup per event: popcount + workgroup scan of the
//                      ballot masks turns them into the ordered moduleStart list
//                      and the ordered error list (word order, as the serial
//                      program produces them), and clears the per-event
//                      cluster arrays for the clusteriser.
//
// Reference behaviour: SiPixelRawToClusterGPUKernel.cc:51-462 (decode),
// gpuCalibPixel.h:22-66 (calibration), gpuClustering.h:19-37 (module starts:
// a valid digi starts a module when the previous valid digi of the event has a
// different module id).
#include "primitives.h"
#include "pxgpu_internal.h"

namespace pxg {

  namespace {

    constexpr uint32_t kSlotsPerFed = kMaxLink * kMaxRoc;  // 384

    struct Digi {
      uint32_t row, col, adcRaw, moduleId, rawId;
    };

    // 0: no digi; 1: digi; 2: error record
    __device__ __forceinline__ int decodeWord(const DevConditions& c, uint32_t w, uint32_t fed, Digi& d,
                                              PixelErrorCompact& e) {
      if (w == 0)
        return 0;
      const SiPixelFedCablingMapGPU& m = *c.cabling;
      const uint32_t link = (w >> kLinkShift) & kLinkMask;
      const uint32_t roc = (w >> kRocShift) & kRocMask;
      // link 0 or a slot past the table is out of the cabling map: the reference
      // reads out of bounds there; this decoder drops the word instead
      const uint32_t slot = fed * kSlotsPerFed + (link - 1) * kMaxRoc + roc;
      if (link == 0 || fed >= kMaxFed)
        return 0;
      if (roc >= kMaxRoc) {
        // ROC numbers 25..31 are error words (checkROC)
        uint32_t type = 0;
        if (roc >= 25) {
          switch (roc) {
            case 25: {
              uint32_t i = fed * kSlotsPerFed + (link - 1) * kMaxRoc + 1;
              bool found = !(i > 1 && i <= m.size && !(link == m.link[i] && 1 == m.roc[i]));
              type = found ? 25 : 0;
              break;
            }
            case 30:
              type = (((w >> 8) & 0xF) == 1) ? 40 : 30;
              break;
            default:
              type = roc;  // 26, 27, 28, 29, 31
          }
        }
        if (type != 0) {
          // raw id of the error (getErrRawID)
          uint32_t rid = 0xffffffffu;
          uint32_t eLink = 0, eRoc = 1;
          bool look = false;  // types 26, 27, 28 carry no module (getErrRawID's default)
          if (type == 29) {
            int ch = int(w & 1) + 2 * int((w >> 1) & 1) + 3 * int((w >> 2) & 1) + 4 * int((w >> 3) & 1) +
                     5 * int((w >> 4) & 1);
            int block = int((w >> 8) & 7);
            int chan = (block & 1) == 0 ? (block / 2) * 9 + ch : ((block - 1) / 2) * 9 + 4 + ch;
            look = chan >= 1 && chan <= 36;
            eLink = uint32_t(chan);
          } else if (type == 25 || type == 30 || type == 31 || type == 40) {
            eLink = link;
            look = true;
          }
          const uint32_t eSlot = fed * kSlotsPerFed + (eLink - 1) * kMaxRoc + eRoc;
          if (look && eSlot < kCablingSize) {
            uint32_t r = m.RawId[eSlot];
            if (r != 9999)
              rid = r;
          }
          e = {rid, w, uint8_t(type), uint8_t(fed)};
          return 2;
        }
      }
      if (slot >= kCablingSize)
        return 0;
      const uint32_t rawId = m.RawId[slot];
      const uint32_t rocInDet = m.rocInDet[slot];
      if (m.badRocs[slot] || c.modToUnp[slot])
        return 0;
      const bool barrel = ((rawId >> 25) & 7) == 1;
      uint32_t layer = 0;
      bool flipped;  // the '-z side / panel' orientation of the frame conversion
      if (barrel) {
        layer = (rawId >> 20) & 0xF;
        const uint32_t module = (rawId >> 2) & 0x3FF;
        flipped = module < 5 && layer != 1;
      } else {
        flipped = true;  // both forward panels share one orientation
      }
      uint32_t lrow, lcol;
      if (layer == 1) {
        lcol = (w >> kColShiftL1) & kColMaskL1;
        lrow = (w >> kRowShiftL1) & kRowMaskL1;
        if (!(lrow < 80 && lcol < 52)) {
          e = {rawId, w, 37, uint8_t(fed)};
          return 2;
        }
      } else {
        const uint32_t dcol = (w >> kDcolShift) & kDcolMask;
        const uint32_t pxid = (w >> kPxidShift) & kPxidMask;
        lrow = 80 - pxid / 2;
        lcol = dcol * 2 + (pxid & 1);
        if (!(dcol < 26 && pxid >= 2 && pxid < 162)) {
          e = {rawId, w, 37, uint8_t(fed)};
          return 2;
        }
      }
      // ROC -> module pixel frame
      int row, col;
      const int r = int(rocInDet);
      if (flipped) {
        if (r < 8) {
          row = int(lrow);
          col = (8 - r) * 52 - 1 - int(lcol);
        } else {
          row = 159 - int(lrow);
          col = (r - 8) * 52 + int(lcol);
        }
      } else {
        if (r < 8) {
          row = 159 - int(lrow);
          col = r * 52 + int(lcol);
        } else {
          row = int(lrow);
          col = (16 - r) * 52 - 1 - int(lcol);
        }
      }
      d.row = uint32_t(row);
      d.col = uint32_t(col);
      d.adcRaw = w & kAdcMask;
      d.moduleId = m.moduleId[slot];
      d.rawId = rawId;
      return 1;
    }

    // gain/pedestal calibration; returns false for dead or noisy columns.
    // The arithmetic is kept in the serial program's order with no FMA
    // contraction (the library is compiled with -ffp-contract=off).
    __device__ __forceinline__ bool calibrate(const DevConditions& c, uint32_t id, uint32_t row, uint32_t col,
                                              uint32_t adcRaw, uint16_t& adcOut) {
      const SiPixelGainForHLTonGPU& g = *c.gain;
      if (id >= kMaxNumModules || row >= kRowsPerModule || col >= kColsPerModule)
        return false;  // outside the calibration tables (the reference reads out of bounds)
      const GainRangeAndCols rc = g.rangeAndCols[id];
      if (rc.nCols <= 0)
        return false;
      const uint32_t colBytes = (rc.second - rc.first) / uint32_t(rc.nCols);
      const uint32_t offset = rc.first + col * colBytes + 2 * (row / g.numberOfRowsAveragedOver_);
      if (offset + 2 > c.gainDataBytes)
        return false;
      const GainDecoding s = c.gainData[offset / 2];
      if (s.ped == g.deadFlag_ || s.ped == g.noisyFlag_)
        return false;
      const float ped = float(s.ped) * g.pedPrecision + g.minPed_;
      const float gain = float(s.gain) * g.gainPrecision + g.minGain_;
      const bool l1 = id < 96;
      const float conv = l1 ? 50.f : 47.f;
      const float offs = l1 ? -670.f : -60.f;
      const float vcal = float(adcRaw) * gain - ped * gain;
      const int e = int(vcal * conv + offs);
      adcOut = uint16_t(e > 100 ? e : 100);
      return true;
    }

    // module id of word g after calibration, or kInvId (used for the look-back)
    __device__ uint32_t validModuleAt(const DevConditions& c, const uint32_t* words, const uint8_t* fedIds,
                                      uint32_t g) {
      Digi d;
      PixelErrorCompact e;
      if (decodeWord(c, words[g], fedIds[g >> 1], d, e) != 1)
        return kInvId;
      uint16_t a;
      if (d.moduleId == kInvId || !calibrate(c, d.moduleId, d.row, d.col, d.adcRaw, a))
        return kInvId;
      return d.moduleId;
    }

  }  // namespace

  // a (blocks per event, events) grid, an event's blocks on one XCD: the event of a
  // workgroup is its grid row, with no search over the events' block offsets
  __global__ void __launch_bounds__(kDecodeBlock) decodeKernel(DevConditions c, Pool p) {
    const LogicalBlock lb = logicalBlock();
    const uint32_t e = lb.y;
    const uint32_t wordOffset = p.events[e].wordOffset;
    const uint32_t nWords = p.events[e].nWords;
    const uint32_t local = lb.x * kDecodeBlock + threadIdx.x;
    if (lb.x * kDecodeBlock >= nWords)
      return;  // uniform per workgroup: the event is shorter than the longest of the launch
    const uint32_t g = wordOffset + local;
    const uint32_t lane = threadIdx.x & 63;
    const bool active = local < nWords;

    Digi d;
    PixelErrorCompact err;
    int kind = 0;
    uint32_t id = kInvId;
    uint16_t adc = 0;
    if (active) {
      const uint32_t w = p.words[g];
      const uint32_t fed = p.fedIds[g >> 1];
      kind = decodeWord(c, w, fed, d, err);
      uint16_t xx = 0, yy = 0, moduleInd = kInvId;
      uint32_t pdigi = 0, rawId = 0;
      if (kind == 1) {
        xx = uint16_t(d.row);
        yy = uint16_t(d.col);
        adc = uint16_t(d.adcRaw);
        pdigi = (uint32_t(xx)) | (uint32_t(yy) << 11) | (min(d.adcRaw, 1023u) << 22);
        rawId = d.rawId;
        moduleInd = uint16_t(d.moduleId);
        if (moduleInd != kInvId) {
          if (calibrate(c, d.moduleId, d.row, d.col, d.adcRaw, adc)) {
            id = moduleInd;
          } else {
            moduleInd = kInvId;
            adc = 0;
          }
        }
      }
      p.xx[g] = xx;
      p.yy[g] = yy;
      p.adc[g] = adc;
      p.moduleInd[g] = moduleInd;
      p.pdigi[g] = pdigi;
      p.rawIdArr[g] = rawId;
      p.clus[g] = int32_t(local);
      if (kind == 2)
        p.errorsSparse[g] = err;  // compacted in word order by moduleStartKernel
    }

    // --- module-start flags: valid digi whose previous valid digi differs
    const bool valid = id != kInvId;
    const uint64_t vmask = __ballot(valid);
    const uint64_t below = vmask & ((1ull << lane) - 1ull);
    const int prevLane = below ? 63 - __clzll(below) : -1;
    const uint32_t prevId = __shfl(id, prevLane < 0 ? 0 : prevLane);
    // the first valid lane of the wave looks back across the wave start
    uint32_t carry = kInvId;
    const uint32_t firstValid = vmask ? __ffsll((unsigned long long)vmask) - 1 : 64;
    if (lane == firstValid) {
      const uint32_t waveStart = local - lane;  // local index of lane 0
      for (uint32_t j = waveStart; j > 0; --j) {
        carry = validModuleAt(c, p.words, p.fedIds, wordOffset + j - 1);
        if (carry != kInvId)
          break;
      }
    }
    const bool flag = valid && (prevLane >= 0 ? prevId != id : carry != id);
    const uint64_t fmask = __ballot(flag);
    const uint64_t emask = __ballot(kind == 2);
    if (lane == 0) {
      p.boundaryMask[2 * (g >> 6)] = fmask;
      p.boundaryMask[2 * (g >> 6) + 1] = emask;
    }
  }

  // ---------------------------------------------------------------------------
  namespace {
    constexpr int kScanBlock = 1024;
  }  // namespace

  __global__ void __launch_bounds__(kScanBlock) moduleStartKernel(Pool p) {
    __shared__ uint64_t lds[kScanBlock / 64];
    const uint32_t e = blockIdx.x;
    const uint32_t wordOffset = p.events[e].wordOffset;
    const uint32_t nWords = p.events[e].nWords;
    uint32_t* moduleStart = p.moduleStart + size_t(e) * kMaxModuleRanges;
    uint32_t* clusInModule = p.clusInModule + size_t(e) * kMaxModuleRanges;
    uint32_t* moduleId = p.moduleId + size_t(e) * kMaxModuleRanges;
    uint32_t* clusModuleStart = p.clusModuleStart + size_t(e) * kMaxModuleRanges;
    PixelErrorCompact* errors = p.errors + wordOffset;
    const PixelErrorCompact* sparse = p.errorsSparse + wordOffset;
    // clear what the clusteriser accumulates into
    for (uint32_t i = threadIdx.x; i < kMaxModuleRanges; i += kScanBlock) {
      clusInModule[i] = 0;
      moduleId[i] = 0;
      clusModuleStart[i] = 0;
    }
    const uint32_t nMasks = (nWords + 63) / 64;
    uint32_t baseMod = 0, baseErr = 0;
    for (uint32_t c0 = 0; c0 < nMasks; c0 += kScanBlock) {
      const uint32_t k = c0 + threadIdx.x;
      uint64_t fm = 0, em = 0;
      if (k < nMasks) {
        fm = p.boundaryMask[2 * (wordOffset / 64 + k)];
        em = p.boundaryMask[2 * (wordOffset / 64 + k) + 1];
      }
      const uint64_t packed = uint64_t(__popcll(fm)) | (uint64_t(__popcll(em)) << 32);
      uint64_t total;
      const uint64_t excl = wgExclusiveScan<kScanBlock>(packed, total, lds);
      uint32_t r = baseMod + uint32_t(excl & 0xffffffffu);
      while (fm) {
        const int b = __ffsll((unsigned long long)fm) - 1;
        fm &= fm - 1;
        if (r < kMaxNumModules)
          moduleStart[1 + r] = k * 64 + b;
        ++r;
      }
      uint32_t re = baseErr + uint32_t(excl >> 32);
      while (em) {
        const int b = __ffsll((unsigned long long)em) - 1;
        em &= em - 1;
        errors[re++] = sparse[k * 64 + b];
      }
      baseMod += uint32_t(total & 0xffffffffu);
      baseErr += uint32_t(total >> 32);
    }
    // the runs the clusteriser's wavefront kernel cannot take, in order: those of 65 to
    // 256 digis from the front of the event's list, the larger ones from its back
    __syncthreads();
    const uint32_t nMod = baseMod < kMaxNumModules ? baseMod : kMaxNumModules;
    uint32_t* bigList = p.bigList + size_t(e) * kMaxModuleRanges;
    uint32_t nMed = 0, nBig = 0;
    for (uint32_t c0 = 0; c0 < nMod; c0 += kScanBlock) {
      const uint32_t k = c0 + threadIdx.x;
      uint32_t n = 0;
      if (k < nMod)
        n = (k + 1 < nMod ? moduleStart[2 + k] : nWords) - moduleStart[1 + k];
      const bool med = n > 64 && n <= 256, big = n > 256;
      uint64_t total;
      const uint64_t at = wgExclusiveScan<kScanBlock>(uint64_t(med) | uint64_t(big) << 32, total, lds);
      if (med)
        bigList[nMed + uint32_t(at)] = k;
      if (big)
        bigList[kMaxModuleRanges - 1 - (nBig + uint32_t(at >> 32))] = k;
      nMed += uint32_t(total);
      nBig += uint32_t(total >> 32);
    }
    if (threadIdx.x == 0) {
      moduleStart[0] = nMod;
      EventCounts& cnt = p.counts[e];
      cnt.nModules = nMod;
      cnt.nErrors = baseErr;
      cnt.nClusters = 0;
      cnt.nMediumRuns = nMed;
      cnt.nBigRuns = nBig;
      cnt.status = baseMod > kMaxNumModules ? kStatusModuleOverflow : 0;
    }
  }

  void launchDecode(const DevConditions& c, const Pool& p, uint32_t nEvents, uint32_t maxWords, hipStream_t s) {
    const uint32_t blocks = (maxWords + kDecodeBlock - 1) / kDecodeBlock;
    if (blocks && nEvents)
      decodeKernel<<<dim3(blocks, nEvents), kDecodeBlock, 0, s>>>(c, p);
  }

  void launchModuleStarts(const Pool& p, uint32_t nEvents, hipStream_t s) {
    moduleStartKernel<<<nEvents, kScanBlock, 0, s>>>(p);
  }

}  // namespace pxg

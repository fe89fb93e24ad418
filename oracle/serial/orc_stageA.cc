// ORACLE -- test infrastructure only (see oracle.h).
// Stage A of the serial chain: FED unpacking, raw-to-digi, calibration and
// clustering, restated from the reference serial program.
#include <algorithm>
#include <array>
#include <cassert>
#include <cstdlib>
#include <cstring>

#include "oracle.h"

namespace orc {

  namespace {
    constexpr uint32_t kLinkRocSlots = kMaxLink * kMaxRoc;  // 384

    inline uint64_t load64(const uint8_t* p) {
      uint64_t v;
      std::memcpy(&v, p, 8);
      return v;
    }
  }  // namespace

  // ---------------------------------------------------------------------------
  // a1: plugin-SiPixelClusterizer/SiPixelRawToClusterCUDA.cc:87-138 with
  // ErrorChecker.cc:38-91 (the checker records no errors: includeErrors is
  // false in its constructor, ErrorChecker.cc:36) and
  // SiPixelRawToClusterGPUKernel.cc:41-47 (WordFedAppender).
  bool unpackFeds(const Conditions& c, const uint8_t* raw, size_t len, Event& ev) {
    // index the raw record: nfeds, then {fedId, nbytes, bytes}
    if (len < 4)
      return false;
    uint32_t nfeds;
    std::memcpy(&nfeds, raw, 4);
    std::vector<std::pair<const uint8_t*, uint32_t>> byFed(4097, {nullptr, 0});
    size_t pos = 4;
    for (uint32_t i = 0; i < nfeds; ++i) {
      if (pos + 8 > len)
        return false;
      uint32_t fed, size;
      std::memcpy(&fed, raw + pos, 4);
      std::memcpy(&size, raw + pos + 4, 4);
      pos += 8;
      if (pos + size > len || fed > 4096)
        return false;
      byFed[fed] = {raw + pos, size};
      pos += size;
    }
    ev.words.assign(kMaxFedWords, 0);
    ev.fedIdx.assign(kMaxFedWords, 0);
    uint32_t wordCounter = 0, fedCounter = 0;
    for (uint32_t k = 0; k < c.nFeds; ++k) {
      uint32_t fedId = c.fedIds[k];
      if (fedId == 40)
        continue;
      assert(fedId >= 1200);
      ++fedCounter;
      const uint8_t* data = byFed[fedId].first;
      uint32_t nWords = byFed[fedId].second / 8;
      if (nWords == 0)
        continue;
      // CRC bit: bit 2 of the last 64-bit word
      const uint8_t* trailer = data + 8 * (nWords - 1);
      if ((load64(trailer) >> 2) & 1)
        continue;
      // headers: marker 0x5 in bits 60-63, "more headers" bit 3
      const uint8_t* header = data - 8;
      bool more = true;
      while (more) {
        header += 8;
        uint64_t h = load64(header);
        bool ok = ((h >> 60) & 0xF) == 0x5;
        more = ok && ((h >> 3) & 1);
      }
      // trailers: marker 0xA in bits 60-63, "more trailers" bit 3
      trailer += 8;
      more = true;
      while (more) {
        trailer -= 8;
        uint64_t t = load64(trailer);
        bool ok = ((t >> 60) & 0xF) == 0xA;
        more = ok && ((t >> 3) & 1);
      }
      const uint8_t* bw = header + 8;
      const uint8_t* ew = trailer;
      uint32_t length = uint32_t((ew - bw) / 4);
      assert(length % 2 == 0);
      if (wordCounter + length > kMaxFedWords)
        return false;
      std::memcpy(ev.words.data() + wordCounter, bw, 4 * size_t(length));
      std::memset(ev.fedIdx.data() + wordCounter / 2, int(fedId - 1200), length / 2);
      wordCounter += length;
    }
    ev.nWords = wordCounter;
    ev.fedCounter = fedCounter;
    ev.words.resize(wordCounter);
    ev.fedIdx.resize(wordCounter / 2);
    return true;
  }

  // ---------------------------------------------------------------------------
  // a2: SiPixelRawToClusterGPUKernel.cc:342-462 (useQualityInfo = includeErrors = true)
  namespace {
    struct DetRef {
      uint32_t rawId, rocInDet, moduleId;
    };
    inline DetRef detFor(const SiPixelFedCablingMapGPU& m, uint32_t fed, uint32_t link, uint32_t roc) {
      uint32_t i = fed * kLinkRocSlots + (link - 1) * kMaxRoc + roc;
      return {m.RawId[i], m.rocInDet[i], m.moduleId[i]};
    }

    // :188-262
    uint8_t rocErrorType(uint32_t w, uint32_t fed, uint32_t link, const SiPixelFedCablingMapGPU& m) {
      uint32_t t = (w >> kRocShift) & kRocMask;
      if (t < 25)
        return 0;
      switch (t) {
        case 25: {
          uint32_t i = fed * kLinkRocSlots + (link - 1) * kMaxRoc + 1;
          bool found = true;
          if (i > 1 && i <= m.size && !(link == m.link[i] && 1 == m.roc[i]))
            found = false;
          return found ? 25 : 0;
        }
        case 26:
        case 27:
        case 28:
        case 29:
        case 31:
          return uint8_t(t);
        case 30: {
          uint32_t stateMatch = (w >> 8) & 0xF;
          return stateMatch == 1 ? 40 : 30;
        }
        default:
          return 0;
      }
    }

    // :264-339
    uint32_t errorRawId(uint32_t fed, uint32_t w, uint32_t type, const SiPixelFedCablingMapGPU& m) {
      uint32_t rid = 0xffffffff;
      auto lookup = [&](uint32_t link, uint32_t roc) {
        uint32_t r = detFor(m, fed, link, roc).rawId;
        if (r != 9999)
          rid = r;
      };
      switch (type) {
        case 25:
        case 30:
        case 31:
        case 36:
        case 40:
          lookup((w >> kLinkShift) & kLinkMask, 1);
          break;
        case 29: {
          int ch = 0;
          for (int b = 0; b < 5; ++b)
            ch += (b + 1) * int((w >> b) & 1);
          int block = int((w >> 8) & 0x7);
          int chan = (block % 2 == 0) ? (block / 2) * 9 + ch : ((block - 1) / 2) * 9 + 4 + ch;
          if (chan < 1 || chan > 36)
            break;
          lookup(uint32_t(chan), 1);
          break;
        }
        case 37:
        case 38:
          lookup((w >> kLinkShift) & kLinkMask, (w >> kRocShift) & kRocMask);
          break;
        default:
          break;
      }
      return rid;
    }

    // :72-138; returns module pixel (row, col) for ROC-local (row, col)
    inline void toModule(bool barrel, int side, uint32_t layer, uint32_t rid, uint32_t lrow, uint32_t lcol,
                         uint32_t& row, uint32_t& col) {
      int sRow, sCol, oRow, oCol;
      bool flippedLike = barrel ? (side == -1 && layer != 1) : true;  // fpix: both panels alike
      if (flippedLike) {
        if (rid < 8) {
          sRow = 1, sCol = -1, oRow = 0, oCol = int(8 - rid) * 52 - 1;
        } else {
          sRow = -1, sCol = 1, oRow = 159, oCol = int(rid - 8) * 52;
        }
      } else {
        if (rid < 8) {
          sRow = -1, sCol = 1, oRow = 159, oCol = int(rid) * 52;
        } else {
          sRow = 1, sCol = -1, oRow = 0, oCol = int(16 - rid) * 52 - 1;
        }
      }
      row = uint32_t(oRow + sRow * int(lrow));
      col = uint32_t(oCol + sCol * int(lcol));
    }
  }  // namespace

  void rawToDigi(const Conditions& c, Event& ev) {
    const auto& m = *c.cabling;
    const uint32_t W = ev.nWords;
    ev.xx.assign(W, 0);
    ev.yy.assign(W, 0);
    ev.adc.assign(W, 0);
    ev.pdigi.assign(W, 0);
    ev.rawIdArr.assign(W, 0);
    ev.moduleInd.assign(W, kInvId);
    ev.clus.assign(W, 0);
    ev.errors.clear();
    for (uint32_t i = 0; i < W; ++i) {
      uint32_t fed = ev.fedIdx[i / 2];
      uint32_t w = ev.words[i];
      if (w == 0)
        continue;
      uint32_t link = (w >> kLinkShift) & kLinkMask;
      uint32_t roc = (w >> kRocShift) & kRocMask;
      DetRef det = detFor(m, fed, link, roc);
      uint8_t etype = rocErrorType(w, fed, link, m);
      bool skip = roc < kMaxRoc ? false : (etype != 0);
      if (skip) {
        ev.errors.push_back({errorRawId(fed, w, etype, m), w, etype, uint8_t(fed)});
        continue;
      }
      uint32_t slot = fed * kLinkRocSlots + (link - 1) * kMaxRoc + roc;
      if (m.badRocs[slot])
        continue;
      if (c.modToUnp[slot])
        continue;
      bool barrel = ((det.rawId >> 25) & 7) == 1;
      uint32_t layer = 0;
      int side;
      if (barrel) {
        layer = (det.rawId >> 20) & 0xF;
        int module = int((det.rawId >> 2) & 0x3FF);
        side = module < 5 ? -1 : 1;
      } else {
        int panel = int((det.rawId >> 10) & 3);
        side = panel == 1 ? -1 : 1;
      }
      uint32_t lrow, lcol;
      if (layer == 1) {
        lcol = (w >> kColShiftL1) & kColMaskL1;
        lrow = (w >> kRowShiftL1) & kRowMaskL1;
        if (!(lrow < 80 && lcol < 52)) {
          ev.errors.push_back({det.rawId, w, 37, uint8_t(fed)});
          continue;
        }
      } else {
        uint32_t dcol = (w >> kDcolShift) & kDcolMask;
        uint32_t pxid = (w >> kPxidShift) & kPxidMask;
        lrow = 80 - pxid / 2;
        lcol = dcol * 2 + pxid % 2;
        if (!(dcol < 26 && 2 <= pxid && pxid < 162)) {
          ev.errors.push_back({det.rawId, w, 37, uint8_t(fed)});
          continue;
        }
      }
      uint32_t row, col;
      toModule(barrel, side, layer, det.rocInDet, lrow, lcol, row, col);
      uint32_t a = w & kAdcMask;
      ev.xx[i] = uint16_t(row);
      ev.yy[i] = uint16_t(col);
      ev.adc[i] = uint16_t(a);
      // pack(row, col, adc): 11 bits row, 11 bits col, adc clamped to 10 bits (:122-137)
      ev.pdigi[i] = (uint32_t(uint16_t(row)) << 0) | (uint32_t(uint16_t(col)) << 11) | (std::min(a, 1023u) << 22);
      ev.moduleInd[i] = uint16_t(det.moduleId);
      ev.rawIdArr[i] = det.rawId;
    }
  }

  // ---------------------------------------------------------------------------
  // a3: gpuCalibPixel.h:22-66 (isRun2) with SiPixelGainForHLTonGPU.h:172-198
  void calibDigis(const Conditions& c, Event& ev) {
    const auto& g = *c.gain;
    const GainDecoding* lp = reinterpret_cast<const GainDecoding*>(c.gainData);
    ev.moduleStart.assign(kMaxNumModules + 1, 0);
    ev.clusInModule.assign(kMaxNumModules, 0);
    ev.moduleId.assign(kMaxNumModules, 0);
    ev.clusModuleStart.assign(kMaxNumModules + 1, 0);
    for (uint32_t i = 0; i < ev.nWords; ++i) {
      uint16_t id = ev.moduleInd[i];
      if (id == kInvId)
        continue;
      float conv = id < 96 ? 50.f : 47.f;
      float offs = id < 96 ? -670.f : -60.f;
      int row = ev.xx[i], col = ev.yy[i];
      const auto& rc = g.rangeAndCols[id];
      uint32_t lengthOfColumnData = (rc.second - rc.first) / uint32_t(rc.nCols);
      uint32_t blocks = uint32_t(row) / g.numberOfRowsAveragedOver_;
      uint32_t offset = rc.first + uint32_t(col) * lengthOfColumnData + 2 * blocks;
      GainDecoding s = lp[offset / 2];
      bool dead = s.ped == g.deadFlag_, noisy = s.ped == g.noisyFlag_;
      float ped = float(s.ped) * g.pedPrecision + g.minPed_;
      float gain = float(s.gain) * g.gainPrecision + g.minGain_;
      if (dead || noisy) {
        ev.moduleInd[i] = kInvId;
        ev.adc[i] = 0;
      } else {
        float vcal = float(ev.adc[i]) * gain - ped * gain;
        ev.adc[i] = uint16_t(std::max(100, int(vcal * conv + offs)));
      }
    }
  }

  // ---------------------------------------------------------------------------
  // a4: gpuClustering.h:19-37; the serial atomicInc saturates at MaxNumModules
  void countModules(Event& ev) {
    const auto& id = ev.moduleInd;
    for (uint32_t i = 0; i < ev.nWords; ++i) {
      ev.clus[i] = int32_t(i);
      if (id[i] == kInvId)
        continue;
      int j = int(i) - 1;
      while (j >= 0 && id[j] == kInvId)
        --j;
      if (j < 0 || id[j] != id[i]) {
        uint32_t loc = ev.moduleStart[0];
        if (ev.moduleStart[0] < kMaxNumModules)
          ev.moduleStart[0]++;
        ev.moduleStart[loc + 1] = i;
      }
    }
    ev.nModules = ev.moduleStart[0];
    ev.nDigis = ev.nWords;
  }

  // ---------------------------------------------------------------------------
  // a5: gpuClustering.h:40-273.  Per module: column histogram (418 bins,
  // filled in decreasing insertion order), nearest-neighbour lists of the
  // pixels later in the same column and in the next column with |dx| <= 1,
  // alternating min-label propagation / pointer jumping until stable, then
  // roots numbered in increasing pixel index.
  void findClus(Event& ev) {
    const auto& id = ev.moduleInd;
    const auto& x = ev.xx;
    const auto& y = ev.yy;
    auto& clus = ev.clus;
    const int n = int(ev.nWords);
    constexpr int nbins = kColsPerModule + 2;
    std::vector<uint32_t> off(nbins + 1);
    std::vector<uint16_t> bins(kMaxPixInModule);
    std::vector<std::array<uint16_t, 10>> nn;
    std::vector<uint8_t> nnn;
    for (uint32_t module = 0; module < ev.moduleStart[0]; ++module) {
      const int first = int(ev.moduleStart[1 + module]);
      const uint16_t me = id[first];
      assert(me < kMaxNumModules);
      int msize = n;
      for (int i = first; i < n; ++i) {
        if (id[i] == kInvId)
          continue;
        if (id[i] != me) {
          msize = i;
          break;
        }
      }
      if (msize - first > int(kMaxPixInModule))
        msize = first + int(kMaxPixInModule);
      // histogram over the column (9 significant bits of a uint16: bin = y)
      std::fill(off.begin(), off.end(), 0);
      for (int i = first; i < msize; ++i)
        if (id[i] != kInvId)
          off[y[i]]++;
      for (int b = 1; b <= nbins; ++b)
        off[b] += off[b - 1];
      for (int i = first; i < msize; ++i)
        if (id[i] != kInvId)
          bins[--off[y[i]]] = uint16_t(i - first);
      const uint32_t hsize = off[nbins];
      nn.assign(hsize, {});
      nnn.assign(hsize, 0);
      for (uint32_t j = 0; j < hsize; ++j) {
        int i = bins[j] + first;
        uint32_t e = off[y[i] + 1 + 1];  // end of the next column's bin
        for (uint32_t p = j + 1; p < e; ++p) {
          int mm = bins[p] + first;
          if (std::abs(int(x[mm]) - int(x[i])) > 1)
            continue;
          assert(nnn[j] < 10);
          nn[j][nnn[j]++] = bins[p];
        }
      }
      bool more = true;
      int nloops = 0;
      while (more) {
        if (nloops % 2 == 1) {
          for (uint32_t j = 0; j < hsize; ++j) {
            int i = bins[j] + first;
            int mm = clus[i];
            while (mm != clus[mm])
              mm = clus[mm];
            clus[i] = mm;
          }
        } else {
          more = false;
          for (uint32_t j = 0; j < hsize; ++j) {
            int i = bins[j] + first;
            for (int kk = 0; kk < nnn[j]; ++kk) {
              int mm = nn[j][kk] + first;
              int old = clus[mm];
              clus[mm] = std::min(clus[mm], clus[i]);
              if (old != clus[i])
                more = true;
              clus[i] = std::min(clus[i], old);
            }
          }
        }
        ++nloops;
      }
      uint32_t found = 0;
      for (int i = first; i < msize; ++i) {
        if (id[i] == kInvId)
          continue;
        if (clus[i] == i)
          clus[i] = -int32_t(++found);
      }
      for (int i = first; i < msize; ++i) {
        if (id[i] == kInvId)
          continue;
        if (clus[i] >= 0)
          clus[i] = clus[clus[i]];
      }
      for (int i = first; i < msize; ++i) {
        if (id[i] == kInvId) {
          clus[i] = -9999;
          continue;
        }
        clus[i] = -clus[i] - 1;
      }
      ev.clusInModule[me] = found;
      ev.moduleId[module] = me;
    }
  }

  // ---------------------------------------------------------------------------
  // a6: gpuClusterChargeCut.h:14-114
  void clusterChargeCut(Event& ev) {
    auto& id = ev.moduleInd;
    auto& clus = ev.clus;
    const uint32_t n = ev.nWords;
    std::vector<int32_t> charge(kMaxClustersPerModule);
    std::vector<uint8_t> ok(kMaxClustersPerModule);
    std::vector<uint16_t> newId(kMaxClustersPerModule);
    for (uint32_t module = 0; module < ev.moduleStart[0]; ++module) {
      const uint32_t first = ev.moduleStart[1 + module];
      const uint16_t me = id[first];
      assert(me == ev.moduleId[module]);
      uint32_t nclus = ev.clusInModule[me];
      if (nclus == 0)
        continue;
      if (nclus > kMaxClustersPerModule) {
        for (uint32_t i = first; i < n; ++i) {
          if (id[i] == kInvId)
            continue;
          if (id[i] != me)
            break;
          if (clus[i] >= int32_t(kMaxClustersPerModule)) {
            id[i] = kInvId;
            clus[i] = kInvId;
          }
        }
        nclus = kMaxClustersPerModule;
      }
      std::fill(charge.begin(), charge.begin() + nclus, 0);
      for (uint32_t i = first; i < n; ++i) {
        if (id[i] == kInvId)
          continue;
        if (id[i] != me)
          break;
        charge[clus[i]] += ev.adc[i];
      }
      const int cut = me < 96 ? 2000 : 4000;
      for (uint32_t i = 0; i < nclus; ++i)
        newId[i] = ok[i] = charge[i] > cut ? 1 : 0;
      for (uint32_t i = 1; i < nclus; ++i)  // inclusive scan in place
        newId[i] = uint16_t(newId[i] + newId[i - 1]);
      if (nclus == newId[nclus - 1])
        continue;
      ev.clusInModule[me] = newId[nclus - 1];
      for (uint32_t i = 0; i < nclus; ++i)
        if (!ok[i])
          newId[i] = kInvId + 1;
      for (uint32_t i = first; i < n; ++i) {
        if (id[i] == kInvId)
          continue;
        if (id[i] != me)
          break;
        clus[i] = int32_t(newId[clus[i]]) - 1;
        if (clus[i] == kInvId)
          id[i] = kInvId;
      }
    }
  }

  // ---------------------------------------------------------------------------
  // a7: SiPixelRawToClusterGPUKernel.cc:464-506
  void fillHitsModuleStart(Event& ev) {
    auto& ms = ev.clusModuleStart;
    ms[0] = 0;
    for (uint32_t i = 0; i < kMaxNumModules; ++i)
      ms[i + 1] = std::min(kMaxClustersPerModule, ev.clusInModule[i]);
    for (uint32_t i = 2; i <= kMaxNumModules; ++i)
      ms[i] += ms[i - 1];
    for (uint32_t i = 0; i <= kMaxNumModules; ++i)
      if (ms[i] > kMaxNumClusters)
        ms[i] = kMaxNumClusters;
    ev.nClusters = ms[kMaxNumModules];
  }

}  // namespace orc

// ORACLE -- test infrastructure only (see oracle.h).
// Stage B of the serial chain: cluster position estimate (CPE), rec-hits in
// global coordinates and the per-layer phi histogram, restated from the
// reference serial program with its floating-point operation order.
#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdlib>
#include <limits>

#include "oracle.h"

namespace orc {

  namespace {
    // Geometry/phase1PixelTopology.h:115-158
    inline uint16_t div52(uint16_t n) { return uint16_t(n / 52); }  // equal to divu52 for n < 512
    inline bool bigX(uint16_t px) { return px == 79 || px == 80; }
    inline bool bigY(uint16_t py) {
      uint16_t ly = uint16_t(py - 52 * div52(py));
      return ly == 0 || ly == 51;
    }
    inline uint16_t unitX(uint16_t px) { return uint16_t(px + (px > 79 ? 1 : 0) + (px > 80 ? 1 : 0)); }
    inline uint16_t unitY(uint16_t py) {
      uint16_t roc = div52(py);
      uint16_t in = uint16_t(py - 52 * roc);
      return uint16_t(py + 2 * roc + (in > 0 ? 1 : 0));
    }

    struct Clus {
      uint32_t minRow, maxRow, minCol, maxCol;
      int32_t qfx, qlx, qfy, qly, charge;
      float xpos, ypos, xerr, yerr;
      int16_t xsize, ysize;
    };

    // CondFormats/pixelCPEforGPU.h:118-174
    float correction(int sizeM1, int qf, int ql, uint16_t upperFirst, uint16_t lowerLast, float lorentz,
                     float thickness, float cotAngle, float pitch, bool firstBig, bool lastBig) {
      if (sizeM1 == 0)
        return 0;
      float weff = 0;
      bool simple = true;
      if (sizeM1 == 1) {
        float winner = pitch * float(int(lowerLast) - int(upperFirst));
        float wpred = thickness * cotAngle - lorentz;
        weff = std::abs(wpred) - winner;
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
      float qdiff = float(ql - qf);
      float qsum = float(ql + qf);
      if (qsum == 0)
        qsum = 1.0f;
      return 0.5f * (qdiff / qsum) * weff;
    }

    // CondFormats/pixelCPEforGPU.h:176-259
    void position(const CPECommonParams& com, const CPEDetParams& det, Clus& cp) {
      uint16_t llx = uint16_t(cp.minRow + 1), lly = uint16_t(cp.minCol + 1);
      uint16_t urx = uint16_t(cp.maxRow), ury = uint16_t(cp.maxCol);
      uint16_t llxl = unitX(llx), llyl = unitY(lly), urxl = unitX(urx), uryl = unitY(ury);
      int mx = llxl + urxl, my = llyl + uryl;
      int xsize = int(urxl) + 2 - int(llxl);
      int ysize = int(uryl) + 2 - int(llyl);
      if (bigX(uint16_t(cp.minRow)))
        ++xsize;
      if (bigX(uint16_t(cp.maxRow)))
        ++xsize;
      if (bigY(uint16_t(cp.minCol)))
        ++ysize;
      if (bigY(uint16_t(cp.maxCol)))
        ++ysize;
      int unbX = int(8. * double(std::abs(float(cp.qfx - cp.qlx))) / double(float(cp.qfx + cp.qlx)));
      int unbY = int(8. * double(std::abs(float(cp.qfy - cp.qly))) / double(float(cp.qfy + cp.qly)));
      xsize = 8 * xsize - unbX;
      ysize = 8 * ysize - unbY;
      cp.xsize = int16_t(std::min(xsize, 1023));
      cp.ysize = int16_t(std::min(ysize, 1023));
      if (cp.minRow == 0 || cp.maxRow == kLastRow)
        cp.xsize = int16_t(-cp.xsize);
      if (cp.minCol == 0 || cp.maxCol == kLastCol)
        cp.ysize = int16_t(-cp.ysize);
      float xPos = det.shiftX + com.thePitchX * (0.5f * float(mx) + float(kXOffset));
      float yPos = det.shiftY + com.thePitchY * (0.5f * float(my) + float(kYOffset));
      // computeAnglesFromDet (:106-116)
      float gvz = -1.f / det.z0;
      float cotalpha = (xPos - det.x0) * gvz;
      float cotbeta = (yPos - det.y0) * gvz;
      float thickness = det.isBarrel ? com.theThicknessB : com.theThicknessE;
      float xcorr = correction(int(cp.maxRow) - int(cp.minRow), cp.qfx, cp.qlx, llxl, urxl, det.chargeWidthX, thickness,
                               cotalpha, com.thePitchX, bigX(uint16_t(cp.minRow)), bigX(uint16_t(cp.maxRow)));
      float ycorr = correction(int(cp.maxCol) - int(cp.minCol), cp.qfy, cp.qly, llyl, uryl, det.chargeWidthY, thickness,
                               cotbeta, com.thePitchY, bigY(uint16_t(cp.minCol)), bigY(uint16_t(cp.maxCol)));
      cp.xpos = xPos + xcorr;
      cp.ypos = yPos + ycorr;
    }

    // CondFormats/pixelCPEforGPU.h:316-340
    void errorFromDB(const CPEDetParams& det, Clus& cp) {
      cp.xerr = 0.0050f;
      cp.yerr = 0.0085f;
      uint32_t sx = cp.maxRow - cp.minRow, sy = cp.maxCol - cp.minCol;
      bool edgeX = cp.minRow == 0 || cp.maxRow == kLastRow;
      bool edgeY = cp.minCol == 0 || cp.maxCol == kLastCol;
      uint32_t ix = (sx == 0) + ((sx == 0) && bigX(uint16_t(cp.minRow)));
      uint32_t iy = (sy == 0) + ((sy == 0) && bigY(uint16_t(cp.minCol)));
      if (!edgeX)
        cp.xerr = det.sx[ix];
      if (!edgeY)
        cp.yerr = det.sy[iy];
    }
  }  // namespace

  // DataFormats/approx_atan2.h:230-263 (degree-7 polynomial, 16-bit angle)
  int16_t atan2s7(float y, float x) {
    float r = (std::abs(x) - std::abs(y)) / (std::abs(x) + std::abs(y));
    if (x < 0)
      r = -r;
    float z = r * r;
    float p = r * ((-10422.177734375f) + z * (3349.97412109375f + z * ((-1525.589599609375f) + z * 406.64190673828125f)));
    int angle = (x >= 0) ? 8192 : 24576;
    angle = int16_t(angle + int(int16_t(p)));
    return int16_t(y < 0 ? -angle : angle);
  }

  // a8: plugin-SiPixelRecHits/gpuPixelRecHits.h:16-207
  void getHits(const Conditions& c, Event& ev) {
    const auto& bs = *c.beamSpot;
    const auto& ag = *c.cpeAverage;
    auto& agc = ev.avgCorrected;
    for (uint32_t il = 0; il < kLaddersInBarrel; ++il) {
      agc.ladderZ[il] = ag.ladderZ[il] - bs.z;
      agc.ladderX[il] = ag.ladderX[il] - bs.x;
      agc.ladderY[il] = ag.ladderY[il] - bs.y;
      agc.ladderR[il] = std::sqrt(agc.ladderX[il] * agc.ladderX[il] + agc.ladderY[il] * agc.ladderY[il]);
      agc.ladderMinZ[il] = ag.ladderMinZ[il] - bs.z;
      agc.ladderMaxZ[il] = ag.ladderMaxZ[il] - bs.z;
    }
    agc.endCapZ[0] = ag.endCapZ[0] - bs.z;
    agc.endCapZ[1] = ag.endCapZ[1] - bs.z;

    const uint32_t H = ev.nClusters;
    ev.nHits = H;
    ev.xl.assign(H, 0);
    ev.yl.assign(H, 0);
    ev.xerr.assign(H, 0);
    ev.yerr.assign(H, 0);
    ev.xg.assign(H, 0);
    ev.yg.assign(H, 0);
    ev.zg.assign(H, 0);
    ev.rg.assign(H, 0);
    ev.iphi.assign(H, 0);
    ev.xsize.assign(H, 0);
    ev.ysize.assign(H, 0);
    ev.charge.assign(H, 0);
    ev.detInd.assign(H, 0);
    if (ev.nModules == 0)
      return;
    const int n = int(ev.nDigis);
    std::vector<Clus> cps(kMaxHitsInIter);
    for (uint32_t module = 0; module < ev.moduleStart[0]; ++module) {
      const uint32_t me = ev.moduleId[module];
      const int nclus = int(ev.clusInModule[me]);
      if (nclus == 0)
        continue;
      const CPEDetParams& det = c.cpeDet[me];
      for (int start = 0; start < nclus; start += int(kMaxHitsInIter)) {
        int first = int(ev.moduleStart[1 + module]);
        int nIn = std::min(int(kMaxHitsInIter), nclus - start);
        int last = start + nIn;
        for (int ic = 0; ic < nIn; ++ic) {
          Clus& cp = cps[ic];
          cp.minRow = cp.minCol = std::numeric_limits<uint32_t>::max();
          cp.maxRow = cp.maxCol = 0;
          cp.charge = cp.qfx = cp.qlx = cp.qfy = cp.qly = 0;
        }
        for (int i = first; i < n; ++i) {
          uint16_t id = ev.moduleInd[i];
          if (id == kInvId)
            continue;
          if (id != me)
            break;
          int cl = ev.clus[i];
          if (cl < start || cl >= last)
            continue;
          Clus& cp = cps[cl - start];
          cp.minRow = std::min<uint32_t>(cp.minRow, ev.xx[i]);
          cp.maxRow = std::max<uint32_t>(cp.maxRow, ev.xx[i]);
          cp.minCol = std::min<uint32_t>(cp.minCol, ev.yy[i]);
          cp.maxCol = std::max<uint32_t>(cp.maxCol, ev.yy[i]);
        }
        for (int i = first; i < n; ++i) {
          uint16_t id = ev.moduleInd[i];
          if (id == kInvId)
            continue;
          if (id != me)
            break;
          int cl = ev.clus[i];
          if (cl < start || cl >= last)
            continue;
          Clus& cp = cps[cl - start];
          int32_t ch = std::min<uint16_t>(ev.adc[i], 65535);
          cp.charge += ch;
          if (cp.minRow == ev.xx[i])
            cp.qfx += ch;
          if (cp.maxRow == ev.xx[i])
            cp.qlx += ch;
          if (cp.minCol == ev.yy[i])
            cp.qfy += ch;
          if (cp.maxCol == ev.yy[i])
            cp.qly += ch;
        }
        first = int(ev.clusModuleStart[me]) + start;
        for (int ic = 0; ic < nIn; ++ic) {
          int h = first + ic;
          if (h >= int(kMaxNumClusters))
            break;
          Clus& cp = cps[ic];
          position(*c.cpeCommon, det, cp);
          errorFromDB(det, cp);
          ev.charge[h] = cp.charge;
          ev.detInd[h] = uint16_t(me);
          float xl = cp.xpos, yl = cp.ypos;
          ev.xl[h] = xl;
          ev.yl[h] = yl;
          ev.xsize[h] = cp.xsize;
          ev.ysize[h] = cp.ysize;
          ev.xerr[h] = cp.xerr * cp.xerr;
          ev.yerr[h] = cp.yerr * cp.yerr;
          // SOAFrame::toGlobal(x, y) (DataFormats/SOARotation.h:89-94, 55-59)
          const auto& f = det.frame;
          float xg = f.R11 * xl + f.R21 * yl;
          float yg = f.R12 * xl + f.R22 * yl;
          float zg = f.R13 * xl + f.R23 * yl;
          xg += f.px;
          yg += f.py;
          zg += f.pz;
          xg -= bs.x;
          yg -= bs.y;
          zg -= bs.z;
          ev.xg[h] = xg;
          ev.yg[h] = yg;
          ev.zg[h] = zg;
          ev.rg[h] = std::sqrt(xg * xg + yg * yg);
          ev.iphi[h] = atan2s7(yg, xg);
        }
      }
    }
  }

  // a9: plugin-SiPixelRecHits/PixelRecHits.cc:15-28
  void setHitsLayerStart(const Conditions& c, Event& ev) {
    ev.hitsLayerStart.assign(kNumLayers + 1, 0);
    if (ev.nHits == 0)
      return;
    for (uint32_t l = 0; l <= kNumLayers; ++l)
      ev.hitsLayerStart[l] = ev.clusModuleStart[c.cpeLayer->layerStart[l]];
  }

  // a10: CUDACore/HistoContainer.h:18-73,238-242 via PixelRecHits.cc:46-48.
  // bin = (iphi >> 9) & 127 within the hit's layer; filled by decrementing the
  // running offsets, so each bin lists its hits in decreasing index order.
  void fillPhiBinner(Event& ev) {
    ev.phiOff.assign(kPhiHistTotBins, 0);
    ev.phiBins.assign(ev.nHits, 0);
    if (ev.nHits == 0)
      return;
    const auto& ls = ev.hitsLayerStart;
    const uint32_t total = ls[kNumLayers];
    auto layerOf = [&](uint32_t i) {
      uint32_t l = uint32_t(std::upper_bound(ls.begin(), ls.end(), i) - ls.begin()) - 1;
      return l;
    };
    auto bin = [&](uint32_t i) { return uint32_t((int(ev.iphi[i]) >> 9) & 127) + kPhiBins * layerOf(i); };
    for (uint32_t i = 0; i < total; ++i)
      ev.phiOff[bin(i)]++;
    for (uint32_t b = 1; b < kPhiHistTotBins; ++b)
      ev.phiOff[b] += ev.phiOff[b - 1];
    for (uint32_t i = 0; i < total; ++i)
      ev.phiBins[--ev.phiOff[bin(i)]] = uint16_t(i);
  }

  bool process(const Conditions& c, const uint8_t* raw, size_t len, int upto, Event& ev) {
    ev.caps = c.caps;
    ev.vertexAlgo = c.vertexAlgo;
    ev.vertexPtMin = c.vertexPtMin;
    ev.earlyFishbone = c.earlyFishbone;
    ev.lateFishbone = c.lateFishbone;
    if (!unpackFeds(c, raw, len, ev))
      return false;
    rawToDigi(c, ev);
    calibDigis(c, ev);
    if (upto < 2)
      return true;
    countModules(ev);
    findClus(ev);
    clusterChargeCut(ev);
    fillHitsModuleStart(ev);
    if (upto < 3)
      return true;
    getHits(c, ev);
    setHitsLayerStart(c, ev);
    fillPhiBinner(ev);
    if (upto < 4)
      return true;
    // CAHitNtupletGeneratorOnGPU::makeTuples (CAHitNtupletGeneratorOnGPU.cc:96-126)
    buildDoublets(ev);
    connectCells(ev);
    if (ev.nHits > 1 && ev.earlyFishbone)
      fishbone(ev);
    findNtuplets(ev);
    earlyDuplicatesAndMultiplicity(ev);
    // CAHitNtupletGeneratorKernels.cc:133-136: the fishbone again over the cells in tuples
    if (ev.nHits > 1 && ev.lateFishbone)
      fishbone(ev, true);
    fillHitDetIndices(ev);
    if (c.riemannFit)  // CAHitNtupletGeneratorOnGPU.cc:168-172
      fitTracksRiemann(c, ev);
    else
      fitTracks(c, ev);
    if (ev.nHits > 0) {
      classifyTracks(ev);
      if (ev.lateFishbone)  // CAHitNtupletGeneratorKernels.cc:161-164
        fishboneCleaner(ev);
      fastDuplicateRemover(ev);
      hitToTupleAndTripletCleaner(ev);
      countStats(ev);
    }
    if (upto < 5)
      return true;
    findVertices(ev);
    return true;
  }

}  // namespace orc

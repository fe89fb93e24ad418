// ORACLE -- test infrastructure only (see oracle.h).
// Stage C of the serial chain, cellular-automaton part: doublets, connection,
// fishbone, ntuplet search, multiplicity binning and the quality cleaners,
// restated from plugin-PixelTriplets of the reference serial program with its
// iteration order and its container capacities.
#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdlib>
#include <limits>

#include "oracle.h"

namespace orc {

  namespace {
    // gpuPixelDoublets.h:14-54
    constexpr int kPairs = 19;
    constexpr uint8_t kLayerPairs[2 * kPairs] = {0, 1, 0, 4, 0, 7, 1, 2, 1, 4, 1, 7, 4, 5, 7, 8, 2, 3, 2, 4,
                                                 2, 7, 5, 6, 8, 9, 0, 2, 1, 3, 0, 5, 0, 8, 4, 6, 7, 9};
    constexpr int16_t kPhi05 = 522, kPhi06 = 626, kPhi07 = 730;
    constexpr int16_t kPhiCuts[kPairs] = {kPhi05, kPhi07, kPhi07, kPhi05, kPhi06, kPhi06, kPhi05,
                                          kPhi05, kPhi06, kPhi06, kPhi06, kPhi05, kPhi05, kPhi05,
                                          kPhi05, kPhi05, kPhi05, kPhi05, kPhi05};
    constexpr float kMinZ[kPairs] = {-20., 0., -30., -22., 10., -30., -70., -70., -22., 15.,
                                     -30,  -70., -70., -20., -22., 0,   -30., -70., -70.};
    constexpr float kMaxZ[kPairs] = {20., 30., 0., 22., 30., -10., 70., 70., 22., 30.,
                                     -15., 70., 70., 20., 22., 30., 0.,  70., 70.};
    constexpr float kMaxR[kPairs] = {20., 9., 9., 20., 7., 7., 5., 5., 20., 6., 6., 5., 5., 20., 20., 9., 9., 9., 9.};

    // CAHitNtupletGeneratorOnGPU.cc:45-65 (the serial program's parameters)
    constexpr float kPtMin = 0.899999976158f;
    constexpr float kThetaCutBarrel = 0.00200000009499f;
    constexpr float kThetaCutForward = 0.00300000002608f;
    constexpr float kHardCurvCut = 0.0328407224959f;
    constexpr float kDcaCutInner = 0.15000000596f;
    constexpr float kDcaCutOuter = 0.25f;

    inline int16_t phiBin(int16_t t) { return int16_t((int(t) >> 9) & 127); }

    // CircleEq.h:33-37,69-105 (float)
    struct Circle {
      float xp, yp, c, alpha, beta;
    };
    Circle circleEq(float x1, float y1, float x2, float y2, float x3, float y3) {
      bool noflip = std::abs(x3 - x1) < std::abs(y3 - y1);
      float x1p = noflip ? x1 - x2 : y1 - y2;
      float y1p = noflip ? y1 - y2 : x1 - x2;
      float d12 = x1p * x1p + y1p * y1p;
      float x3p = noflip ? x3 - x2 : y3 - y2;
      float y3p = noflip ? y3 - y2 : x3 - x2;
      float d32 = x3p * x3p + y3p * y3p;
      float num = x1p * y3p - y1p * x3p;
      float det = d12 * y3p - d32 * y1p;
      float st2 = (d12 * x3p - d32 * x1p);
      float seq = det * det + st2 * st2;
      float al2 = 1.f / std::sqrt(seq);
      float be2 = -st2 * al2;
      float ct = 2.f * num * al2;
      al2 *= det;
      Circle e;
      e.xp = x2;
      e.yp = y2;
      e.c = noflip ? ct : -ct;
      e.alpha = noflip ? al2 : -be2;
      e.beta = noflip ? be2 : -al2;
      return e;
    }
    float dca0(const Circle& e) {
      float x = e.c * e.xp + e.alpha;
      float y = e.c * e.yp + e.beta;
      return std::sqrt(x * x + y * y) - 1.f;
    }

    // GPUCACell.h:175-185
    bool alignedRZ(float r1, float z1, float ri, float zi, float ro, float zo, float ptmin, float thetaCut) {
      float rd = std::abs(r1 - ro);
      float d13 = rd * rd + (z1 - zo) * (z1 - zo);
      float pMin = ptmin * std::sqrt(d13);
      float tan1213 = std::fabs(z1 * (ri - ro) + zi * (ro - r1) + zo * (r1 - ri));
      return tan1213 * pMin <= thetaCut * d13 * rd;
    }
  }  // namespace

  void circleForTest(float x1, float y1, float x2, float y2, float x3, float y3, float* curvature, float* dca) {
    Circle e = circleEq(x1, y1, x2, y2, x3, y3);
    *curvature = e.c;
    *dca = dca0(e);
  }

  // a11 initDoublets + a12 getDoubletsFromHisto: gpuPixelDoublets.h:64-124,
  // gpuPixelDoubletsAlgos.h:25-238 (all cuts on, ideal conditions, 19 pairs)
  void buildDoublets(Event& ev) {
    auto& C = ev.ca;
    const uint32_t H = ev.nHits;
    C = CAState{};
    C.outerHitOfCell.assign(std::max(1u, H), {});
    C.cells.clear();
    C.neighborStore.assign(1, {});  // entry 0 is the shared empty list
    C.trackStore.assign(1, {});
    if (H == 0)
      return;
    const uint32_t* offsets = ev.hitsLayerStart.data();
    auto layerSize = [&](uint8_t l) { return offsets[l + 1] - offsets[l]; };
    uint32_t cum[kPairs];
    cum[0] = layerSize(kLayerPairs[0]);
    for (int p = 1; p < kPairs; ++p)
      cum[p] = cum[p - 1] + layerSize(kLayerPairs[2 * p]);
    const uint32_t ntot = cum[kPairs - 1];

    constexpr int minYsizeB1 = 36, minYsizeB2 = 28, maxDYsize12 = 28, maxDYsize = 20, maxDYPred = 20;
    constexpr float dzdrFact = 8 * 0.0285 / 0.015;
    constexpr float z0cut = 12.f;
    constexpr float hardPtCut = 0.5f;
    constexpr float minRadius = hardPtCut * 87.78f;
    constexpr float minRadius2T4 = 4.f * minRadius * minRadius;
    constexpr float i2p = M_PI / ((int)(std::numeric_limits<short>::max()) + 1);

    uint32_t pairId = 0;
    for (uint32_t j = 0; j < ntot; ++j) {
      while (j >= cum[pairId++])
        ;
      --pairId;
      const uint8_t inner = kLayerPairs[2 * pairId], outer = kLayerPairs[2 * pairId + 1];
      const uint32_t hoff = kPhiBins * outer;
      uint32_t i = (pairId == 0) ? j : j - cum[pairId - 1];
      i += offsets[inner];
      const uint16_t mi = ev.detInd[i];
      if (mi > 2000)
        continue;
      const float mez = ev.zg[i];
      if (mez < kMinZ[pairId] || mez > kMaxZ[pairId])
        continue;
      const int16_t mes = ev.ysize[i];  // ideal conditions: every ladder is an outer ladder
      if (inner == 0 && outer > 3 && mes > 0 && mes < minYsizeB1)
        continue;
      if (inner == 1 && outer > 3 && mes > 0 && mes < minYsizeB2)
        continue;
      const int16_t mep = ev.iphi[i];
      const float mer = ev.rg[i];
      const int16_t iphicut = kPhiCuts[pairId];
      const uint32_t maxr = 0;
      (void)maxr;
      auto z0cutoff = [&](uint32_t o) {
        float zo = ev.zg[o], ro = ev.rg[o];
        float dr = ro - mer;
        return dr > kMaxR[pairId] || dr < 0 || std::abs((mez * ro - mer * zo)) > z0cut * dr;
      };
      auto zsizeCut = [&](uint32_t o) {
        bool onlyBarrel = outer < 4;
        int16_t so = ev.ysize[o];
        int dy = inner == 0 ? maxDYsize12 : maxDYsize;
        float zo = ev.zg[o], ro = ev.rg[o];
        return onlyBarrel ? (mes > 0 && so > 0 && std::abs(so - mes) > dy)
                          : (inner < 4) && mes > 0 &&
                                std::abs(mes - int(std::abs((mez - zo) / (mer - ro)) * dzdrFact + 0.5f)) > maxDYPred;
      };
      auto ptcut = [&](uint32_t o, int16_t idphi) {
        float ri = mer, ro = ev.rg[o];
        float dphi = float(idphi) * i2p;
        return dphi * dphi * (minRadius2T4 - ri * ro) > (ro - ri) * (ro - ri);
      };
      const int16_t kl = phiBin(int16_t(mep - iphicut));
      const int16_t kh = phiBin(int16_t(mep + iphicut));
      const int16_t khh = int16_t((kh + 1) % kPhiBins);
      for (int16_t kk = kl; kk != khh; kk = int16_t((kk + 1) % kPhiBins)) {
        const uint32_t b = kk + hoff;
        for (uint32_t p = ev.phiOff[b]; p < ev.phiOff[b + 1]; ++p) {
          const uint32_t oi = ev.phiBins[p];
          if (ev.detInd[oi] > 2000)
            continue;
          if (z0cutoff(oi))
            continue;
          const int16_t mop = ev.iphi[oi];
          const uint16_t idphi = uint16_t(std::min(std::abs(int16_t(mop - mep)), std::abs(int16_t(mep - mop))));
          if (idphi > iphicut)
            continue;
          if (zsizeCut(oi))
            continue;
          if (ptcut(oi, int16_t(idphi)))
            continue;
          const uint32_t ind = uint32_t(C.cells.size());
          if (ind >= ev.caps.maxDoublets) {
            C.status |= kCaCellOverflow;
            break;
          }
          Cell c;
          c.innerHit = uint16_t(i);
          c.outerHit = uint16_t(oi);
          c.doubletId = int32_t(ind);
          c.layerPairId = int16_t(pairId);
          c.used = 0;
          c.innerZ = ev.zg[i];
          c.innerR = ev.rg[i];
          c.neighbors = 0;
          c.tracks = 0;
          C.cells.push_back(c);
          if (C.outerHitOfCell[oi].push(ind, ev.caps.maxCellsPerHit) < 0)
            C.status |= kCaOuterHitOverflow;
        }
      }
    }
    C.nCells = uint32_t(C.cells.size());
  }

  namespace {
    template <class Store>
    int addTo(Store& store, uint32_t capacity, int32_t& slot, uint32_t value, uint32_t& status, uint32_t flagFull,
              uint32_t flagList) {
      if (store[slot].size == 0) {
        if (store.size() >= capacity) {
          status |= flagFull;
          return -1;
        }
        slot = int32_t(store.size());
        store.emplace_back();
      }
      int r = store[slot].push(value);
      if (r < 0)
        status |= flagList;
      return r;
    }
  }  // namespace

  // a13 kernel_connect: CAHitNtupletGeneratorKernelsImpl.h:191-262
  void connectCells(Event& ev) {
    auto& C = ev.ca;
    for (uint32_t idx = 0; idx < C.nCells; ++idx) {
      Cell& me = C.cells[idx];
      const auto& list = C.outerHitOfCell[me.innerHit];
      const float ri = me.innerR, zi = me.innerZ;
      const float ro = ev.rg[me.outerHit], zo = ev.zg[me.outerHit];
      const bool isBarrel = float(ev.detInd[me.innerHit]) < 1184;
      for (int j = 0; j < list.size; ++j) {
        Cell& oc = C.cells[list.data[j]];
        const float r1 = oc.innerR, z1 = oc.innerZ;
        bool aligned = alignedRZ(r1, z1, ri, zi, ro, zo, kPtMin, isBarrel ? kThetaCutBarrel : kThetaCutForward);
        if (!aligned)
          continue;
        // dcaCut (GPUCACell.h:187-206)
        const float cut = float(ev.detInd[oc.innerHit]) < 96 ? kDcaCutInner : kDcaCutOuter;
        Circle eq = circleEq(ev.xg[oc.innerHit], ev.yg[oc.innerHit], ev.xg[me.innerHit], ev.yg[me.innerHit],
                             ev.xg[me.outerHit], ev.yg[me.outerHit]);
        if (eq.c > kHardCurvCut)
          continue;
        if (!(std::abs(dca0(eq)) < cut * std::abs(eq.c)))
          continue;
        addTo(C.neighborStore, ev.caps.maxCellLists, oc.neighbors, idx, C.status, kCaNeighborStoreOverflow, kCaNeighborOverflow);
        me.used |= 1;
        oc.used |= 1;
      }
    }
  }

  // a14 fishbone: gpuFishbone.h:21-90 (checkTrack: the late pass, over the cells in tuples)
  void fishbone(Event& ev, bool checkTrack) {
    auto& C = ev.ca;
    constexpr int kMax = kCaMaxCellsPerHitCap;
    float x[kMax], y[kMax], z[kMax], n[kMax];
    uint16_t d[kMax];
    uint32_t cc[kMax];
    for (uint32_t h = 0; h < ev.nHits; ++h) {
      const auto& vc = C.outerHitOfCell[h];
      if (vc.size < 2)
        continue;
      const Cell& c0 = C.cells[vc.data[0]];
      const float xo = ev.xg[c0.outerHit], yo = ev.yg[c0.outerHit], zo = ev.zg[c0.outerHit];
      int sg = 0;
      for (int ic = 0; ic < vc.size; ++ic) {
        const Cell& ci = C.cells[vc.data[ic]];
        if (ci.used == 0)
          continue;
        if (checkTrack && (ci.tracks == 0 || C.trackStore[ci.tracks].size == 0))
          continue;
        cc[sg] = vc.data[ic];
        d[sg] = uint16_t(float(ev.detInd[ci.innerHit]));
        x[sg] = ev.xg[ci.innerHit] - xo;
        y[sg] = ev.yg[ci.innerHit] - yo;
        z[sg] = ci.innerZ - zo;
        n[sg] = x[sg] * x[sg] + y[sg] * y[sg] + z[sg] * z[sg];
        ++sg;
      }
      if (sg < 2)
        continue;
      for (int ic = 0; ic < sg - 1; ++ic) {
        Cell& ci = C.cells[cc[ic]];
        for (int jc = ic + 1; jc < sg; ++jc) {
          Cell& cj = C.cells[cc[jc]];
          float cos12 = x[ic] * x[jc] + y[ic] * y[jc] + z[ic] * z[jc];
          if (d[ic] != d[jc] && cos12 * cos12 >= 0.99999f * n[ic] * n[jc]) {
            if (n[ic] > n[jc]) {
              ci.doubletId = -1;
              break;
            } else {
              cj.doubletId = -1;
            }
          }
        }
      }
    }
  }

  namespace {
    struct Walker {
      Event& ev;
      uint32_t stack[6];
      int depth = 0;

      // GPUCACell.h:278-331 (minHitsPerNtuplet = 3)
      void visit(uint32_t cellIdx) {
        auto& C = ev.ca;
        if (depth >= 4) {  // the reference asserts at most 4 cells per tuple
          C.status |= kCaDepthOverflow;
          return;
        }
        const Cell& cell = C.cells[cellIdx];
        stack[depth++] = uint32_t(cell.doubletId);
        bool last = true;
        const auto& nb = C.neighborStore[cell.neighbors];
        for (int j = 0; j < nb.size; ++j) {
          const uint32_t other = nb.data[j];
          if (C.cells[other].doubletId < 0)
            continue;
          last = false;
          visit(other);
        }
        if (last && depth >= 2) {
          uint16_t hits[6];
          int nh = 0;
          for (int k = 0; k < depth; ++k)
            hits[nh++] = C.cells[stack[k]].innerHit;
          hits[nh++] = cell.outerHit;
          // OneToManyAssoc::bulkFill with the atomic pair counter
          const uint32_t m = C.apcM, first = C.apcN;
          C.apcM += 1;
          C.apcN += uint32_t(nh);
          if (m >= ev.caps.maxTuples) {
            C.status |= kCaTupleOverflow;
          } else {
            C.tupleOff[m] = first;
            for (int k = 0; k < nh; ++k)
              C.tupleBins[first + k] = hits[k];
            for (int k = 0; k < depth; ++k)
              addTo(C.trackStore, ev.caps.maxCellLists, C.cells[stack[k]].tracks, m, C.status, kCaTrackStoreOverflow, kCaTrackOverflow);
            C.quality[m] = kQualityBad;
          }
        }
        --depth;
      }
    };
  }  // namespace

  // a15 kernel_find_ntuplets + a16 finalizeBulk
  void findNtuplets(Event& ev) {
    auto& C = ev.ca;
    C.tupleOff.assign(ev.caps.maxTuples + 1, 0);
    C.tupleBins.assign(5 * ev.caps.maxTuples, 0);
    C.quality.assign(ev.caps.maxTuples, 0);
    C.apcM = C.apcN = 0;
    Walker w{ev};
    for (uint32_t idx = 0; idx < C.nCells; ++idx) {
      const Cell& c = C.cells[idx];
      if (c.doubletId < 0)
        continue;
      const int pid = c.layerPairId;
      if (pid < 8 || pid > 12) {
        w.depth = 0;
        w.visit(idx);
      }
    }
    // bulkFinalizeFill (HistoContainer.h:193-204)
    if (C.apcM >= ev.caps.maxTuples) {
      C.tupleOff[ev.caps.maxTuples] = C.tupleOff[ev.caps.maxTuples - 1];
    } else {
      for (uint32_t i = C.apcM; i <= ev.caps.maxTuples; ++i)
        C.tupleOff[i] = C.apcN;
    }
    C.nTuples = std::min(C.apcM, ev.caps.maxTuples);
  }

  // a16 kernel_earlyDuplicateRemover, kernel_countMultiplicity/fillMultiplicity
  void earlyDuplicatesAndMultiplicity(Event& ev) {
    auto& C = ev.ca;
    auto size = [&](uint32_t t) { return C.tupleOff[t + 1] - C.tupleOff[t]; };
    for (uint32_t idx = 0; idx < C.nCells; ++idx) {
      const auto& tr = C.trackStore[C.cells[idx].tracks];
      if (tr.size < 2)
        continue;
      uint32_t maxNh = 0;
      for (int k = 0; k < tr.size; ++k)
        maxNh = std::max(size(tr.data[k]), maxNh);
      for (int k = 0; k < tr.size; ++k)
        if (size(tr.data[k]) != maxNh)
          C.quality[tr.data[k]] = kQualityDup;
    }
    C.multOff.assign(9, 0);
    C.multBins.assign(ev.caps.maxTuplesMult, 0);
    for (uint32_t t = 0; t < ev.caps.maxTuples; ++t) {
      uint32_t nh = size(t);
      if (nh < 3 || C.quality[t] == kQualityDup)
        continue;
      C.multOff[nh]++;
    }
    for (int b = 1; b < 9; ++b)
      C.multOff[b] += C.multOff[b - 1];
    if (C.multOff[8] > ev.caps.maxTuplesMult)
      C.status |= kCaMultiplicityOverflow;
    for (uint32_t t = 0; t < ev.caps.maxTuples; ++t) {
      uint32_t nh = size(t);
      if (nh < 3 || C.quality[t] == kQualityDup)
        continue;
      uint32_t w = --C.multOff[nh];
      if (w < ev.caps.maxTuplesMult)
        C.multBins[w] = t;
    }
    // The reference writes past the container's capacity when it overflows;
    // here the entries beyond it are dropped and the offsets clamped, so that
    // every reader sees only stored entries.
    for (auto& o : C.multOff)
      o = std::min(o, ev.caps.maxTuplesMult);
  }

  // a17 kernel_fillHitDetIndices
  void fillHitDetIndices(Event& ev) {
    auto& C = ev.ca;
    C.detIndices.assign(5 * ev.caps.maxTuples, 0);
    const uint32_t total = C.tupleOff[ev.caps.maxTuples];
    for (uint32_t i = 0; i < total && i < 5 * ev.caps.maxTuples; ++i)
      C.detIndices[i] = ev.detInd[C.tupleBins[i]];
  }

  // a20 kernel_classifyTracks: CAHitNtupletGeneratorKernelsImpl.h:340-406
  void classifyTracks(Event& ev) {
    auto& C = ev.ca;
    const float coeff[4] = {(float)0.68177776, (float)0.74609577, (float)-0.08035491, (float)0.00315399};
    for (uint32_t t = 0; t < ev.caps.maxTuples; ++t) {
      uint32_t nh = C.tupleOff[t + 1] - C.tupleOff[t];
      if (nh == 0)
        break;
      if (C.quality[t] == kQualityDup)
        continue;
      if (nh < 3)
        continue;
      bool isNaN = false;
      for (int i = 0; i < 5; ++i)
        isNaN |= std::isnan(C.state[t][i]);
      if (isNaN)
        continue;
      float pt = std::min<float>(C.pt[t], 10.f);
      float chi2Cut = 30.f * (coeff[0] + pt * (coeff[1] + pt * (coeff[2] + pt * coeff[3])));
      if (3.f * C.chi2[t] >= chi2Cut)
        continue;
      const bool quad = nh > 3;
      const float maxTip = quad ? 0.5f : 0.3f, minPt = quad ? 0.3f : 0.5f, maxZip = 12.f;
      bool ok = (std::abs(C.state[t][1]) < maxTip) && (C.pt[t] > minPt) && (std::abs(C.state[t][4]) < maxZip);
      if (ok)
        C.quality[t] = kQualityLoose;
    }
  }

  // a21 kernel_fastDuplicateRemover: CAHitNtupletGeneratorKernelsImpl.h:151-189
  void fastDuplicateRemover(Event& ev) {
    auto& C = ev.ca;
    for (uint32_t idx = 0; idx < C.nCells; ++idx) {
      const auto& tr = C.trackStore[C.cells[idx].tracks];
      if (tr.size < 2)
        continue;
      float mc = 10000.f;
      TupleIndex im = kNoTuple;
      for (int k = 0; k < tr.size; ++k) {
        const TupleIndex it = tr.data[k];
        float score = std::abs(C.state[it][1]);
        if (C.quality[it] == kQualityLoose && score < mc) {
          mc = score;
          im = it;
        }
      }
      for (int k = 0; k < tr.size; ++k) {
        const TupleIndex it = tr.data[k];
        if (C.quality[it] != kQualityBad && it != im)
          C.quality[it] = kQualityDup;
      }
    }
  }

  // a22 kernel_countHitInTracks / fillHitInTracks / tripletCleaner: :421-447, :479-534
  void hitToTupleAndTripletCleaner(Event& ev) {
    auto& C = ev.ca;
    auto size = [&](uint32_t t) { return C.tupleOff[t + 1] - C.tupleOff[t]; };
    C.hitToTupleOff.assign(kMaxNumClusters + 1, 0);
    C.hitToTupleBins.assign(ev.caps.maxHitToTuple, 0);
    for (uint32_t t = 0; t < ev.caps.maxTuples; ++t) {
      if (size(t) == 0)
        break;
      if (C.quality[t] != kQualityLoose)
        continue;
      for (uint32_t k = C.tupleOff[t]; k < C.tupleOff[t + 1]; ++k)
        C.hitToTupleOff[C.tupleBins[k]]++;
    }
    for (uint32_t b = 1; b <= kMaxNumClusters; ++b)
      C.hitToTupleOff[b] += C.hitToTupleOff[b - 1];
    if (C.hitToTupleOff[kMaxNumClusters] > ev.caps.maxHitToTuple)
      C.status |= kCaHitToTupleOverflow;
    for (uint32_t t = 0; t < ev.caps.maxTuples; ++t) {
      if (size(t) == 0)
        break;
      if (C.quality[t] != kQualityLoose)
        continue;
      for (uint32_t k = C.tupleOff[t]; k < C.tupleOff[t + 1]; ++k) {
        uint32_t w = --C.hitToTupleOff[C.tupleBins[k]];
        if (w < ev.caps.maxHitToTuple)
          C.hitToTupleBins[w] = t;
      }
    }
    for (auto& o : C.hitToTupleOff)
      o = std::min(o, ev.caps.maxHitToTuple);
    for (uint32_t h = 0; h < kMaxNumClusters; ++h) {
      const uint32_t b = C.hitToTupleOff[h], e = C.hitToTupleOff[h + 1];
      if (e - b < 2)
        continue;
      uint32_t maxNh = 0;
      for (uint32_t k = b; k < e; ++k)
        maxNh = std::max(size(C.hitToTupleBins[k]), maxNh);
      for (uint32_t k = b; k < e; ++k)
        if (size(C.hitToTupleBins[k]) != maxNh)
          C.quality[C.hitToTupleBins[k]] = kQualityDup;
      if (maxNh > 3)
        continue;
      float mc = 10000.f;
      TupleIndex im = kNoTuple;
      for (uint32_t k = b; k < e; ++k) {
        const TupleIndex it = C.hitToTupleBins[k];
        if (C.quality[it] != kQualityBad && std::abs(C.state[it][1]) < mc) {
          mc = std::abs(C.state[it][1]);
          im = it;
        }
      }
      for (uint32_t k = b; k < e; ++k) {
        const TupleIndex it = C.hitToTupleBins[k];
        if (C.quality[it] != kQualityBad && it != im)
          C.quality[it] = kQualityDup;
      }
    }
  }

  // kernel_fishboneCleaner (CAHitNtupletGeneratorKernelsImpl.h:113-125): the tuples of
  // the cells the fishbone killed are bad
  void fishboneCleaner(Event& ev) {
    auto& C = ev.ca;
    for (uint32_t idx = 0; idx < C.nCells; ++idx) {
      const Cell& cell = C.cells[idx];
      if (cell.doubletId >= 0 || cell.tracks == 0)
        continue;
      const auto& tr = C.trackStore[cell.tracks];
      for (int k = 0; k < tr.size; ++k)
        C.quality[tr.data[k]] = kQualityBad;
    }
  }

  // doStats counters of one event: kernel_checkOverflows (CAHitNtupletGeneratorKernelsImpl.h:30-104,
  // called after the late fishbone, while the hit -> tuple map is still empty, so that every
  // cell counts as a zero-track cell), kernel_doStatsForTracks (:410-421) and
  // kernel_doStatsForHitInTracks (:468-478) after the cleaners
  void countStats(Event& ev) {
    auto& C = ev.ca;
    auto& s = C.stats;
    s.fill(0);
    s[0] = 1;
    s[1] = ev.nHits;
    s[2] = C.nCells;
    s[3] = C.apcM;
    s[4] = C.multOff.empty() ? 0 : C.multOff[8];
    for (uint32_t t = 0; t < C.nTuples; ++t) {
      if (C.tupleOff[t + 1] == C.tupleOff[t])
        break;
      if (C.quality[t] == kQualityLoose)
        ++s[5];
    }
    if (!C.hitToTupleOff.empty())
      for (size_t h = 0; h + 1 < C.hitToTupleOff.size(); ++h) {
        const uint32_t n = C.hitToTupleOff[h + 1] - C.hitToTupleOff[h];
        s[6] += n > 0;
        s[7] += n > 1;
      }
    for (uint32_t c = 0; c < C.nCells; ++c) {
      s[8] += C.cells[c].doubletId < 0;
      s[9] += C.cells[c].used == 0;
    }
    s[10] = C.nCells;
  }

}  // namespace orc

// ORACLE -- test infrastructure only (see oracle.h).
// Stage D of the serial chain: primary-vertex finding along the beam line,
// restated from plugin-PixelVertexFinding (density clustering, weighted z
// fit with outlier rejection, 2-means splitting, refit, sort by sum pt^2)
// with the serial program's iteration and summation order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>

#include "oracle.h"

namespace orc {

  // PixelVertexProducerCUDA.cc:26-40 and gpuVertexFinderImpl.h:84-111
  void findVertices(Event& ev) {
    auto& V = ev.vtx;
    const auto& C = ev.ca;
    V = VertexState{};
    V.idv.assign(ev.caps.maxTuples, 0);
    V.ndof.assign(ev.caps.maxTuples, 0);
    const float ptMin = ev.vertexPtMin;

    // loadTracks (gpuVertexFinderImpl.h:10-43)
    std::vector<TupleIndex> itrk;
    std::vector<float> zt, ezt2, ptt2;
    for (uint32_t idx = 0; idx < ev.caps.maxTuples; ++idx) {
      const uint32_t nHits = C.tupleOff[idx + 1] - C.tupleOff[idx];
      if (nHits == 0)
        break;
      V.idv[idx] = -1;
      if (nHits < 4)
        continue;
      if (C.quality[idx] != kQualityLoose)
        continue;
      const float pt = C.pt[idx];
      if (pt < ptMin)
        continue;
      itrk.push_back(idx);
      zt.push_back(C.state[idx][4]);
      ezt2.push_back(C.cov[idx][14]);
      ptt2.push_back(pt * pt);
    }
    clusterAndFit(V, itrk, zt, ezt2, ptt2, ev.vertexAlgo);
  }

  // clusterTracksByDensity (gpuClusterTracksByDensity.h:20-207), clusterTracksDBSCAN
  // (gpuClusterTracksDBSCAN.h:17-215) or clusterTracksIterative (gpuClusterTracksIterative.h:17-193)
  // on nt tracks: iv[i] becomes the track's cluster (0 ... found-1, numbered by seed index) or
  // 9997 (noise), nn[i] its neighbour count; returns the number of clusters found
  uint32_t clusterTracks(int algo, uint32_t nt, const float* zt, const float* ezt2, int32_t* iv, int32_t* nn) {
    constexpr int minT = 2;
    constexpr float eps = 0.07f, errmax = 0.01f, chi2max = 9.f;
    std::vector<uint8_t> izt(nt);

    // clusterTracksByDensity (gpuClusterTracksByDensity.h:20-207)
    const float er2mx = errmax * errmax;
    std::vector<uint32_t> off(257, 0);
    std::vector<uint16_t> bins(nt);
    for (uint32_t i = 0; i < nt; ++i) {
      int iz = int(zt[i] * 10.);
      iz = std::min(std::max(iz, INT8_MIN), INT8_MAX);
      izt[i] = uint8_t(iz - INT8_MIN);
      off[izt[i]]++;
      iv[i] = int32_t(i);
      nn[i] = 0;
    }
    for (int b = 1; b < 257; ++b)
      off[b] += off[b - 1];
    for (uint32_t i = 0; i < nt; ++i)
      bins[--off[izt[i]]] = uint16_t(i);
    auto forEachNear = [&](uint8_t value, auto&& f) {
      int bs = value;
      int be = std::min(255, bs + 1);
      bs = std::max(0, bs - 1);
      for (uint32_t p = off[bs]; p < off[be + 1]; ++p)
        f(uint32_t(bins[p]));
    };
    if (algo == 1) {
      // clusterTracksDBSCAN (gpuClusterTracksDBSCAN.h:17-215): neighbours within
      // eps only; a core track links to its lowest-z core neighbour (first in
      // histogram order on a tie); the chains are followed to their roots; the
      // other tracks take the root of the nearest core (last on a tie) within
      // eps and the error-scaled cut
      for (uint32_t i = 0; i < nt; ++i) {
        if (ezt2[i] > er2mx)
          continue;
        forEachNear(izt[i], [&](uint32_t j) {
          if (i == j)
            return;
          float dist = std::abs(zt[i] - zt[j]);
          if (dist > eps)
            return;
          nn[i]++;
        });
      }
      for (uint32_t i = 0; i < nt; ++i) {
        if (nn[i] < minT)
          continue;
        float mz = zt[i];
        forEachNear(izt[i], [&](uint32_t j) {
          if (zt[j] >= mz)
            return;
          if (nn[j] < minT)
            return;
          float dist = std::abs(zt[i] - zt[j]);
          if (dist > eps)
            return;
          mz = zt[j];
          iv[i] = int32_t(j);
        });
      }
      for (uint32_t i = 0; i < nt; ++i) {
        int32_t m = iv[i];
        while (m != iv[m])
          m = iv[m];
        iv[i] = m;
      }
    } else if (algo == 2) {
      // clusterTracksIterative (gpuClusterTracksIterative.h:17-193): neighbours
      // within eps and the error-scaled cut; the core tracks merge labels with
      // their core neighbours later in histogram order (within the next bin)
      // and the labels are followed to their roots, in alternate passes, until
      // a merging pass changes nothing
      for (uint32_t i = 0; i < nt; ++i) {
        if (ezt2[i] > er2mx)
          continue;
        forEachNear(izt[i], [&](uint32_t j) {
          if (i == j)
            return;
          float dist = std::abs(zt[i] - zt[j]);
          if (dist > eps)
            return;
          if (dist * dist > chi2max * (ezt2[i] + ezt2[j]))
            return;
          nn[i]++;
        });
      }
      int nloops = 0;
      bool more = true;
      while (more) {
        if (1 == nloops % 2) {
          for (uint32_t i = 0; i < nt; ++i) {
            int32_t m = iv[i];
            while (m != iv[m])
              m = iv[m];
            iv[i] = m;
          }
        } else {
          more = false;
          for (uint32_t k = 0; k < nt; ++k) {
            const uint32_t i = bins[k];
            const int be = std::min(int(izt[i]) + 1, 255);
            if (nn[i] < minT)
              continue;
            for (uint32_t p = k + 1; p < off[be + 1]; ++p) {
              const uint32_t j = bins[p];
              if (nn[j] < minT)
                continue;
              float dist = std::abs(zt[i] - zt[j]);
              if (dist > eps)
                continue;
              if (dist * dist > chi2max * (ezt2[i] + ezt2[j]))
                continue;
              const int32_t old = iv[j];
              iv[j] = std::min(old, iv[i]);
              if (old != iv[i])
                more = true;
              iv[i] = std::min(iv[i], old);
            }
          }
        }
        ++nloops;
      }
    } else {
      // clusterTracksByDensity (gpuClusterTracksByDensity.h:20-207)
      for (uint32_t i = 0; i < nt; ++i) {
        if (ezt2[i] > er2mx)
          continue;
        forEachNear(izt[i], [&](uint32_t j) {
          if (i == j)
            return;
          float dist = std::abs(zt[i] - zt[j]);
          if (dist > eps)
            return;
          if (g_mutation != kMutDensityNoChi2 && dist * dist > chi2max * (ezt2[i] + ezt2[j]))
            return;
          nn[i]++;
        });
      }
      for (uint32_t i = 0; i < nt; ++i) {
        float mdist = eps;
        forEachNear(izt[i], [&](uint32_t j) {
          if (nn[j] < nn[i])
            return;
          if (nn[j] == nn[i] && (g_mutation == kMutDensityTie ? zt[j] <= zt[i] : zt[j] >= zt[i]))
            return;
          float dist = std::abs(zt[i] - zt[j]);
          if (dist > mdist)
            return;
          if (dist * dist > chi2max * (ezt2[i] + ezt2[j]))
            return;
          mdist = dist;
          iv[i] = int32_t(j);
        });
      }
      for (uint32_t i = 0; i < nt; ++i) {
        int32_t m = iv[i];
        while (m != iv[m])
          m = iv[m];
        iv[i] = m;
      }
    }
    if (algo == 1 || algo == 2) {
      // edges (both algorithms): non-core tracks join the nearest core track
      for (uint32_t i = 0; i < nt; ++i) {
        if (nn[i] >= minT)
          continue;
        float mdist = eps;
        forEachNear(izt[i], [&](uint32_t j) {
          if (nn[j] < minT)
            return;
          float dist = std::abs(zt[i] - zt[j]);
          if (dist > mdist)
            return;
          if (dist * dist > chi2max * (ezt2[i] + ezt2[j]))
            return;
          mdist = dist;
          iv[i] = iv[j];
        });
      }
    }
    uint32_t found = 0;
    for (uint32_t i = 0; i < nt; ++i) {
      if (iv[i] == int32_t(i)) {
        if (nn[i] >= minT)
          iv[i] = -int32_t(++found);
        else
          iv[i] = -9998;
      }
    }
    for (uint32_t i = 0; i < nt; ++i)
      if (iv[i] >= 0)
        iv[i] = iv[iv[i]];
    for (uint32_t i = 0; i < nt; ++i)
      iv[i] = -iv[i] - 1;
    return found;
  }

  // density clustering, fit, split, refit and sort over loaded tracks
  void clusterAndFit(VertexState& V, const std::vector<TupleIndex>& itrk, const std::vector<float>& zt,
                     const std::vector<float>& ezt2, const std::vector<float>& ptt2, int algo) {
    V.zv.assign(kMaxVertices, 0);
    V.wv.assign(kMaxVertices, 0);
    V.chi2.assign(kMaxVertices, 0);
    V.ptv2.assign(kMaxVertices, 0);
    V.sortInd.assign(kMaxVertices, 0);
    const uint32_t nt = uint32_t(itrk.size());
    std::vector<int32_t> iv(nt);
    int32_t* nn = V.ndof.data();
    V.ntrks = nt;
    const uint32_t found = clusterTracks(algo, nt, zt.data(), ezt2.data(), iv.data(), nn);
    uint32_t nvIntermediate = found;
    uint32_t nvFinal = found;
    V.nFound = found;
    if (found >= kMaxVertices)
      V.status |= 1;

    // fitVertices (gpuFitVertices.h:15-98)
    auto fit = [&](float chi2Max) {
      nvFinal = nvIntermediate;
      const uint32_t nc = std::min(nvFinal, kMaxVertices);
      for (uint32_t i = 0; i < nc; ++i) {
        V.zv[i] = 0;
        V.wv[i] = 0;
        V.chi2[i] = 0;
      }
      // (vertex ids beyond the container, possible only when it overflowed,
      // count as unassigned)
      for (uint32_t i = 0; i < nt; ++i) {
        if (iv[i] > 9990 || iv[i] >= int32_t(nc))
          continue;
        float w = 1.f / ezt2[i];
        V.zv[iv[i]] += zt[i] * w;
        V.wv[iv[i]] += w;
      }
      for (uint32_t i = 0; i < nc; ++i) {
        V.zv[i] /= V.wv[i];
        nn[i] = -1;
      }
      for (uint32_t i = 0; i < nt; ++i) {
        if (iv[i] > 9990 || iv[i] >= int32_t(nc))
          continue;
        float c2 = V.zv[iv[i]] - zt[i];
        c2 *= c2 / ezt2[i];
        if (c2 > chi2Max) {
          iv[i] = 9999;
          continue;
        }
        V.chi2[iv[i]] += c2;
        nn[iv[i]] += 1;
      }
      for (uint32_t i = 0; i < nc; ++i)
        if (nn[i] > 0)
          V.wv[i] *= float(nn[i]) / V.chi2[i];
    };
    fit(50.f);

    // splitVertices (gpuSplitVertices.h:15-125)
    const float maxChi2 = 9.f;
    for (uint32_t kv = 0; kv < nvFinal && kv < kMaxVertices; ++kv) {
      if (nn[kv] < 4)
        continue;
      if (V.chi2[kv] < maxChi2 * float(nn[kv]))
        continue;
      constexpr int MAXTK = 512;
      if (nn[kv] >= MAXTK)
        continue;
      uint32_t it[MAXTK];
      float zz[MAXTK], ww[MAXTK];
      uint8_t newV[MAXTK];
      uint32_t nq = 0;
      for (uint32_t k = 0; k < nt; ++k) {
        if (iv[k] == int32_t(kv)) {
          uint32_t old = nq;
          if (nq < MAXTK)
            ++nq;
          zz[old] = zt[k] - V.zv[kv];
          newV[old] = zz[old] < 0 ? 0 : 1;
          ww[old] = 1.f / ezt2[k];
          it[old] = k;
        }
      }
      float znew[2], wnew[2];
      int maxiter = 20;
      bool more = true;
      while (more) {
        more = false;
        znew[0] = znew[1] = 0;
        wnew[0] = wnew[1] = 0;
        for (uint32_t k = 0; k < nq; ++k) {
          auto i = newV[k];
          znew[i] += zz[k] * ww[k];
          wnew[i] += ww[k];
        }
        znew[0] /= wnew[0];
        znew[1] /= wnew[1];
        for (uint32_t k = 0; k < nq; ++k) {
          auto d0 = std::fabs(zz[k] - znew[0]);
          auto d1 = std::fabs(zz[k] - znew[1]);
          auto newer = d0 < d1 ? 0 : 1;
          more |= newer != newV[k];
          newV[k] = uint8_t(newer);
        }
        --maxiter;
        if (maxiter <= 0)
          more = false;
      }
      if (0 == wnew[0] || 0 == wnew[1])
        continue;
      float dist2 = (znew[0] - znew[1]) * (znew[0] - znew[1]);
      float chi2Dist = dist2 / (1.f / wnew[0] + 1.f / wnew[1]);
      if (chi2Dist < 4)
        continue;
      uint32_t igv = nvIntermediate++;
      for (uint32_t k = 0; k < nq; ++k)
        if (1 == newV[k])
          iv[it[k]] = int32_t(igv);
    }
    fit(5000.f);

    // sortByPt2 (gpuSortByPt2.h:15-55)
    V.nvFinal = nvFinal;
    if (nvFinal >= 1) {
      for (uint32_t i = 0; i < nt; ++i)
        V.idv[itrk[i]] = int16_t(iv[i]);
      const uint32_t nc = std::min(nvFinal, kMaxVertices);
      for (uint32_t i = 0; i < nc; ++i)
        V.ptv2[i] = 0;
      for (uint32_t i = 0; i < nt; ++i) {
        if (iv[i] > 9990 || iv[i] >= int32_t(nc))
          continue;
        V.ptv2[iv[i]] += ptt2[i];
      }
      if (nvFinal == 1) {
        V.sortInd[0] = 0;
      } else {
        for (uint16_t i = 0; i < nc; ++i)
          V.sortInd[i] = i;
        std::stable_sort(V.sortInd.begin(), V.sortInd.begin() + nc,
                         [&](uint16_t a, uint16_t b) { return V.ptv2[a] < V.ptv2[b]; });
      }
    }
    V.zt.assign(zt.begin(), zt.end());
    V.iv.assign(iv.begin(), iv.end());
  }

  int verticesDirect(uint32_t nt, const float* zt, const float* ezt2, const float* ptt2, int16_t* iv, float* zv,
                     float* wv, float* chi2, float* ptv2, int32_t* ndof, uint16_t* sortInd, int algo) {
    VertexState V;
    V.idv.assign(nt, -1);
    V.ndof.assign(std::max<uint32_t>(nt, kMaxVertices), 0);
    std::vector<TupleIndex> itrk(nt);
    for (uint32_t i = 0; i < nt; ++i)
      itrk[i] = i;
    clusterAndFit(V, itrk, std::vector<float>(zt, zt + nt), std::vector<float>(ezt2, ezt2 + nt),
                  std::vector<float>(ptt2, ptt2 + nt), algo);
    for (uint32_t i = 0; i < nt; ++i)
      iv[i] = V.idv[i];
    for (uint32_t k = 0; k < V.nvFinal && k < kMaxVertices; ++k) {
      zv[k] = V.zv[k];
      wv[k] = V.wv[k];
      chi2[k] = V.chi2[k];
      ptv2[k] = V.ptv2[k];
      ndof[k] = V.ndof[k];
      sortInd[k] = V.sortInd[k];
    }
    return int(V.nvFinal);
  }

}  // namespace orc

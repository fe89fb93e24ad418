// Primary vertices on gfx950 (a23 loadTracks, a24 clusterTracksByDensity +
// fitVertices + splitVertices, a25 sortByPt2: plugin-PixelVertexFinding
// gpuVertexFinderImpl.h:10-111, gpuClusterTracksByDensity.h:20-207,
// gpuFitVertices.h:15-98, gpuSplitVertices.h:15-125, gpuSortByPt2.h:15-55).
//
// One workgroup per event.  Every float sum of the serial program runs over
// tracks in ascending order; here each vertex owns a list of its tracks in
// ascending order (a stable counting sort) and one lane per vertex sums its
// list, so the sums are the serial ones bit for bit.
#include <hip/hip_runtime.h>

#include "ca_device.h"

namespace pxg {

  namespace {
    constexpr int kVB = kEventBlock;
    constexpr int kUnassigned = 9999;
    constexpr uint32_t kLdsTracks = 2048;

    struct VtxLds {
      uint32_t counts[kMaxVertices];
      uint32_t start[kMaxVertices + 1];
      uint32_t running[kMaxVertices];
      uint32_t scan[kVB / 64];
      uint32_t nvIntermediate, nvFinal, found, nt;
      uint8_t waveCnt[kVB / 64][kMaxVertices];  // items of a bin in each wave of a batch (<= 64)
    };

    // Stable counting sort of items 0..n-1 by key(i) in [0, K) (items with a
    // key >= K are left out): list gets the items bin by bin, ascending item
    // index within a bin (descending if `descending`); L.start[0..K] the bin starts.
    template <typename Key>
    __device__ void countingSort(uint32_t n, uint32_t K, Key key, bool descending, uint32_t* list, VtxLds& L) {
      for (uint32_t k = threadIdx.x; k < K; k += kVB) {
        L.counts[k] = 0;
        L.running[k] = 0;
      }
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < n; i += kVB) {
        const uint32_t k = key(i);
        if (k < K)
          atomicAdd(&L.counts[k], 1u);
      }
      __syncthreads();
      uint32_t carry = 0;
      for (uint32_t base = 0; base < K; base += kVB) {
        const uint32_t k = base + threadIdx.x;
        const uint32_t v = k < K ? L.counts[k] : 0;
        uint32_t tot;
        const uint32_t ex = wgExclusiveScan<kVB>(v, tot, L.scan);
        if (k < K)
          L.start[k] = carry + ex;
        carry += tot;
      }
      if (threadIdx.x == 0)
        L.start[K] = carry;
      __syncthreads();
      // rank of an item among the earlier items of its bin: within its wave by
      // shuffles, from the earlier waves of the batch by their per-bin counts,
      // from the earlier batches by the running count
      uint32_t* wc = reinterpret_cast<uint32_t*>(&L.waveCnt[0][0]);
      for (uint32_t w = threadIdx.x; w < (kVB / 64) * kMaxVertices / 4; w += kVB)
        wc[w] = 0;
      __syncthreads();
      const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
      for (uint32_t base = 0; base < n; base += kVB) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t k = i < n ? key(i) : 0xffffffffu;
        uint32_t below = 0, same = 0;
        for (uint32_t src = 0; src < 64; ++src) {
          const uint32_t ks = __shfl(k, src);
          same += ks == k;
          below += ks == k && src < lane;
        }
        const bool mine = i < n && k < K;
        const bool last = mine && below == same - 1;  // the wave's last item of the bin
        if (last)
          L.waveCnt[wid][k] = uint8_t(same);
        __syncthreads();
        if (mine) {
          uint32_t rank = L.running[k] + below;
          for (uint32_t w = 0; w < wid; ++w)
            rank += L.waveCnt[w][k];
          const uint32_t pos = descending ? L.start[k] + L.counts[k] - 1 - rank : L.start[k] + rank;
          list[pos] = i;
        }
        __syncthreads();
        if (last) {
          atomicAdd(&L.running[k], same);
          L.waveCnt[wid][k] = 0;
        }
        __syncthreads();
      }
    }

    __global__ void __launch_bounds__(kVB) vertexKernel(CAPool ca, int algo, float ptMin) {
      const uint32_t e = blockIdx.x;
      __shared__ VtxLds L;
      const size_t tb = ca.tupleBase(e);
      const uint32_t M = ca.caps.maxTuples;
      const uint32_t nT = ca.cnt[e].nTuples;
      const uint32_t* toff = ca.tupleOff + ca.offBase(e);
      const uint8_t* q = ca.quality + tb;
      int16_t* idv = ca.idv + tb;
      int32_t* const nnOut = ca.ndof + tb;
      int32_t* nn = nnOut;
      TupleIndex* itrk = ca.vItrk + tb;
      float* zt = ca.vZt + tb;
      float* ezt2 = ca.vEzt2 + tb;
      float* ptt2 = ca.vPtt2 + tb;
      int32_t* iv = ca.vIv + tb;
      int32_t* tmp = ca.vTmp + tb;
      uint8_t* izt = ca.vIzt + tb;
      uint32_t* zbins = ca.vTrk + tb;  // tracks by z bin, then tracks by vertex
      const uint32_t* voff = L.start;  // bin starts of the last counting sort
      // the per-vertex results live in LDS and are stored once at the end
      __shared__ float zv[kMaxVertices], wv[kMaxVertices], chi2[kMaxVertices], ptv2[kMaxVertices];
      uint16_t* sortInd = ca.sortInd + size_t(e) * kMaxVertices;
      constexpr int minT = 2;
      constexpr float eps = 0.07f, errmax = 0.01f, chi2max = 9.f;

      // ---- loadTracks: loose tuples with 4+ hits and pt >= 0.5 GeV, in tuple order.
      // The tuples go in chunks of 64 * kVB; in a chunk each thread takes a contiguous
      // run of at most 64 tuples and one scan over the threads orders the taken ones.
      {
        uint32_t taken = 0;  // tracks of the earlier chunks
        for (uint32_t c0 = 0; c0 < nT; c0 += 64 * kVB) {
          const uint32_t n = min(nT - c0, 64u * kVB);
          const uint32_t R = (n + kVB - 1) / kVB;
          const uint32_t t0 = c0 + min(threadIdx.x * R, n), t1 = min(t0 + R, c0 + n);
          uint64_t takeMask = 0;
          uint32_t cnt = 0;
          for (uint32_t t = t0; t < t1; ++t) {
            const uint32_t nh = toff[t + 1] - toff[t];
            idv[t] = nh == 0 ? 0 : -1;  // only the last stored tuple can be empty
            const bool take = nh >= 4 && q[t] == kQualityLoose && !(ca.pt[tb + t] < ptMin);
            takeMask |= uint64_t(take) << (t - t0);
            cnt += take;
          }
          uint32_t tot;
          uint32_t k = taken + wgExclusiveScan<kVB>(cnt, tot, L.scan);
          for (uint64_t m = takeMask; m; m &= m - 1) {
            const uint32_t t = t0 + __ffsll((unsigned long long)m) - 1;
            const float pt = ca.pt[tb + t];
            itrk[k] = t;
            zt[k] = ca.state[5 * tb + 4 * size_t(M) + t];
            ezt2[k] = ca.cov[15 * tb + 14 * size_t(M) + t];
            ptt2[k] = pt * pt;
            ++k;
          }
          taken += tot;
        }
        if (threadIdx.x == 0)
          L.nt = taken;
        __syncthreads();
      }
      const uint32_t nt = L.nt;
      // events with up to kLdsTracks tracks keep the per-track arrays in LDS
      __shared__ float sZt[kLdsTracks], sEzt2[kLdsTracks];
      __shared__ int32_t sIv[kLdsTracks], sTmp[kLdsTracks], sNn[kLdsTracks];
      __shared__ uint32_t sBins[kLdsTracks];
      __shared__ uint8_t sIzt[kLdsTracks];
      __shared__ float sZs[kLdsTracks], sEs[kLdsTracks];
      __shared__ int32_t sNnS[kLdsTracks];
      float* zS = ca.vZs + tb;
      float* eS = ca.vEs + tb;
      int32_t* nnS = ca.vNnS + tb;
      const bool inLds = nt <= kLdsTracks;
      if (inLds) {
        for (uint32_t i = threadIdx.x; i < nt; i += kVB) {
          sZt[i] = zt[i];
          sEzt2[i] = ezt2[i];
        }
        zt = sZt;
        ezt2 = sEzt2;
        iv = sIv;
        tmp = sTmp;
        nn = sNn;
        zbins = sBins;
        izt = sIzt;
        zS = sZs;
        eS = sEs;
        nnS = sNnS;
        __syncthreads();
      }

      // ---- clusterTracksByDensity
      const float er2mx = errmax * errmax;
      for (uint32_t i = threadIdx.x; i < nt; i += kVB) {
        int iz = int(zt[i] * 10.);
        iz = min(max(iz, -128), 127);
        izt[i] = uint8_t(iz + 128);
        iv[i] = int32_t(i);
        nn[i] = 0;
      }
      __syncthreads();
      countingSort(nt, 256, [&](uint32_t i) { return uint32_t(izt[i]); }, true, zbins, L);
      __syncthreads();
      // z-ordered copies of z, its error and (below) the neighbour count: the
      // density loops read their candidates by z-ordered position, without
      // going through the track index
      for (uint32_t pp = threadIdx.x; pp < nt; pp += kVB) {
        const uint32_t i = zbins[pp];
        zS[pp] = zt[i];
        eS[pp] = ezt2[i];
        nnS[pp] = 0;
      }
      __syncthreads();
      // z-ordered positions of the tracks of the neighbouring bins, in histogram order
      auto nearRange = [&](uint32_t i, uint32_t& pb, uint32_t& pe) {
        const int bs = max(0, int(izt[i]) - 1);
        const int be = min(255, int(izt[i]) + 1);
        pb = voff[bs];
        pe = voff[be + 1];
      };
      // tracks taken in z order, so the lanes of a wave mostly scan the same bins
      for (uint32_t pp = threadIdx.x; pp < nt; pp += kVB) {
        const uint32_t i = zbins[pp];
        const float zi = zS[pp], ei = eS[pp];
        if (ei > er2mx)
          continue;
        uint32_t pb, pe;
        nearRange(i, pb, pe);
        // a neighbour: another track within eps and (not for DBSCAN) within the
        // error-scaled cut
        auto near = [&](uint32_t p, float zj, float ej) -> int32_t {
          const float dist = fabsf(zi - zj);
          return p != pp && !(dist > eps) && !(algo != kVertexDBSCAN && dist * dist > chi2max * (ei + ej));
        };
        int32_t c = 0;
        uint32_t p = pb;
        for (; p + 4 <= pe; p += 4) {
          const float z0 = zS[p], z1 = zS[p + 1], z2 = zS[p + 2], z3 = zS[p + 3];
          const float e0 = eS[p], e1 = eS[p + 1], e2 = eS[p + 2], e3 = eS[p + 3];
          c += near(p, z0, e0) + near(p + 1, z1, e1) + near(p + 2, z2, e2) + near(p + 3, z3, e3);
        }
        for (; p < pe; ++p)
          c += near(p, zS[p], eS[p]);
        nn[i] = c;
        nnS[pp] = c;
      }
      __syncthreads();
      if (algo == kVertexDensity) {
        // clusterTracksByDensity: each track links to the closest track with more
        // neighbours (or as many at smaller z)
        for (uint32_t pp = threadIdx.x; pp < nt; pp += kVB) {
          const uint32_t i = zbins[pp];
          const float zi = zS[pp], ei = eS[pp];
          const int32_t ni = nnS[pp];
          uint32_t pb, pe;
          nearRange(i, pb, pe);
          float mdist = eps;
          uint32_t bestP = 0xffffffffu;
          // the closest qualifying track, first one on ties (candidates in order)
          auto visit = [&](uint32_t p, int32_t nj, float zj, float ej) {
            if (nj < ni)
              return;
            if (nj == ni && zj >= zi)
              return;
            const float dist = fabsf(zi - zj);
            if (dist > mdist)
              return;
            if (dist * dist > chi2max * (ei + ej))
              return;
            mdist = dist;
            bestP = p;
          };
          uint32_t p = pb;
          for (; p + 4 <= pe; p += 4) {
            const int32_t n0 = nnS[p], n1 = nnS[p + 1], n2 = nnS[p + 2], n3 = nnS[p + 3];
            const float z0 = zS[p], z1 = zS[p + 1], z2 = zS[p + 2], z3 = zS[p + 3];
            const float e0 = eS[p], e1 = eS[p + 1], e2 = eS[p + 2], e3 = eS[p + 3];
            visit(p, n0, z0, e0);
            visit(p + 1, n1, z1, e1);
            visit(p + 2, n2, z2, e2);
            visit(p + 3, n3, z3, e3);
          }
          for (; p < pe; ++p)
            visit(p, nnS[p], zS[p], eS[p]);
          iv[i] = bestP == 0xffffffffu ? int32_t(i) : int32_t(zbins[bestP]);
        }
      } else if (algo == kVertexDBSCAN) {
        // clusterTracksDBSCAN (gpuClusterTracksDBSCAN.h:100-119): a core track
        // links to its lowest-z core neighbour within eps, the first one in
        // histogram order on a tie
        for (uint32_t pp = threadIdx.x; pp < nt; pp += kVB) {
          if (nnS[pp] < minT)
            continue;
          const uint32_t i = zbins[pp];
          const float zi = zS[pp];
          uint32_t pb, pe;
          nearRange(i, pb, pe);
          float mz = zi;
          uint32_t linkP = 0xffffffffu;
          for (uint32_t p = pb; p < pe; ++p) {
            const float zj = zS[p];
            if (zj >= mz)
              continue;
            if (nnS[p] < minT)
              continue;
            if (fabsf(zi - zj) > eps)
              continue;
            mz = zj;
            linkP = p;
          }
          if (linkP != 0xffffffffu)
            iv[i] = int32_t(zbins[linkP]);
        }
      } else {
        // clusterTracksIterative (gpuClusterTracksIterative.h:101-146): the core
        // tracks connected by neighbour pairs share the smallest track index
        // among them.  Here: each core track takes the smallest label of its
        // core neighbours, then labels are followed to their roots, until
        // nothing changes (labels only decrease and never exceed the index).
        for (uint32_t r = 0; r <= nt; ++r) {
          bool changed = false;
          for (uint32_t pp = threadIdx.x; pp < nt; pp += kVB) {
            if (nnS[pp] < minT)
              continue;
            const uint32_t i = zbins[pp];
            const float zi = zS[pp], ei = eS[pp];
            uint32_t pb, pe;
            nearRange(i, pb, pe);
            int32_t m = iv[i];
            for (uint32_t p = pb; p < pe; ++p) {
              if (p == pp || nnS[p] < minT)
                continue;
              const float ej = eS[p];
              const float dist = fabsf(zi - zS[p]);
              if (dist > eps || dist * dist > chi2max * (ei + ej))
                continue;
              m = min(m, iv[zbins[p]]);
            }
            if (m < iv[i]) {
              atomicMin(&iv[i], m);
              changed = true;
            }
          }
          __syncthreads();
          for (uint32_t i = threadIdx.x; i < nt; i += kVB) {
            int32_t m = iv[i];
            for (uint32_t s2 = 0; m != iv[m] && s2 < nt; ++s2)
              m = iv[m];
            atomicMin(&iv[i], m);
          }
          if (!__syncthreads_or(changed))
            break;
        }
      }
      __syncthreads();
      // follow the links to the root of each tree
      for (uint32_t i = threadIdx.x; i < nt; i += kVB) {
        // links go to strictly more neighbours, or as many at smaller z: no cycles
        int32_t m = iv[i];
        for (uint32_t s = 0; m != iv[m] && s < nt; ++s)
          m = iv[m];
        tmp[i] = m;
      }
      __syncthreads();
      if (algo != kVertexDensity) {
        // edges (gpuClusterTracksDBSCAN.h:156-175): a track that is not core
        // joins the root of the nearest core track within eps and the
        // error-scaled cut, the last one in histogram order on a tie
        for (uint32_t pp = threadIdx.x; pp < nt; pp += kVB) {
          if (nnS[pp] >= minT)
            continue;
          const uint32_t i = zbins[pp];
          const float zi = zS[pp], ei = eS[pp];
          uint32_t pb, pe;
          nearRange(i, pb, pe);
          float mdist = eps;
          uint32_t nearP = 0xffffffffu;
          for (uint32_t p = pb; p < pe; ++p) {
            if (nnS[p] < minT)
              continue;
            const float dist = fabsf(zi - zS[p]);
            if (dist > mdist)
              continue;
            if (dist * dist > chi2max * (ei + eS[p]))
              continue;
            mdist = dist;
            nearP = p;
          }
          if (nearP != 0xffffffffu)
            tmp[i] = tmp[zbins[nearP]];
        }
        __syncthreads();
      }
      // roots with enough neighbours become vertices, numbered in track order
      if (threadIdx.x == 0)
        L.found = 0;
      __syncthreads();
      for (uint32_t base = 0; base < nt; base += kVB) {
        const uint32_t i = base + threadIdx.x;
        const bool root = i < nt && tmp[i] == int32_t(i);
        const bool seed = root && nn[i] >= minT;
        uint32_t tot;
        const uint32_t ex = wgExclusiveScan<kVB>(seed ? 1u : 0u, tot, L.scan);
        if (root)
          iv[i] = seed ? -int32_t(L.found + ex + 1) : -9998;
        __syncthreads();
        if (threadIdx.x == 0)
          L.found += tot;
        __syncthreads();
      }
      for (uint32_t i = threadIdx.x; i < nt; i += kVB)
        if (tmp[i] != int32_t(i))
          iv[i] = iv[tmp[i]];
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < nt; i += kVB)
        iv[i] = -iv[i] - 1;
      if (threadIdx.x == 0) {
        L.nvIntermediate = L.found;
        L.nvFinal = L.found;
        ca.cnt[e].vstatus = L.found >= kMaxVertices ? 1 : 0;
        ca.cnt[e].vtxTracks = nt;
      }
      __syncthreads();

      // ---- fitVertices (twice) and splitVertices
      // the tracks of vertex v in ascending order (four indices loaded at a time)
      auto forEachOf = [&](uint32_t v, auto&& f) {
        uint32_t p = voff[v];
        const uint32_t pe = voff[v + 1];
        for (; p + 4 <= pe; p += 4) {
          const uint32_t i0 = zbins[p], i1 = zbins[p + 1], i2 = zbins[p + 2], i3 = zbins[p + 3];
          f(i0);
          f(i1);
          f(i2);
          f(i3);
        }
        for (; p < pe; ++p)
          f(zbins[p]);
      };
      auto fit = [&](float chi2Max) {
        const uint32_t nvf = L.nvIntermediate;
        const uint32_t nc = min(nvf, kMaxVertices);
        __syncthreads();
        if (threadIdx.x == 0)
          L.nvFinal = nvf;
        countingSort(nt, nc, [&](uint32_t i) { return uint32_t(iv[i]); }, false, zbins, L);
        __syncthreads();
        for (uint32_t v = threadIdx.x; v < nc; v += kVB) {
          float z = 0, w = 0;
          forEachOf(v, [&](uint32_t i) {
            const float wi = 1.f / ezt2[i];
            z += zt[i] * wi;
            w += wi;
          });
          zv[v] = z / w;
          wv[v] = w;
          nn[v] = -1;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nt; i += kVB) {
          if (iv[i] > 9990 || iv[i] >= int32_t(nc))
            continue;
          float c2 = zv[iv[i]] - zt[i];
          c2 *= c2 / ezt2[i];
          if (c2 > chi2Max)
            iv[i] = kUnassigned;
          else
            tmp[i] = __float_as_int(c2);
        }
        __syncthreads();
        for (uint32_t v = threadIdx.x; v < nc; v += kVB) {
          float c = 0;
          int32_t k = -1;
          forEachOf(v, [&](uint32_t i) {
            // both loads unconditional (i < nt), so a batch's loads issue together;
            // tmp[i] is used only for the vertex's own tracks
            const int32_t ivi = iv[i];
            const float ci = __int_as_float(tmp[i]);
            if (ivi != int32_t(v))
              return;
            c += ci;
            ++k;
          });
          chi2[v] = c;
          nn[v] = k;
          if (k > 0)
            wv[v] *= float(k) / c;
        }
        __syncthreads();
      };
      fit(50.f);

      // splitVertices: each vertex's 2-means runs on its own tracks; new
      // vertices are numbered in the order of the vertices they split from.
      // A vertex to split gets a wave: every lane runs the same sequential sums
      // over the vertex's tracks in ascending order (the serial sums, bit for
      // bit), and the lanes reassign the tracks in parallel.
      {
        const uint32_t nvf = L.nvFinal;
        const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        for (uint32_t base = 0; base < nvf; base += kVB) {
          const uint32_t kv = base + threadIdx.x;
          // the vertex's assigned tracks in ascending order, compacted into its own
          // segment [b, b + m) of arrays that are free here (the z-ordered copies
          // and tmp): offset from the vertex, weight, offset x weight, and the
          // track index with the side in bit 31
          const uint32_t b = kv < nvf && kv < kMaxVertices ? voff[kv] : 0;
          float* const czz = zS;
          float* const cww = eS;
          float* const cprod = reinterpret_cast<float*>(nnS);
          uint32_t* const cidx = reinterpret_cast<uint32_t*>(tmp);
          uint32_t m = 0;
          const bool cand =
              kv < nvf && kv < kMaxVertices && nn[kv] >= 4 && !(chi2[kv] < 9.f * float(nn[kv])) && nn[kv] < 512;
          if (cand) {
            forEachOf(kv, [&](uint32_t i) {
              if (iv[i] != int32_t(kv))
                return;
              const float zz = zt[i] - zv[kv];
              const float ww = 1.f / ezt2[i];
              czz[b + m] = zz;
              cww[b + m] = ww;
              cprod[b + m] = zz * ww;
              cidx[b + m] = i | (zz < 0 ? 0u : 0x80000000u);
              ++m;
            });
          }
          // L.counts: the vertex's track count (bit 31: split); L.running: the candidates
          if (kv < kMaxVertices)
            L.counts[kv] = m;
          uint32_t nCand;
          const uint32_t ci = wgExclusiveScan<kVB>(cand ? 1u : 0u, nCand, L.scan);
          if (cand)
            L.running[ci] = kv;
          __syncthreads();
          for (uint32_t c = wid; c < nCand; c += kVB / 64) {
            const uint32_t v = L.running[c];
            const uint32_t vb = voff[v], vm = L.counts[v];
            float znew[2], wnew[2];
            int maxiter = 20;
            bool more = true;
            while (more) {
              znew[0] = znew[1] = 0;
              wnew[0] = wnew[1] = 0;
              // eight tracks' inputs are loaded before they are added, in order
              uint32_t u = 0;
              for (; u + 8 <= vm; u += 8) {
                float pr[8], ww[8];
                uint32_t sd[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                  pr[t] = cprod[vb + u + t];
                  ww[t] = cww[vb + u + t];
                  sd[t] = cidx[vb + u + t] >> 31;
                }
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                  if (sd[t]) {
                    znew[1] += pr[t];
                    wnew[1] += ww[t];
                  } else {
                    znew[0] += pr[t];
                    wnew[0] += ww[t];
                  }
                }
              }
              for (; u < vm; ++u) {
                const float pr = cprod[vb + u], ww = cww[vb + u];
                if (cidx[vb + u] >> 31) {
                  znew[1] += pr;
                  wnew[1] += ww;
                } else {
                  znew[0] += pr;
                  wnew[0] += ww;
                }
              }
              znew[0] /= wnew[0];
              znew[1] /= wnew[1];
              bool changed = false;
              for (uint32_t u = lane; u < vm; u += 64) {
                const float zz = czz[vb + u];
                const float d0 = fabsf(zz - znew[0]);
                const float d1 = fabsf(zz - znew[1]);
                const uint32_t newer = d0 < d1 ? 0u : 0x80000000u;
                const uint32_t x = cidx[vb + u];
                changed |= newer != (x & 0x80000000u);
                cidx[vb + u] = (x & 0x7fffffffu) | newer;
              }
              // the next pass's sums read what other lanes wrote (in LDS, or in
              // global memory for events with more than kLdsTracks tracks)
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
              __builtin_amdgcn_wave_barrier();
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
              more = __ballot(changed) != 0;
              --maxiter;
              if (maxiter <= 0)
                more = false;
            }
            bool split = false;
            if (!(0 == wnew[0] || 0 == wnew[1])) {
              const float dist2 = (znew[0] - znew[1]) * (znew[0] - znew[1]);
              const float chi2Dist = dist2 / (1.f / wnew[0] + 1.f / wnew[1]);
              split = !(chi2Dist < 4);
            }
            if (lane == 0 && split)
              L.counts[v] = vm | 0x80000000u;
          }
          __syncthreads();
          const bool split = kv < kMaxVertices && (L.counts[kv] >> 31) != 0;
          uint32_t tot;
          const uint32_t ex = wgExclusiveScan<kVB>(split ? 1u : 0u, tot, L.scan);
          if (split) {
            // the new id is >= nvf, so no other vertex of this pass sees the moved tracks
            const int32_t igv = int32_t(L.nvIntermediate + ex);
            for (uint32_t u = 0; u < m; ++u) {
              const uint32_t x = cidx[b + u];
              if (x >> 31)
                iv[x & 0x7fffffffu] = igv;
            }
          }
          __syncthreads();
          if (threadIdx.x == 0)
            L.nvIntermediate += tot;
          __syncthreads();
        }
      }
      fit(5000.f);

      // ---- sortByPt2
      const uint32_t nvFinal = L.nvFinal;
      if (inLds)
        for (uint32_t i = threadIdx.x; i < nt; i += kVB)
          nnOut[i] = sNn[i];
      if (threadIdx.x == 0)
        ca.cnt[e].nVertices = nvFinal;
      if (nvFinal >= 1) {
        for (uint32_t i = threadIdx.x; i < nt; i += kVB)
          idv[itrk[i]] = int16_t(iv[i]);
        const uint32_t nc = min(nvFinal, kMaxVertices);
        // the tracks' pt^2 next to the other per-track arrays (tmp is free now)
        const float* pt2 = ptt2;
        if (inLds) {
          for (uint32_t i = threadIdx.x; i < nt; i += kVB)
            tmp[i] = __float_as_int(ptt2[i]);
          __syncthreads();
          pt2 = reinterpret_cast<const float*>(tmp);
        }
        for (uint32_t v = threadIdx.x; v < nc; v += kVB) {
          float s = 0;
          for (uint32_t p = voff[v]; p < voff[v + 1]; ++p) {
            const uint32_t i = zbins[p];
            if (iv[i] == int32_t(v))
              s += pt2[i];
          }
          ptv2[v] = s;
        }
        __syncthreads();
        for (uint32_t v = threadIdx.x; v < nc; v += kVB) {
          uint32_t r = 0;
          const float pv = ptv2[v];
          for (uint32_t u = 0; u < nc; ++u)
            r += (ptv2[u] < pv) || (ptv2[u] == pv && u < v);
          sortInd[r] = uint16_t(v);
        }
        float* gZv = ca.zv + size_t(e) * kMaxVertices;
        float* gWv = ca.wv + size_t(e) * kMaxVertices;
        float* gChi2 = ca.vchi2 + size_t(e) * kMaxVertices;
        float* gPtv2 = ca.ptv2 + size_t(e) * kMaxVertices;
        for (uint32_t v = threadIdx.x; v < nc; v += kVB) {
          gZv[v] = zv[v];
          gWv[v] = wv[v];
          gChi2[v] = chi2[v];
          gPtv2[v] = ptv2[v];
        }
      }
    }
  }  // namespace

  void launchVertices(const CALaunch& l) {
    l.t->b("vertices");
    vertexKernel<<<l.nEvents, kVB, 0, l.s>>>(*l.ca, l.vertexAlgo, l.vertexPtMin);
    l.t->e();
  }

}  // namespace pxg

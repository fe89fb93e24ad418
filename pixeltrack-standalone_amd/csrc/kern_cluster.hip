// Stage A, part 2: per-module clustering with the cluster charge cut, and
// the clusters-per-module prefix.
//
// The serial program labels every pixel of a module with the smallest pixel
// index of its 8-connected component and numbers the components by that
// index (gpuClustering.h:40-273), then drops clusters under the charge cut
// and renumbers the survivors in order (gpuClusterChargeCut.h:14-114).  The
// result depends only on the connected components, so any exact algorithm
// gives the same labels; two are used here, picked per module run:
//
//   clusterSmallKernel  two module runs of at most 64 digis per wavefront
//                       (most modules at PU50 carry ~30: on the two halves when
//                       both fit 32 lanes): a neighbour mask per lane from
//                       shuffles, min-label propagation
//                       with pointer jumping through a per-wave LDS row until
//                       a ballot sees no change, roots ranked with popcounts.
//                       No workgroup barrier, no atomics but the charge sums.
//                       The surviving clusters' extents and charges, which the
//                       rec-hit stage needs, are accumulated in LDS on the way
//                       and written once per cluster (ClusterExt).
//                       Larger runs are listed per event in order by the
//                       module-start scan.
//   clusterMediumKernel runs of 65 to 256 digis, one wavefront each (a
//   clusterBigKernel    workgroup of one wave), and the larger ones, a few
//                       256-thread workgroups per event: pixels staged in LDS,
//                       column histogram, lock-free union-find that hooks the
//                       larger root under the smaller (so the root of a
//                       component is its smallest index), roots numbered by a
//                       workgroup scan.
//   clusStartKernel     one workgroup per event: the clamped prefix over the
//                       2000 cluster counts (SiPixelRawToClusterGPUKernel.cc:
//                       464-506).
#include "primitives.h"
#include "pxgpu_internal.h"

namespace pxg {

  namespace {
    constexpr int kBlock = 256;
    constexpr int kWaves = kBlock / 64;
    constexpr uint32_t kNoPixel = 0xffffffffu;
    constexpr int kCols = kColsPerModule;  // 416
    // workgroups per event for the runs of 65 to 256 digis (one wave each, ~4096 per
    // launch) and for the larger ones (a few per event at most)
    constexpr uint32_t kMediumWorkgroups = 4096, kMediumPerEventMax = 64, kBigPerEvent = 2;

    __device__ __forceinline__ uint32_t findRoot(volatile uint32_t* parent, uint32_t x) {
      uint32_t p = parent[x];
      while (p != x) {
        x = p;
        p = parent[x];
      }
      return x;
    }

    // unite the sets of a and b; the larger root is hooked under the smaller
    __device__ __forceinline__ void unite(uint32_t* parent, uint32_t a, uint32_t b) {
      while (true) {
        a = findRoot(parent, a);
        b = findRoot(parent, b);
        if (a == b)
          return;
        if (a > b) {
          uint32_t t = a;
          a = b;
          b = t;
        }
        const uint32_t old = atomicMin(&parent[b], a);
        if (old == b)
          return;
        b = old;  // b was hooked meanwhile: unite a with its new parent
      }
    }
  }  // namespace

  // LDS of one module run of at most MAXPIX digis on a workgroup of NT threads
  template <int NT, uint32_t MAXPIX>
  struct RunLds {
    static constexpr uint32_t kClusCap = MAXPIX < kMaxClustersPerModule ? MAXPIX : kMaxClustersPerModule;
    uint32_t sxy[MAXPIX];     // x | y << 16, kNoPixel if invalid
    uint32_t parent[MAXPIX];  // union-find; later the cluster id
    uint32_t work[MAXPIX > 2 * kClusCap ? MAXPIX : 2 * kClusCap];  // column-sorted pixels; later charges/renumbering
    // column starts [0, kCols + 2) and fill cursors [kCols + 2, 2 kCols + 2); once the
    // components are found, the cluster extents of a run of at most 256 digis
    uint32_t cols[2 * kCols + 2];
    uint32_t wsum[NT / 64];
  };

  // one module run with a whole workgroup of NT threads (a run of up to MAXPIX digis;
  // the largest runs, up to 4000 digis, on 256 threads)
  template <int NT, uint32_t MAXPIX>
  __device__ void clusterRunWG(const Pool& p, uint32_t e, uint32_t k, RunLds<NT, MAXPIX>& L) {
    constexpr uint32_t kClusCap = RunLds<NT, MAXPIX>::kClusCap;
    uint32_t* sxy = L.sxy;
    uint32_t* parent = L.parent;
    uint32_t* work = L.work;
    uint32_t* colStart = L.cols;
    uint32_t* wsum = L.wsum;
    const EventDesc& ev = p.events[e];
    EventCounts& cnt = p.counts[e];
    const uint32_t nModules = cnt.nModules;

    const uint32_t* moduleStart = p.moduleStart + size_t(e) * kMaxModuleRanges;
    const uint32_t wo = ev.wordOffset;
    const uint32_t first = moduleStart[1 + k];
    const uint32_t end = (k + 1 < nModules) ? moduleStart[2 + k] : ev.nWords;
    uint32_t n = end - first;
    if (n > kMaxPixInModule) {
      if (threadIdx.x == 0)
        atomicOr(&cnt.status, kStatusPixelOverflow);
      n = kMaxPixInModule;  // truncated as in the reference
    }
    n = n < MAXPIX ? n : MAXPIX;  // runs are given to a workgroup whose LDS holds them
    const uint16_t me = p.moduleInd[wo + first];
    if (me >= kNumModules)
      n = 0;  // a run always starts on a valid digi; never taken for valid conditions

    // ---- stage pixels, column counts
    for (int i = threadIdx.x; i < kCols + 2; i += NT)
      colStart[i] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
      const uint32_t g = wo + first + i;
      const bool valid = p.moduleInd[g] != kInvId;
      uint32_t xy = kNoPixel;
      const uint32_t x = p.xx[g], y = p.yy[g];
      if (valid && x < kRowsPerModule && y < uint32_t(kCols)) {
        xy = x | (y << 16);
        atomicAdd(&colStart[y + 1], 1u);
      }
      sxy[i] = xy;
      parent[i] = i;
    }
    __syncthreads();
    // prefix over the 416 column counts (one wave, 7 columns per lane)
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
      uint32_t v[7], s = 0;
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        const int c = 1 + lane * 7 + j;
        v[j] = c <= kCols ? colStart[c] : 0;
        s += v[j];
      }
      uint32_t run = waveExclusiveScan(s);
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        const int c = 1 + lane * 7 + j;
        run += v[j];
        if (c <= kCols)
          colStart[c] = run;  // colStart[c] = end of column c-1 = start of column c
      }
    }
    __syncthreads();
    // fill the column lists (order inside a column is irrelevant)
    uint32_t* colFill = L.cols + kCols + 2;
    for (int i = threadIdx.x; i < kCols; i += NT)
      colFill[i] = colStart[i];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
      const uint32_t xy = sxy[i];
      if (xy != kNoPixel)
        work[atomicAdd(&colFill[xy >> 16], 1u)] = i;
    }
    __syncthreads();

    // ---- union over 8-neighbours: same column (pairs once) and next column
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
      const uint32_t xy = sxy[i];
      if (xy == kNoPixel)
        continue;
      const int x = int(xy & 0xffff), y = int(xy >> 16);
      for (uint32_t q = colStart[y]; q < colStart[y + 1]; ++q) {
        const uint32_t j = work[q];
        if (j <= i)
          continue;
        const int dx = int(sxy[j] & 0xffff) - x;
        if (dx >= -1 && dx <= 1)
          unite(parent, i, j);
      }
      if (y + 1 < kCols) {
        for (uint32_t q = colStart[y + 1]; q < colStart[y + 2]; ++q) {
          const uint32_t j = work[q];
          const int dx = int(sxy[j] & 0xffff) - x;
          if (dx >= -1 && dx <= 1)
            unite(parent, i, j);
        }
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += NT)
      if (sxy[i] != kNoPixel)
        parent[i] = findRoot(parent, i);
    __syncthreads();

    // ---- number the roots in increasing pixel index
    uint32_t nclus = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += NT) {
      const uint32_t i = c0 + threadIdx.x;
      const bool root = i < n && sxy[i] != kNoPixel && parent[i] == i;
      uint32_t tot;
      const uint32_t excl = wgExclusiveScan<NT>(root ? 1u : 0u, tot, wsum);
      if (root)
        work[i] = nclus + excl;  // rank of the root
      nclus += tot;
    }
    __syncthreads();
    const uint32_t found = nclus;
    // cluster id per pixel (roots' ranks); ids >= 1024 are dropped as in the reference
    for (uint32_t i = threadIdx.x; i < n; i += NT)
      if (sxy[i] != kNoPixel)
        parent[i] = work[parent[i]];
    __syncthreads();
    uint32_t ncut = found;
    if (found > kMaxClustersPerModule) {
      ncut = kMaxClustersPerModule;
      if (threadIdx.x == 0)
        atomicOr(&cnt.status, kStatusClusterOverflow);
    }

    // ---- charge cut
    int32_t* charge = reinterpret_cast<int32_t*>(work);                        // [kClusCap]
    uint32_t* newId = work + kClusCap;                                         // [kClusCap]
    for (uint32_t c = threadIdx.x; c < ncut; c += NT)
      charge[c] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
      if (sxy[i] != kNoPixel && parent[i] < ncut)
        atomicAdd(&charge[parent[i]], int32_t(p.adc[wo + first + i]));
    }
    __syncthreads();
    const int cut = me < 96 ? 2000 : 4000;
    uint32_t kept = 0;
    for (uint32_t c0 = 0; c0 < ncut; c0 += NT) {
      const uint32_t c = c0 + threadIdx.x;
      const uint32_t ok = (c < ncut && charge[c] > cut) ? 1u : 0u;
      uint32_t tot;
      const uint32_t excl = wgExclusiveScan<NT>(ok, tot, wsum);
      if (c < ncut)
        newId[c] = ok ? (kept + excl) : 9999u;
      kept += tot;
    }
    __syncthreads();
    const bool renumber = ncut > 0 && kept != ncut;

    // ---- write back
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
      const uint32_t g = wo + first + i;
      if (sxy[i] == kNoPixel) {
        p.clus[g] = -9999;
        continue;
      }
      uint32_t cid = parent[i];
      if (cid >= ncut) {
        p.clus[g] = kInvId;
        p.moduleInd[g] = kInvId;
        continue;
      }
      if (renumber) {
        cid = newId[cid];
        if (cid == 9999u)
          p.moduleInd[g] = kInvId;
      }
      p.clus[g] = int32_t(cid);
    }
    uint32_t* clusInModule = p.clusInModule + size_t(e) * kMaxModuleRanges;
    if (threadIdx.x == 0 && me < kNumModules) {
      clusInModule[me] = renumber ? kept : found;
      p.moduleId[size_t(e) * kMaxModuleRanges + k] = me;
    }

    // Runs of at most 256 digis also record each cluster's extent and charges for the
    // rec-hit stage (ClusterExt, as the wavefront kernel does), accumulated with LDS
    // atomics in the column area, which holds the nine fields of 92 clusters at a time
    // (nearly always all of the run's).  Every cluster of such a run has pixels.
    if constexpr (MAXPIX <= 256) {
      if (p.clExt && n > 0) {
        const uint32_t nf = renumber ? kept : found;
        constexpr uint32_t kChunk = (2 * kCols + 2) / 9;
        uint32_t* ext = L.cols;  // field j of chunk cluster c at ext[j * kChunk + c]
        auto finalId = [&](uint32_t i) -> uint32_t {
          if (sxy[i] == kNoPixel || parent[i] >= ncut)
            return 0xffffffffu;
          const uint32_t cid = renumber ? newId[parent[i]] : parent[i];
          return cid == 9999u ? 0xffffffffu : cid;
        };
        for (uint32_t f0 = 0; f0 < nf; f0 += kChunk) {
          const uint32_t m = min(kChunk, nf - f0);
          __syncthreads();  // the area is free (column lists, or the previous chunk)
          for (uint32_t c = threadIdx.x; c < m; c += NT) {
            ext[0 * kChunk + c] = ext[2 * kChunk + c] = 0xffffffffu;
            for (uint32_t j : {1u, 3u, 4u, 5u, 6u, 7u, 8u})
              ext[j * kChunk + c] = 0;
          }
          __syncthreads();
          for (uint32_t i = threadIdx.x; i < n; i += NT) {
            const uint32_t f = finalId(i) - f0;
            if (f < m) {
              const uint32_t x = sxy[i] & 0xffff, y = sxy[i] >> 16;
              atomicMin(&ext[0 * kChunk + f], x);
              atomicMax(&ext[1 * kChunk + f], x);
              atomicMin(&ext[2 * kChunk + f], y);
              atomicMax(&ext[3 * kChunk + f], y);
            }
          }
          __syncthreads();
          for (uint32_t i = threadIdx.x; i < n; i += NT) {
            const uint32_t f = finalId(i) - f0;
            if (f < m) {
              const uint32_t x = sxy[i] & 0xffff, y = sxy[i] >> 16;
              const uint32_t q = p.adc[wo + first + i];
              atomicAdd(&ext[4 * kChunk + f], q);
              if (ext[0 * kChunk + f] == x)
                atomicAdd(&ext[5 * kChunk + f], q);
              if (ext[1 * kChunk + f] == x)
                atomicAdd(&ext[6 * kChunk + f], q);
              if (ext[2 * kChunk + f] == y)
                atomicAdd(&ext[7 * kChunk + f], q);
              if (ext[3 * kChunk + f] == y)
                atomicAdd(&ext[8 * kChunk + f], q);
            }
          }
          __syncthreads();
          for (uint32_t c = threadIdx.x; c < m; c += NT) {
            ClusterExt r;
            r.minRow = uint16_t(ext[0 * kChunk + c]);
            r.maxRow = uint16_t(ext[1 * kChunk + c]);
            r.minCol = uint16_t(ext[2 * kChunk + c]);
            r.maxCol = uint16_t(ext[3 * kChunk + c]);
            r.qsum = int32_t(ext[4 * kChunk + c]);
            r.qfx = int32_t(ext[5 * kChunk + c]);
            r.qlx = int32_t(ext[6 * kChunk + c]);
            r.qfy = int32_t(ext[7 * kChunk + c]);
            r.qly = int32_t(ext[8 * kChunk + c]);
            r.pad = 0;
            p.clExt[wo + first + f0 + c] = r;
          }
        }
      }
    }
  }

  namespace {
    __device__ __forceinline__ void waveSync() {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __device__ __forceinline__ uint64_t lanesBelow(uint32_t lane) { return (1ull << lane) - 1ull; }
  }  // namespace

  // ---- module runs of at most 64 digis ---------------------------------------------
  // One run on the lanes [base, base + G) of the wavefront (G = 64, or G = 32 with the
  // two halves on two runs at once).  Every lane of the wavefront calls this together:
  // the propagation loop ends on a wavefront-wide ballot.  `live`: this group has a run.
  namespace {
    // per-wavefront LDS of the cluster extents: 64 clusters of each of the 9 fields
    enum : int { kMinRow, kMaxRow, kMinCol, kMaxCol, kQsum, kQfx, kQlx, kQfy, kQly, kExtFields };

    template <int G>
    __device__ void clusterRunLanes(const Pool& p, uint32_t e, uint32_t k, uint32_t first, uint32_t n,
                                    const EventDesc& ev, int32_t* rowWave, uint32_t (*extWave)[64], uint32_t base,
                                    bool live) {
      const uint32_t lane = (threadIdx.x & 63) - base;
      constexpr uint64_t kGroupMask = G == 64 ? ~0ull : (1ull << G) - 1;
      auto ballot = [&](bool v) { return (__ballot(v) >> base) & kGroupMask; };
      int32_t* row = rowWave + base;
      const uint32_t wo = ev.wordOffset;
      const uint16_t me = live ? p.moduleInd[wo + first] : kInvId;
      // a run always starts on a valid digi; never false for valid conditions
      live = live && me < kNumModules;
      const bool act = live && lane < n;
      const uint32_t g = wo + first + lane;
      uint32_t x = 0, y = 0;
      int32_t adc = 0;
      bool valid = false;
      if (act) {
        x = p.xx[g];
        y = p.yy[g];
        adc = p.adc[g];
        valid = p.moduleInd[g] != kInvId && x < kRowsPerModule && y < kColsPerModule;
      }
      // 8-neighbourhood as a lane mask (duplicated pixels are neighbours too); a
      // valid pixel is x | y << 16 with x < 160, y < 416, an invalid one ~0
      uint64_t nb = 0;
      const uint32_t xy = valid ? x | y << 16 : 0xffffffffu;
      const uint32_t nMax = live ? n : 0;
      const uint32_t nLoop = max(nMax, uint32_t(__shfl_xor(int(nMax), G == 64 ? 0 : 32)));  // both halves' runs
      for (uint32_t j = 0; j < nLoop; ++j) {
        const uint32_t xyj = __shfl(xy, int(base + j));
        const int dx = int(xyj & 0xffff) - int(x), dy = int(xyj >> 16) - int(y);
        if (j < n && valid && xyj != 0xffffffffu && j != lane && dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1)
          nb |= 1ull << j;
      }
      // min-label propagation with pointer jumping; every label is a lane of the same component
      uint32_t label = lane;
      while (true) {
        row[lane] = int32_t(label);
        waveSync();
        uint32_t m = label;
        for (uint64_t b = nb; b; b &= b - 1) {
          const uint32_t l = uint32_t(row[__ffsll((unsigned long long)b) - 1]);
          m = l < m ? l : m;
        }
        const uint32_t jump = uint32_t(row[m]);
        m = jump < m ? jump : m;
        waveSync();
        const bool changed = m != label;
        label = m;
        if (__ballot(changed) == 0)
          break;
      }
      // roots (label == own index) numbered in index order
      const uint64_t roots = ballot(valid && label == lane);
      const uint32_t nclus = __popcll(roots);
      const uint32_t cid = __popcll(roots & lanesBelow(label));
      // charge cut
      row[lane] = 0;
      waveSync();
      if (valid)
        atomicAdd(&row[cid], adc);
      waveSync();
      const bool isRoot = (roots >> lane) & 1;
      const int cut = me < 96 ? 2000 : 4000;
      const uint64_t okRoots = ballot(isRoot && row[__popcll(roots & lanesBelow(lane))] > cut);
      const uint32_t kept = __popcll(okRoots);
      const bool renumber = nclus > 0 && kept != nclus;
      if (act) {
        if (!valid) {
          p.clus[g] = -9999;
        } else if (!renumber) {
          p.clus[g] = int32_t(cid);
        } else if ((okRoots >> label) & 1) {
          p.clus[g] = int32_t(__popcll(okRoots & lanesBelow(label)));
        } else {
          p.clus[g] = 9999;
          p.moduleInd[g] = kInvId;
        }
      }
      if (live && lane == 0) {
        p.clusInModule[size_t(e) * kMaxModuleRanges + me] = renumber ? kept : nclus;
        p.moduleId[size_t(e) * kMaxModuleRanges + k] = me;
      }
      // Each surviving cluster's extent and charges for the rec-hit stage (the pool of
      // the no-allocation entry point has no room for them, and no rec-hit stage):
      // LDS atomics over the pixels, exact in any order, then one lane per cluster
      // writes the record at the run's first digi + cluster id.
      if (p.clExt) {
        const uint32_t nf = renumber ? kept : nclus;
        int32_t fid = -1;  // the pixel's final cluster id
        if (act && valid) {
          if (!renumber)
            fid = int32_t(cid);
          else if ((okRoots >> label) & 1)
            fid = int32_t(__popcll(okRoots & lanesBelow(label)));
        }
        const uint32_t slot = base + lane;
        if (live && lane < nf) {
          extWave[kMinRow][slot] = extWave[kMinCol][slot] = 0xffffffffu;
          extWave[kMaxRow][slot] = extWave[kMaxCol][slot] = 0;
          extWave[kQsum][slot] = extWave[kQfx][slot] = extWave[kQlx][slot] = 0;
          extWave[kQfy][slot] = extWave[kQly][slot] = 0;
        }
        waveSync();
        const uint32_t f = base + uint32_t(fid);
        if (fid >= 0) {
          atomicMin(&extWave[kMinRow][f], x);
          atomicMax(&extWave[kMaxRow][f], x);
          atomicMin(&extWave[kMinCol][f], y);
          atomicMax(&extWave[kMaxCol][f], y);
        }
        waveSync();
        if (fid >= 0) {
          const uint32_t q = uint32_t(adc);
          atomicAdd(&extWave[kQsum][f], q);
          if (extWave[kMinRow][f] == x)
            atomicAdd(&extWave[kQfx][f], q);
          if (extWave[kMaxRow][f] == x)
            atomicAdd(&extWave[kQlx][f], q);
          if (extWave[kMinCol][f] == y)
            atomicAdd(&extWave[kQfy][f], q);
          if (extWave[kMaxCol][f] == y)
            atomicAdd(&extWave[kQly][f], q);
        }
        waveSync();
        if (live && lane < nf) {
          ClusterExt c;
          c.minRow = uint16_t(extWave[kMinRow][slot]);
          c.maxRow = uint16_t(extWave[kMaxRow][slot]);
          c.minCol = uint16_t(extWave[kMinCol][slot]);
          c.maxCol = uint16_t(extWave[kMaxCol][slot]);
          c.qsum = int32_t(extWave[kQsum][slot]);
          c.qfx = int32_t(extWave[kQfx][slot]);
          c.qlx = int32_t(extWave[kQlx][slot]);
          c.qfy = int32_t(extWave[kQfy][slot]);
          c.qly = int32_t(extWave[kQly][slot]);
          c.pad = 0;
          p.clExt[wo + first + lane] = c;
        }
      }
      waveSync();  // the row and the extents are reused by the next run of the wavefront
    }
  }  // namespace

  // kRuns (2) consecutive module runs per wavefront: on its two halves when both hold
  // at most 32 digis (most at PU50), else one after the other on the whole wavefront
  // (one run per wavefront for a few events, to fill the device); runs of more than
  // 64 digis go to the work list of the workgroup kernel
  template <uint32_t kRuns>
  __global__ void __launch_bounds__(kBlock) clusterSmallKernel(Pool p) {
    __shared__ int32_t rowLds[kWaves][64];
    __shared__ uint32_t extLds[kWaves][kExtFields][64];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const LogicalBlock lb = logicalBlock();
    const uint32_t k0 = kRuns * (lb.x * kWaves + wave);
    const uint32_t e = lb.y;
    const EventDesc& ev = p.events[e];
    const uint32_t nModules = p.counts[e].nModules;
    if (k0 >= nModules)
      return;  // uniform per wavefront; this kernel has no workgroup barrier
    const uint32_t* ms = p.moduleStart + size_t(e) * kMaxModuleRanges;
    auto runOf = [&](uint32_t k, uint32_t& first) -> uint32_t {
      if (k >= nModules)
        return 0;
      first = ms[1 + k];
      const uint32_t end = (k + 1 < nModules) ? ms[2 + k] : ev.nWords;
      return end - first;
    };
    uint32_t f0 = 0, f1 = 0;
    const uint32_t n0 = runOf(k0, f0), n1 = kRuns > 1 ? runOf(k0 + 1, f1) : 0;
    int32_t* row = rowLds[wave];
    uint32_t(*ext)[64] = extLds[wave];
    if (n0 <= 32 && n1 <= 32) {
      const uint32_t half = lane >> 5;
      clusterRunLanes<32>(p, e, k0 + half, half ? f1 : f0, half ? n1 : n0, ev, row, ext, 32 * half,
                          (half ? n1 : n0) > 0);
    } else {
      clusterRunLanes<64>(p, e, k0, f0, n0, ev, row, ext, 0, n0 > 0 && n0 <= 64);
      clusterRunLanes<64>(p, e, k0 + 1, f1, n1, ev, row, ext, 0, n1 > 0 && n1 <= 64);
    }
  }

  // ---- larger module runs: each event's lists, their runs strided over its workgroups ----
  // runs of 65 to 256 digis: one wavefront each (a workgroup of one wave, 7.4 KB of LDS)
  constexpr uint32_t kMediumMaxPix = 256;
  __global__ void __launch_bounds__(64) clusterMediumKernel(Pool p) {
    __shared__ RunLds<64, kMediumMaxPix> L;
    const uint32_t e = blockIdx.y;
    const uint32_t n = p.counts[e].nMediumRuns;
    const uint32_t* list = p.bigList + size_t(e) * kMaxModuleRanges;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {  // uniform per workgroup
      clusterRunWG<64, kMediumMaxPix>(p, e, list[i], L);
      __syncthreads();
    }
  }

  // runs of more than 256 digis (up to 4000), listed from the back of the event's list
  __global__ void __launch_bounds__(kBlock) clusterBigKernel(Pool p) {
    __shared__ RunLds<kBlock, kMaxPixInModule> L;
    const uint32_t e = blockIdx.y;
    const uint32_t n = p.counts[e].nBigRuns;
    const uint32_t* list = p.bigList + size_t(e) * kMaxModuleRanges;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {  // uniform per workgroup
      clusterRunWG<kBlock, kMaxPixInModule>(p, e, list[kMaxModuleRanges - 1 - i], L);
      __syncthreads();
    }
  }

  // ---- clamped prefix of the clusters per module -----------------------------------
  __global__ void __launch_bounds__(1024) clusStartKernel(Pool p) {
    __shared__ uint32_t wsum[16];
    const uint32_t e = blockIdx.x;
    const uint32_t* cim = p.clusInModule + size_t(e) * kMaxModuleRanges;
    uint32_t* cms = p.clusModuleStart + size_t(e) * kMaxModuleRanges;
    uint32_t run = 0;
    for (uint32_t c0 = 0; c0 < kMaxNumModules; c0 += 1024) {
      const uint32_t m = c0 + threadIdx.x;
      uint32_t v = m < kMaxNumModules ? cim[m] : 0;
      v = v < kMaxClustersPerModule ? v : kMaxClustersPerModule;
      uint32_t tot;
      const uint32_t excl = wgExclusiveScan<1024>(v, tot, wsum);
      if (m < kMaxNumModules) {
        const uint32_t sum = run + excl + v;
        cms[m + 1] = sum < kMaxNumClusters ? sum : kMaxNumClusters;
      }
      run += tot;
    }
    if (threadIdx.x == 0) {
      cms[0] = 0;
      EventCounts& cnt = p.counts[e];
      cnt.nClusters = run < kMaxNumClusters ? run : kMaxNumClusters;
      if (run > kMaxNumClusters)
        cnt.status |= kStatusHitOverflow;
    }
  }

  void launchClusterize(const Pool& p, uint32_t nEvents, hipStream_t s) {
    if (nEvents < 8)
      clusterSmallKernel<1><<<dim3((kMaxNumModules + kWaves - 1) / kWaves, nEvents), kBlock, 0, s>>>(p);
    else
      clusterSmallKernel<2><<<dim3((kMaxNumModules + 2 * kWaves - 1) / (2 * kWaves), nEvents), kBlock, 0, s>>>(p);
    const uint32_t med = nEvents ? std::max(1u, std::min(kMediumPerEventMax, kMediumWorkgroups / nEvents)) : 1u;
    clusterMediumKernel<<<dim3(med, nEvents), 64, 0, s>>>(p);
    clusterBigKernel<<<dim3(kBigPerEvent, nEvents), kBlock, 0, s>>>(p);
    clusStartKernel<<<nEvents, 1024, 0, s>>>(p);
  }

}  // namespace pxg

// Stage A, part 2: per-module clustering, cluster charge cut and the
// clusters-per-module prefix (one fused kernel).
//
// One workgroup per (module run, event).  The module's pixels are staged in
// LDS; a column histogram sorts them by column; every pixel unites with the
// pixels of its own and the next column within one row (8-connectivity)
// through a lock-free union-find in LDS that always hooks the larger root
// under the smaller one, so every component ends with its smallest pixel
// index as root.  Roots are then numbered in increasing pixel index by a
// workgroup scan, which is exactly the serial program's cluster numbering
// (gpuClustering.h:40-273: labels converge to the component's minimum index,
// roots are numbered in index order).  The charge cut (gpuClusterChargeCut.h:
// 14-114) runs on the same LDS state: integer LDS atomics for the charges, a
// scan for the renumbering.  The last workgroup of an event to finish builds
// clusModuleStart (SiPixelRawToClusterGPUKernel.cc:464-506).
#include "pxgpu_internal.h"

namespace pxg {

  namespace {
    constexpr int kBlock = 256;
    constexpr int kWaves = kBlock / 64;
    constexpr uint32_t kNoPixel = 0xffffffffu;
    constexpr int kCols = kColsPerModule;  // 416

    // exclusive workgroup scan of one 32-bit value per thread; returns the total
    __device__ __forceinline__ uint32_t scanExclusive(uint32_t v, uint32_t& excl, uint32_t* wsum) {
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      uint32_t incl = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(incl, o);
        if (lane >= o)
          incl += t;
      }
      if (lane == 63)
        wsum[wave] = incl;
      __syncthreads();
      uint32_t before = 0, total = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        uint32_t s = wsum[w];
        before += (w < wave) ? s : 0;
        total += s;
      }
      excl = before + incl - v;
      __syncthreads();
      return total;
    }

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

  __global__ void __launch_bounds__(kBlock) clusterKernel(Pool p) {
    const uint32_t k = blockIdx.x;  // module run index
    const uint32_t e = blockIdx.y;
    const EventDesc& ev = p.events[e];
    EventCounts& cnt = p.counts[e];
    const uint32_t nModules = cnt.nModules;
    if (k >= nModules)
      return;

    __shared__ uint32_t sxy[kMaxPixInModule];     // x | y << 16, kNoPixel if invalid
    __shared__ uint32_t parent[kMaxPixInModule];  // union-find; later the cluster id
    __shared__ uint32_t work[kMaxPixInModule];    // column-sorted pixels; later charges/renumbering
    __shared__ uint32_t colStart[kCols + 2];
    __shared__ uint32_t wsum[kWaves];
    __shared__ uint32_t shared1[4];

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
    const uint16_t me = p.moduleInd[wo + first];
    if (me >= kNumModules)
      n = 0;  // a run always starts on a valid digi; never taken for valid conditions

    // ---- stage pixels, column counts
    for (int i = threadIdx.x; i < kCols + 2; i += kBlock)
      colStart[i] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
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
      uint32_t incl = s;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(incl, o);
        if (lane >= o)
          incl += t;
      }
      uint32_t run = incl - s;
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
    __shared__ uint32_t colFill[kCols];
    for (int i = threadIdx.x; i < kCols; i += kBlock)
      colFill[i] = colStart[i];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
      const uint32_t xy = sxy[i];
      if (xy != kNoPixel)
        work[atomicAdd(&colFill[xy >> 16], 1u)] = i;
    }
    __syncthreads();

    // ---- union over 8-neighbours: same column (pairs once) and next column
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
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
    for (uint32_t i = threadIdx.x; i < n; i += kBlock)
      if (sxy[i] != kNoPixel)
        parent[i] = findRoot(parent, i);
    __syncthreads();

    // ---- number the roots in increasing pixel index
    uint32_t nclus = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += kBlock) {
      const uint32_t i = c0 + threadIdx.x;
      const bool root = i < n && sxy[i] != kNoPixel && parent[i] == i;
      uint32_t excl;
      const uint32_t tot = scanExclusive(root ? 1u : 0u, excl, wsum);
      if (root)
        work[i] = nclus + excl;  // rank of the root
      nclus += tot;
    }
    __syncthreads();
    const uint32_t found = nclus;
    // cluster id per pixel (roots' ranks); ids >= 1024 are dropped as in the reference
    for (uint32_t i = threadIdx.x; i < n; i += kBlock)
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
    int32_t* charge = reinterpret_cast<int32_t*>(work);                        // [1024]
    uint32_t* newId = work + kMaxClustersPerModule;                            // [1024]
    for (uint32_t c = threadIdx.x; c < ncut; c += kBlock)
      charge[c] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
      if (sxy[i] != kNoPixel && parent[i] < ncut)
        atomicAdd(&charge[parent[i]], int32_t(p.adc[wo + first + i]));
    }
    __syncthreads();
    const int cut = me < 96 ? 2000 : 4000;
    uint32_t kept = 0;
    for (uint32_t c0 = 0; c0 < ncut; c0 += kBlock) {
      const uint32_t c = c0 + threadIdx.x;
      const uint32_t ok = (c < ncut && charge[c] > cut) ? 1u : 0u;
      uint32_t excl;
      const uint32_t tot = scanExclusive(ok, excl, wsum);
      if (c < ncut)
        newId[c] = ok ? (kept + excl) : 9999u;
      kept += tot;
    }
    __syncthreads();
    const bool renumber = ncut > 0 && kept != ncut;

    // ---- write back
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
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

    // ---- the last workgroup of the event builds clusModuleStart
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t done = atomicAdd(&cnt.modulesDone, 1u);
      shared1[0] = (done == nModules - 1) ? 1u : 0u;
    }
    __syncthreads();
    if (shared1[0] == 0)
      return;
    __threadfence();
    uint32_t* cms = p.clusModuleStart + size_t(e) * kMaxModuleRanges;
    uint32_t run = 0;
    for (uint32_t c0 = 0; c0 < kMaxNumModules; c0 += kBlock) {
      const uint32_t m = c0 + threadIdx.x;
      uint32_t v = 0;
      if (m < kMaxNumModules) {
        v = __hip_atomic_load(&clusInModule[m], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v = v < kMaxClustersPerModule ? v : kMaxClustersPerModule;
      }
      uint32_t excl;
      const uint32_t tot = scanExclusive(v, excl, wsum);
      if (m < kMaxNumModules) {
        const uint32_t s = run + excl + v;
        cms[m + 1] = s < kMaxNumClusters ? s : kMaxNumClusters;
      }
      run += tot;
    }
    if (threadIdx.x == 0) {
      cms[0] = 0;
      cnt.nClusters = run < kMaxNumClusters ? run : kMaxNumClusters;
      if (run > kMaxNumClusters)
        atomicOr(&cnt.status, kStatusHitOverflow);
    }
  }

  void launchClusterize(const Pool& p, uint32_t nEvents, hipStream_t s) {
    clusterKernel<<<dim3(kMaxNumModules, nEvents), kBlock, 0, s>>>(p);
  }

}  // namespace pxg

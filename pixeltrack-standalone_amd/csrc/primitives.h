// Building blocks shared by the stage kernels (the reference's CUDACore
// primitives, restated for 64-wide wavefronts): prefix scans over a wavefront
// and over a workgroup (prefixScan.h:13-26 blockPrefixScan), the one-workgroup
// array scan that finalises every count -> scan -> fill container, the
// one-to-many container itself (OneToManyAssoc / HistoContainer,
// HistoContainer.h:103-262, with a capacity like the reference's fixed
// content size), and per-wavefront counter allocation and sums (one atomic per
// wavefront where AtomicPairCounter.h:33-45 and the containers take one per
// thread).
//
// All scans are exact integer scans: the results do not depend on the
// workgroup size or on the order of the shuffles.  tests/native holds the
// self-test kernels that check each of them against a sequential scan.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pxg {

  __device__ __forceinline__ uint32_t laneId() { return threadIdx.x & 63; }
  __device__ __forceinline__ uint64_t lanesBelow() { return (1ull << laneId()) - 1; }

  // inclusive scan over the 64 lanes of a wavefront (all lanes active)
  template <typename T>
  __device__ __forceinline__ T waveInclusiveScan(T v) {
    const uint32_t lane = laneId();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const T t = __shfl_up(v, o);
      if (lane >= uint32_t(o))
        v += t;
    }
    return v;
  }

  template <typename T>
  __device__ __forceinline__ T waveExclusiveScan(T v) {
    return waveInclusiveScan(v) - v;
  }

  template <typename T>
  __device__ __forceinline__ T waveMax(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const T t = __shfl_xor(v, o);
      v = t > v ? t : v;
    }
    return v;
  }

  template <typename T>
  __device__ __forceinline__ T waveSum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      v += __shfl_xor(v, o);
    return v;
  }

  // Exclusive scan over a workgroup of NT threads (every thread calls it); returns the
  // thread's exclusive prefix and sets the workgroup total.  lds: NT / 64 elements.
  // Ends with a barrier, so consecutive calls may reuse lds.
  template <int NT, typename T>
  __device__ __forceinline__ T wgExclusiveScan(T v, T& total, T* lds) {
    static_assert(NT % 64 == 0 && NT <= 4096, "workgroup size");
    constexpr int kWaves = NT / 64;
    const uint32_t lane = laneId(), wid = threadIdx.x >> 6;
    const T incl = waveInclusiveScan(v);
    if (lane == 63)
      lds[wid] = incl;
    __syncthreads();
    if (wid == 0) {
      T w = lane < uint32_t(kWaves) ? lds[lane] : T(0);
#pragma unroll
      for (int o = 1; o < kWaves; o <<= 1) {
        const T t = __shfl_up(w, o);
        if (lane >= uint32_t(o))
          w += t;
      }
      if (lane < uint32_t(kWaves))
        lds[lane] = w;
    }
    __syncthreads();
    const T before = wid ? lds[wid - 1] : T(0);
    total = lds[kWaves - 1];
    __syncthreads();
    return before + incl - v;
  }

  // In-place exclusive scan of a[0, n) by one workgroup of NT threads; a[n] = total.
  // Four consecutive elements per thread and pass, so that a pass takes one workgroup
  // scan (three barriers) per 4 NT elements.  lds: NT / 64 words.
  template <int NT>
  __device__ uint32_t wgScanArray(uint32_t* a, uint32_t n, uint32_t* lds) {
    constexpr uint32_t kV = 4;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += NT * kV) {
      const uint32_t i0 = base + threadIdx.x * kV;
      uint32_t v[kV], sum = 0;
#pragma unroll
      for (uint32_t u = 0; u < kV; ++u) {
        v[u] = i0 + u < n ? a[i0 + u] : 0;
        sum += v[u];
      }
      uint32_t tot;
      uint32_t ex = carry + wgExclusiveScan<NT>(sum, tot, lds);
#pragma unroll
      for (uint32_t u = 0; u < kV; ++u) {
        if (i0 + u < n)
          a[i0 + u] = ex;
        ex += v[u];
      }
      carry += tot;
    }
    if (threadIdx.x == 0)
      a[n] = carry;
    return carry;
  }

  // Counter allocation with one atomic per wavefront: the active lanes asking for a
  // slot get consecutive values in lane order.  Lanes that diverged away do not take
  // part, so it may be called in divergent code.
  __device__ __forceinline__ uint32_t waveAlloc(uint32_t* counter, bool need) {
    const uint64_t m = __ballot(need);
    if (!m)
      return 0;
    const uint32_t leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (laneId() == leader)
      base = atomicAdd(counter, uint32_t(__popcll(m)));
    base = __shfl(base, leader);
    return base + __popcll(m & lanesBelow());
  }

  // sum over the active lanes, added to *counter once per wavefront
  __device__ __forceinline__ void waveAdd(uint32_t* counter, uint32_t v) {
    const uint64_t m = __ballot(v != 0);
    if (!m)
      return;
    uint32_t sum = v;
    for (int o = 32; o > 0; o >>= 1)
      sum += __shfl_xor(sum, o);
    if (laneId() == __ffsll((unsigned long long)m) - 1)
      atomicAdd(counter, sum);
  }

  // One-to-many container over nKeys keys with at most `capacity` entries
  // (OneToManyAssoc<T, nKeys, capacity>, HistoContainer.h:103-262), built in three
  // passes: count() per entry, scan() by one workgroup (offsets clamped to the
  // capacity, counts reset to serve as fill cursors), fill() per entry.  An entry
  // past the capacity is dropped and scan() reports the overflow by returning a
  // total above the capacity.  Within a key the entries are in atomic order.
  struct OneToMany {
    uint32_t* count;  // [nKeys]: counts, then fill cursors
    uint32_t* off;    // [nKeys + 1]: offsets into the content
    uint32_t nKeys, capacity;

    __device__ __forceinline__ void countEntry(uint32_t key) const { atomicAdd(&count[key], 1u); }

    // by one workgroup of NT threads (lds: NT / 64 words); returns the unclamped total
    template <int NT>
    __device__ uint32_t scan(uint32_t* lds) const {
      uint32_t carry = 0;
      for (uint32_t base = 0; base < nKeys; base += NT) {
        const uint32_t k = base + threadIdx.x;
        const uint32_t v = k < nKeys ? count[k] : 0;
        uint32_t tot;
        const uint32_t ex = wgExclusiveScan<NT>(v, tot, lds);
        if (k < nKeys) {
          off[k] = min(carry + ex, capacity);
          count[k] = 0;
        }
        carry += tot;
      }
      if (threadIdx.x == 0)
        off[nKeys] = min(carry, capacity);
      return carry;
    }

    // stores value under key; false when the key's range is full (capacity reached)
    template <typename V>
    __device__ __forceinline__ bool fill(uint32_t key, V value, V* content) const {
      const uint32_t pos = off[key] + atomicAdd(&count[key], 1u);
      if (pos >= off[key + 1])
        return false;
      content[pos] = value;
      return true;
    }
  };

}  // namespace pxg

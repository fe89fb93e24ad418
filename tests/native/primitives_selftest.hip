// Self-test kernels of the shared device primitives (csrc/primitives.h), for the
// GPU tests in tests/test_gpu_primitives.py: each entry point runs one primitive
// over device arrays the caller owns and waits for the result.  Test support
// only; the reconstruction library does not load it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../pixeltrack-standalone_amd/csrc/primitives.h"

using namespace pxg;

namespace {
  // one wavefront per 64 inputs: inclusive and exclusive scans, max and sum
  __global__ void waveKernel(const uint32_t* in, uint32_t* incl, uint32_t* excl, uint32_t* wmax, uint32_t* wsum,
                             uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t v = i < n ? in[i] : 0;
    const uint32_t a = waveInclusiveScan(v), b = waveExclusiveScan(v), m = waveMax(v), s = waveSum(v);
    if (i < n) {
      incl[i] = a;
      excl[i] = b;
    }
    if (laneId() == 0) {
      wmax[i >> 6] = m;
      wsum[i >> 6] = s;
    }
  }

  // one workgroup of NT threads per NT inputs: exclusive prefix within the workgroup and its total
  template <int NT, typename T>
  __global__ void __launch_bounds__(NT) wgKernel(const T* in, T* out, T* totals, uint32_t n) {
    __shared__ T lds[NT / 64];
    const uint32_t i = blockIdx.x * NT + threadIdx.x;
    T total;
    const T ex = wgExclusiveScan<NT>(i < n ? in[i] : T(0), total, lds);
    if (i < n)
      out[i] = ex;
    if (threadIdx.x == 0)
      totals[blockIdx.x] = total;
  }

  template <int NT>
  __global__ void __launch_bounds__(NT) scanArrayKernel(uint32_t* a, uint32_t n) {
    __shared__ uint32_t lds[NT / 64];
    wgScanArray<NT>(a, n, lds);
  }

  __global__ void allocKernel(uint32_t* counter, const uint8_t* need, uint32_t* slot, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
      return;  // divergent exit: the remaining lanes still allocate together
    if (need[i])
      slot[i] = waveAlloc(counter, true);
    else
      slot[i] = 0xffffffffu;
  }

  __global__ void addKernel(uint32_t* counter, const uint32_t* v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    waveAdd(counter, i < n ? v[i] : 0u);
  }

  __global__ void countKernel(OneToMany a, const uint32_t* keys, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
      a.countEntry(keys[i]);
  }

  __global__ void __launch_bounds__(1024) scanKernel(OneToMany a, uint32_t* total) {
    __shared__ uint32_t lds[1024 / 64];
    const uint32_t t = a.scan<1024>(lds);
    if (threadIdx.x == 0)
      *total = t;
  }

  __global__ void fillKernel(OneToMany a, const uint32_t* keys, uint32_t n, uint32_t* content, uint32_t* dropped) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
      if (!a.fill(keys[i], i, content))
        atomicAdd(dropped, 1u);
  }

  int finish() {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipDeviceSynchronize();
    return e == hipSuccess ? 0 : -int(e);
  }

  template <int NT, typename T>
  int wg(const void* in, void* out, void* totals, uint32_t n) {
    const uint32_t blocks = (n + NT - 1) / NT;
    if (blocks)
      wgKernel<NT, T><<<blocks, NT>>>(static_cast<const T*>(in), static_cast<T*>(out), static_cast<T*>(totals), n);
    return finish();
  }
}  // namespace

extern "C" {

int pxs_wave(const uint32_t* in, uint32_t* incl, uint32_t* excl, uint32_t* wmax, uint32_t* wsum, uint32_t n) {
  const uint32_t blocks = (n + 255) / 256;
  if (blocks)
    waveKernel<<<blocks, 256>>>(in, incl, excl, wmax, wsum, n);
  return finish();
}

// nt: 64, 256, 640 or 1024 threads; bits: 32 or 64
int pxs_wg_scan(int nt, int bits, const void* in, void* out, void* totals, uint32_t n) {
  if (bits == 32) {
    switch (nt) {
      case 64: return wg<64, uint32_t>(in, out, totals, n);
      case 256: return wg<256, uint32_t>(in, out, totals, n);
      case 640: return wg<640, uint32_t>(in, out, totals, n);
      case 1024: return wg<1024, uint32_t>(in, out, totals, n);
    }
  } else if (bits == 64) {
    switch (nt) {
      case 64: return wg<64, uint64_t>(in, out, totals, n);
      case 256: return wg<256, uint64_t>(in, out, totals, n);
      case 640: return wg<640, uint64_t>(in, out, totals, n);
      case 1024: return wg<1024, uint64_t>(in, out, totals, n);
    }
  }
  return -1;
}

int pxs_scan_array(int nt, uint32_t* a, uint32_t n) {
  if (nt == 256)
    scanArrayKernel<256><<<1, 256>>>(a, n);
  else if (nt == 1024)
    scanArrayKernel<1024><<<1, 1024>>>(a, n);
  else
    return -1;
  return finish();
}

int pxs_wave_alloc(uint32_t* counter, const uint8_t* need, uint32_t* slot, uint32_t n) {
  if (n)
    allocKernel<<<(n + 255) / 256, 256>>>(counter, need, slot, n);
  return finish();
}

int pxs_wave_add(uint32_t* counter, const uint32_t* v, uint32_t n) {
  if (n)
    addKernel<<<(n + 255) / 256, 256>>>(counter, v, n);
  return finish();
}

// keys[n] < nKeys; count[nKeys] must be zero; off[nKeys + 1]; content[capacity];
// total: the scan's unclamped total; dropped: entries past the capacity
int pxs_one_to_many(const uint32_t* keys, uint32_t n, uint32_t nKeys, uint32_t capacity, uint32_t* count,
                    uint32_t* off, uint32_t* content, uint32_t* total, uint32_t* dropped) {
  const OneToMany a{count, off, nKeys, capacity};
  if (n)
    countKernel<<<64, 256>>>(a, keys, n);
  scanKernel<<<1, 1024>>>(a, total);
  if (n)
    fillKernel<<<64, 256>>>(a, keys, n, content, dropped);
  return finish();
}

}  // extern "C"

// Raw words -> digis -> clusters of one event without allocating device memory:
// pxg_raw_to_cluster_noalloc.  It runs the kernels of pxg_raw_to_cluster (decode,
// module starts, clustering) on the caller's device arrays; what those kernels need
// beyond the inputs and outputs (the event descriptor and counters, the per-wave
// module-boundary and error masks, the error records at their word index, the
// clusteriser's list of large module runs) lives in `scratch`, a static device
// array each GPU holds once for the life of the process.
//
// Host code never passes `scratch` itself to a kernel or copy: its name is a
// host-side stand-in, not a device address.  noallocSetup (called by pxg_init
// with the context's device current) looks up each device's own address of the
// array with hipGetSymbolAddress, and every use goes through that address.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "pxgpu_stream.h"

namespace pxg {
  namespace {
    constexpr uint32_t kNoAllocWordCap = (PXG_NOALLOC_MAX_WORDS + kWordAlign - 1) / kWordAlign * kWordAlign;

    // byte layout of the scratch array, every part 256-byte aligned
    constexpr size_t align256(size_t n) { return (n + 255) / 256 * 256; }
    constexpr size_t kOffEvents = 0;
    constexpr size_t kOffCounts = kOffEvents + align256(sizeof(EventDesc));
    constexpr size_t kOffBigList = kOffCounts + align256(sizeof(EventCounts));
    constexpr size_t kOffMasks = kOffBigList + align256(kMaxModuleRanges * sizeof(uint32_t));
    constexpr size_t kOffSparse = kOffMasks + align256(2 * (kNoAllocWordCap / 64) * sizeof(uint64_t));
    constexpr size_t kScratchBytes = kOffSparse + align256(size_t(kNoAllocWordCap) * sizeof(PixelErrorCompact));
    static_assert(kScratchBytes == PXG_NOALLOC_SCRATCH_BYTES, "scratch size published in pxgpu.h");

    static __device__ __align__(256) uint8_t scratch[kScratchBytes];

    constexpr int kMaxDevices = 64;
    std::mutex g_scratchMutex;
    uint8_t* g_scratchPtr[kMaxDevices] = {};  // each device's address of `scratch`

    __global__ void setEventKernel(EventDesc* dst, EventDesc v) { *dst = v; }
  }  // namespace

  namespace detail {
    // Looks up the current device's address of the scratch array (once per device).
    int noallocSetup(int device) {
      if (device < 0 || device >= kMaxDevices)
        return fail(PXG_ERR_ARG, "device %d out of range", device);
      std::lock_guard<std::mutex> lock(g_scratchMutex);
      if (g_scratchPtr[device])
        return 0;
      void* ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !ptr)
        return fail(PXG_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
      g_scratchPtr[device] = static_cast<uint8_t*>(ptr);
      return 0;
    }
  }  // namespace detail
}  // namespace pxg

using namespace pxg;
using namespace pxg::detail;

extern "C" int pxg_raw_to_cluster_noalloc(pxg_context* ctx, void* hipStream, const uint32_t* words,
                                          const uint8_t* fedIds, uint32_t nWords, pxg_digis* d, pxg_clusters* c,
                                          pxg_errors* er) {
  if (!ctx || !d || !c || !er || (nWords && (!words || !fedIds)))
    return fail(PXG_ERR_ARG, "pxg_raw_to_cluster_noalloc: null argument");
  if (nWords % 2)
    return fail(PXG_ERR_ARG, "pxg_raw_to_cluster_noalloc: odd number of words %u", nWords);
  if (nWords > PXG_NOALLOC_MAX_WORDS)
    return fail(PXG_ERR_CAPACITY, "pxg_raw_to_cluster_noalloc: %u words > %u, the scratch array's capacity", nWords,
                PXG_NOALLOC_MAX_WORDS);
  if (!d->xx || !d->yy || !d->adc || !d->moduleInd || !d->clus || !d->pdigi || !d->rawIdArr || !c->moduleStart ||
      !c->clusInModule || !c->moduleId || !c->clusModuleStart || !er->data)
    return fail(PXG_ERR_ARG, "pxg_raw_to_cluster_noalloc: an output array is null");
  int dev = -1;
  PXG_HIP(hipGetDevice(&dev));
  if (dev != ctx->device)
    return fail(PXG_ERR_ARG, "pxg_raw_to_cluster_noalloc: device %d is current, the context is on %d", dev,
                ctx->device);
  uint8_t* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    base = dev < kMaxDevices ? g_scratchPtr[dev] : nullptr;
  }
  if (!base)
    return fail(PXG_ERR_ARG, "pxg_raw_to_cluster_noalloc: no scratch address for device %d (pxg_init not run)",
                dev);
  const hipStream_t s = static_cast<hipStream_t>(hipStream);

  // one event: the pool view puts the caller's arrays at offset 0 and the scratch parts
  // in place of the pool's own
  Pool p{};
  p.maxEvents = 1;
  p.maxWords = kNoAllocWordCap;
  p.words = const_cast<uint32_t*>(words);
  p.fedIds = const_cast<uint8_t*>(fedIds);
  p.events = reinterpret_cast<EventDesc*>(base + kOffEvents);
  p.counts = reinterpret_cast<EventCounts*>(base + kOffCounts);
  p.bigList = reinterpret_cast<uint32_t*>(base + kOffBigList);
  p.boundaryMask = reinterpret_cast<uint64_t*>(base + kOffMasks);
  p.errorsSparse = reinterpret_cast<PixelErrorCompact*>(base + kOffSparse);
  p.xx = d->xx;
  p.yy = d->yy;
  p.adc = d->adc;
  p.moduleInd = d->moduleInd;
  p.clus = d->clus;
  p.pdigi = d->pdigi;
  p.rawIdArr = d->rawIdArr;
  p.errors = static_cast<PixelErrorCompact*>(er->data);
  p.moduleStart = c->moduleStart;
  p.clusInModule = c->clusInModule;
  p.moduleId = c->moduleId;
  p.clusModuleStart = c->clusModuleStart;

  EventDesc ev{};
  ev.nWords = nWords;
  // the descriptor is written by a kernel (no host staging, no memset); the decode
  // and module-start kernels clear the counters and the work list themselves
  setEventKernel<<<1, 1, 0, s>>>(p.events, ev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(PXG_ERR_HIP, "launch of the event descriptor failed: %s", hipGetErrorString(e));
  launchDecode(ctx->cond, p, 1, nWords, s);
  launchModuleStarts(p, 1, s);
  launchClusterize(p, 1, s);
  e = hipGetLastError();
  if (e != hipSuccess)
    return fail(PXG_ERR_HIP, "pxg_raw_to_cluster_noalloc: launch failed: %s", hipGetErrorString(e));
  EventCounts cnt{};
  PXG_HIP(hipMemcpyAsync(&cnt, p.counts, sizeof(EventCounts), hipMemcpyDeviceToHost, s));
  PXG_HIP(hipStreamSynchronize(s));
  d->nModules = cnt.nModules;
  d->nDigis = nWords;
  c->nClusters = cnt.nClusters;
  er->size = cnt.nErrors;
  return int(cnt.status);
}

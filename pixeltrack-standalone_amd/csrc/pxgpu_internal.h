// Internal declarations of the gfx950 pixel reconstruction library: device
// conditions, per-stream device pools and kernel launchers.
//
// Batch layout.  Every stage kernel takes a table of per-event descriptors,
// so one launch covers one event (the plain per-event path) or a batch of
// events of different sizes (the batched path); both run the same kernels and
// give identical products.  Inside a batch the raw words of event e occupy
// [wordOffset[e], wordOffset[e] + nWords[e]) of the packed word/digi arrays,
// with wordOffset a multiple of 256 so that no workgroup or wavefront of the
// decoder straddles two events; per-event fixed-size arrays (module starts,
// cluster counts, hits, histograms) sit at e * capacity.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../common/pxformats.h"

namespace pxg {

  using namespace pxf;

  constexpr int kDecodeBlock = 256;         // words per decode workgroup
  constexpr uint32_t kWordAlign = 256;      // event word offsets in a batch
  constexpr uint32_t kMaxModuleRanges = kMaxNumModules + 1;

  // status bits reported per event (a set bit never silently truncates a product)
  enum : uint32_t {
    kStatusModuleOverflow = 1u << 0,   // more than 2000 module runs
    kStatusPixelOverflow = 1u << 1,    // a module with more than 4000 pixels (truncated like the reference)
    kStatusClusterOverflow = 1u << 2,  // a module with more than 1024 clusters
    kStatusHitOverflow = 1u << 3,      // more than 48k clusters in the event
  };

  // Conditions resident on the device, uploaded once (immutable afterwards).
  struct DevConditions {
    const SiPixelFedCablingMapGPU* cabling;
    const uint8_t* modToUnp;
    const SiPixelGainForHLTonGPU* gain;  // v_pedestals unused: see gainData
    const GainDecoding* gainData;
    uint32_t gainDataBytes;
    const CPECommonParams* cpeCommon;
    const CPEDetParams* cpeDet;
    const CPELayerGeometry* cpeLayer;
    const AverageGeometry* cpeAverage;
  };

  // Per-event descriptor (device copy written by the host before a launch).
  struct EventDesc {
    uint32_t wordOffset;   // first word / digi of the event in the batch arrays
    uint32_t nWords;
    uint32_t blockOffset;  // first decode workgroup of the event
    uint32_t pad;
    BeamSpotPOD bs;
  };

  // Per-event scalars written by the kernels and read back by the host.
  struct EventCounts {
    uint32_t nModules;
    uint32_t nClusters;
    uint32_t nErrors;
    uint32_t status;
    uint32_t modulesDone;  // completion counter of the clustering grid
    uint32_t pad[3];
  };

  // Device pointers of one stream's pool (batch capacity fixed at creation).
  struct Pool {
    uint32_t maxEvents, maxWords;
    // inputs
    uint32_t* words;
    uint8_t* fedIds;
    EventDesc* events;
    EventCounts* counts;
    // digis [maxWords]
    uint16_t *xx, *yy, *adc, *moduleInd;
    int32_t* clus;
    uint32_t *pdigi, *rawIdArr;
    uint64_t* boundaryMask;  // [maxWords / 64]
    PixelErrorCompact* errorsSparse;  // [maxWords], error records at their word index
    PixelErrorCompact* errors;        // event e: errors[wordOffset[e] ...], word order
    uint32_t* work;     // [0] large module runs listed, [1] next one to take
    uint32_t* bigList;  // large module runs of the launch: event << 16 | run
    // clusters, per event
    uint32_t *moduleStart, *clusInModule, *moduleId, *clusModuleStart;  // stride kMaxModuleRanges
    // hits, per event (stride kMaxNumClusters)
    float *xl, *yl, *xerr, *yerr, *xg, *yg, *zg, *rg;
    int16_t* iphi;
    int32_t* charge;
    int16_t *xsize, *ysize;
    uint16_t* detInd;
    uint32_t* hitsLayerStart;  // stride 16
    uint32_t* phiOff;          // stride kPhiHistTotBins (padded to 1288)
    uint16_t* phiBins;         // stride kMaxNumClusters
    AverageGeometry* avg;      // per event
  };

  constexpr uint32_t kLayerStartStride = 16;
  constexpr uint32_t kPhiOffStride = 1288;

  // kernel launchers (stage A, B); nBlocks = total decode workgroups of the batch
  void launchDecode(const DevConditions& c, const Pool& p, uint32_t nEvents, uint32_t nBlocks, uint32_t blockBase,
                    hipStream_t s);
  void launchModuleStarts(const Pool& p, uint32_t nEvents, hipStream_t s);
  void launchClusterize(const Pool& p, uint32_t nEvents, hipStream_t s);
  void launchHits(const DevConditions& c, const Pool& p, uint32_t nEvents, hipStream_t s);
  void launchPhiBinner(const DevConditions& c, const Pool& p, uint32_t nEvents, hipStream_t s);

}  // namespace pxg

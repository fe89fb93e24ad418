// ORACLE -- test infrastructure only (see oracle.h).
// C entry points used by the tests, the smoke check and the bench's CPU
// baseline: run the serial chain on one raw record and read its products by
// field name.
#include <atomic>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "oracle.h"

using namespace orc;

extern "C" {

struct orc_conditions {
  Conditions c;
};

struct orc_event {
  Event ev;
};

// dtype codes shared with the Python side: 1 u8, 2 u16, 3 i16, 4 u32, 5 i32, 6 f32, 7 f64
orc_conditions* orc_conditions_create(const void* cabling, const uint8_t* modToUnp, const void* gain,
                                      const uint8_t* gainData, const void* cpeCommon, const void* cpeDet,
                                      const void* cpeLayer, const void* cpeAverage, const void* beamSpot,
                                      const uint32_t* fedIds, uint32_t nFeds) {
  auto* h = new orc_conditions;
  h->c = {static_cast<const SiPixelFedCablingMapGPU*>(cabling),
          modToUnp,
          static_cast<const SiPixelGainForHLTonGPU*>(gain),
          gainData,
          static_cast<const CPECommonParams*>(cpeCommon),
          static_cast<const CPEDetParams*>(cpeDet),
          static_cast<const CPELayerGeometry*>(cpeLayer),
          static_cast<const AverageGeometry*>(cpeAverage),
          static_cast<const BeamSpotPOD*>(beamSpot),
          fedIds,
          nFeds};
  return h;
}

void orc_conditions_destroy(orc_conditions* h) { delete h; }

// upto: 1 digis, 2 clusters, 3 rec-hits
orc_event* orc_process(const orc_conditions* h, const uint8_t* raw, size_t len, int upto) {
  auto ev = std::make_unique<orc_event>();
  if (!process(h->c, raw, len, upto, ev->ev))
    return nullptr;
  return ev.release();
}

void orc_event_destroy(orc_event* e) { delete e; }

const void* orc_field(const orc_event* e, const char* name, size_t* count, int* dtype) {
  const Event& ev = e->ev;
  std::string n(name);
#define F(NAME, VEC, CODE)         \
  if (n == NAME) {                 \
    *count = ev.VEC.size();        \
    *dtype = CODE;                 \
    return ev.VEC.data();          \
  }
  F("words", words, 4)
  F("fedIds", fedIdx, 1)
  F("xx", xx, 2)
  F("yy", yy, 2)
  F("adc", adc, 2)
  F("moduleInd", moduleInd, 2)
  F("clus", clus, 5)
  F("pdigi", pdigi, 4)
  F("rawIdArr", rawIdArr, 4)
  F("moduleStart", moduleStart, 4)
  F("clusInModule", clusInModule, 4)
  F("moduleId", moduleId, 4)
  F("clusModuleStart", clusModuleStart, 4)
  F("xl", xl, 6)
  F("yl", yl, 6)
  F("xerr", xerr, 6)
  F("yerr", yerr, 6)
  F("xg", xg, 6)
  F("yg", yg, 6)
  F("zg", zg, 6)
  F("rg", rg, 6)
  F("iphi", iphi, 3)
  F("charge", charge, 5)
  F("xsize", xsize, 3)
  F("ysize", ysize, 3)
  F("detInd", detInd, 2)
  F("hitsLayerStart", hitsLayerStart, 4)
  F("phiOff", phiOff, 4)
  F("phiBins", phiBins, 2)
#undef F
  if (n == "errors") {  // 12-byte records
    *count = ev.errors.size() * sizeof(PixelErrorCompact);
    *dtype = 1;
    return ev.errors.data();
  }
  if (n == "avgCorrected") {
    *count = sizeof(AverageGeometry) / 4;
    *dtype = 6;
    return &ev.avgCorrected;
  }
  *count = 0;
  *dtype = 0;
  return nullptr;
}

// scalar counters: nWords, nModules, nDigis, nClusters, nHits, fedCounter
uint32_t orc_scalar(const orc_event* e, const char* name) {
  const Event& ev = e->ev;
  std::string n(name);
  if (n == "nWords")
    return ev.nWords;
  if (n == "nModules")
    return ev.nModules;
  if (n == "nDigis")
    return ev.nDigis;
  if (n == "nClusters")
    return ev.nClusters;
  if (n == "nHits")
    return ev.nHits;
  if (n == "fedCounter")
    return ev.fedCounter;
  return 0xffffffffu;
}

int16_t orc_atan2s7(float y, float x) { return atan2s7(y, x); }

// CPU baseline: runs the chain over nEvents raw records (concatenated, with
// offsets) with nThreads worker threads, each taking events from a shared
// counter, one event at a time (one "stream" per thread); returns the number
// of events processed.
uint32_t orc_run_many(const orc_conditions* h, const uint8_t* raws, const uint64_t* offsets, uint32_t nEvents,
                      int upto, int nThreads, uint32_t repeat) {
  std::atomic<uint32_t> next{0};
  const uint32_t total = nEvents * repeat;
  auto worker = [&]() {
    Event ev;
    for (;;) {
      uint32_t k = next.fetch_add(1);
      if (k >= total)
        break;
      uint32_t i = k % nEvents;
      process(h->c, raws + offsets[i], offsets[i + 1] - offsets[i], upto, ev);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < nThreads; ++t)
    pool.emplace_back(worker);
  for (auto& t : pool)
    t.join();
  return total;
}

}  // extern "C"

// ORACLE -- test infrastructure only.
//
// A CPU restatement of the reference serial pixel chain (src/serial of
// pixeltrack-standalone), written from the reference's algorithm and operation
// order, used as the parity checker of the GPU library.  Only tests/, the
// smoke check and the bench's cpu_baseline leg may load it.  Nothing in the
// product links or calls it.
//
// Pinning (see tests/test_oracle_*.py):
//   stage A (a1-a7) and stage B (a8-a10): bit-exact against the reference's own
//     stage A/B translation units compiled from /root/reference (oracle/ref/)
//     on seeded events, and against committed golden fixtures made from them;
//     plus brute-force connected components, closed-form clamped prefix sums,
//     atan2 accuracy bounds and frame orthogonality.
#pragma once

#include <cstdint>
#include <vector>

#include "../../pixeltrack-standalone_amd/common/pxformats.h"

namespace orc {

  using namespace pxf;

  struct Conditions {
    const SiPixelFedCablingMapGPU* cabling;
    const uint8_t* modToUnp;
    const SiPixelGainForHLTonGPU* gain;
    const uint8_t* gainData;
    const CPECommonParams* cpeCommon;
    const CPEDetParams* cpeDet;
    const CPELayerGeometry* cpeLayer;
    const AverageGeometry* cpeAverage;
    const BeamSpotPOD* beamSpot;
    const uint32_t* fedIds;
    uint32_t nFeds;
  };

  struct Event {
    // a1: host FED unpacking
    std::vector<uint32_t> words;
    std::vector<uint8_t> fedIdx;  // one per two words: fedId - 1200
    uint32_t nWords = 0, fedCounter = 0;
    // a2-a3: digis
    std::vector<uint16_t> xx, yy, adc, moduleInd;
    std::vector<int32_t> clus;
    std::vector<uint32_t> pdigi, rawIdArr;
    std::vector<PixelErrorCompact> errors;
    // a4-a7: clusters
    std::vector<uint32_t> moduleStart, clusInModule, moduleId, clusModuleStart;
    uint32_t nModules = 0, nDigis = 0, nClusters = 0;
    // a8-a10: rec-hits
    uint32_t nHits = 0;
    std::vector<float> xl, yl, xerr, yerr, xg, yg, zg, rg;
    std::vector<int16_t> iphi, xsize, ysize;
    std::vector<int32_t> charge;
    std::vector<uint16_t> detInd;
    std::vector<uint32_t> hitsLayerStart;
    std::vector<uint32_t> phiOff;     // 1281
    std::vector<uint16_t> phiBins;    // nHits
    AverageGeometry avgCorrected;
  };

  // a1 SiPixelRawToClusterCUDA.cc:87-138 on one raw record (raw.bin event layout)
  bool unpackFeds(const Conditions& c, const uint8_t* raw, size_t len, Event& ev);
  // a2 RawToDigi (also writes errors), a3 calibDigis
  void rawToDigi(const Conditions& c, Event& ev);
  void calibDigis(const Conditions& c, Event& ev);
  // a4 countModules, a5 findClus, a6 clusterChargeCut, a7 fillHitsModuleStart
  void countModules(Event& ev);
  void findClus(Event& ev);
  void clusterChargeCut(Event& ev);
  void fillHitsModuleStart(Event& ev);
  // a8 getHits, a9 setHitsLayerStart, a10 phi binner
  void getHits(const Conditions& c, Event& ev);
  void setHitsLayerStart(const Conditions& c, Event& ev);
  void fillPhiBinner(Event& ev);

  // runs the chain up to and including the given stage: 1 = digis, 2 = clusters, 3 = rec-hits
  bool process(const Conditions& c, const uint8_t* raw, size_t len, int upto, Event& ev);

  // exposed for unit tests
  int16_t atan2s7(float y, float x);

}  // namespace orc

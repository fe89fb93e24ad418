// ORACLE -- test infrastructure only.
//
// Driver of the reference's own stage A/B code, compiled from the reference
// sources where they lie (see oracle/Makefile; nothing of the reference is
// copied here).  It replays the FED loop of the reference plugin
// (src/serial/plugin-SiPixelClusterizer/SiPixelRawToClusterCUDA.cc:60-158,
// which itself cannot be compiled without TBB) with the reference
// ErrorChecker, then calls the reference's SiPixelRawToClusterGPUKernel::
// makeClusters and PixelRecHitGPUKernel::makeHits, and exposes the products
// by field name with the same C interface as the oracle restatement.
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "CUDADataFormats/SiPixelClustersSoA.h"
#include "CUDADataFormats/SiPixelDigiErrorsSoA.h"
#include "CUDADataFormats/SiPixelDigisSoA.h"
#include "CUDADataFormats/TrackingRecHit2DHeterogeneous.h"
#include "CondFormats/SiPixelFedCablingMapGPU.h"
#include "CondFormats/SiPixelGainForHLTonGPU.h"
#include "CondFormats/pixelCPEforGPU.h"
#include "DataFormats/BeamSpotPOD.h"
#include "DataFormats/FEDRawData.h"
#include "DataFormats/FEDRawDataCollection.h"
#include "DataFormats/approx_atan2.h"
#include "Geometry/phase1PixelTopology.h"
#include "plugin-SiPixelClusterizer/ErrorChecker.h"
#include "plugin-SiPixelClusterizer/SiPixelRawToClusterGPUKernel.h"
#include "plugin-SiPixelRecHits/PixelRecHits.h"

namespace {
  struct Cond {
    const SiPixelFedCablingMapGPU* cabling;
    const unsigned char* modToUnp;
    SiPixelGainForHLTonGPU gain;
    pixelCPEforGPU::ParamsOnGPU params;
    BeamSpotPOD bs;
    std::vector<unsigned int> fedIds;
  };

  struct Result {
    uint32_t nWords = 0, fedCounter = 0, nModules = 0, nDigis = 0, nClusters = 0, nHits = 0;
    std::vector<uint32_t> words;
    std::vector<uint8_t> fedIds;
    std::vector<uint16_t> xx, yy, adc, moduleInd;
    std::vector<int32_t> clus;
    std::vector<uint32_t> pdigi, rawIdArr;
    std::vector<uint8_t> errors;  // 12-byte records
    std::vector<uint32_t> moduleStart, clusInModule, moduleId, clusModuleStart;
    std::vector<float> xl, yl, xerr, yerr, xg, yg, zg, rg, avg;
    std::vector<int16_t> iphi, xsize, ysize;
    std::vector<int32_t> charge;
    std::vector<uint16_t> detInd, phiBins;
    std::vector<uint32_t> hitsLayerStart, phiOff;
  };

  template <typename T>
  std::vector<T> take(const T* p, size_t n) {
    return std::vector<T>(p, p + n);
  }
}  // namespace

extern "C" {

struct ref_conditions {
  Cond c;
};
struct ref_event {
  Result r;
};

ref_conditions* ref_conditions_create(const void* cabling, const uint8_t* modToUnp, const void* gain,
                                      const uint8_t* gainData, const void* cpeCommon, const void* cpeDet,
                                      const void* cpeLayer, const void* cpeAverage, const void* beamSpot,
                                      const uint32_t* fedIds, uint32_t nFeds) {
  auto* h = new ref_conditions;
  h->c.cabling = static_cast<const SiPixelFedCablingMapGPU*>(cabling);
  h->c.modToUnp = modToUnp;
  std::memcpy(static_cast<void*>(&h->c.gain), gain, sizeof(SiPixelGainForHLTonGPU));
  h->c.gain.v_pedestals =
      reinterpret_cast<SiPixelGainForHLTonGPU_DecodingStructure*>(const_cast<uint8_t*>(gainData));
  h->c.params.m_commonParams = static_cast<const pixelCPEforGPU::CommonParams*>(cpeCommon);
  h->c.params.m_detParams = static_cast<const pixelCPEforGPU::DetParams*>(cpeDet);
  h->c.params.m_layerGeometry = static_cast<const pixelCPEforGPU::LayerGeometry*>(cpeLayer);
  h->c.params.m_averageGeometry = static_cast<const pixelCPEforGPU::AverageGeometry*>(cpeAverage);
  std::memcpy(&h->c.bs, beamSpot, sizeof(BeamSpotPOD));
  h->c.fedIds.assign(fedIds, fedIds + nFeds);
  return h;
}

void ref_conditions_destroy(ref_conditions* h) { delete h; }

ref_event* ref_process(const ref_conditions* h, const uint8_t* raw, size_t len, int upto) {
  const Cond& c = h->c;
  // raw record -> FEDRawDataCollection (as bin/Source.cc:9-21)
  FEDRawDataCollection buffers;
  uint32_t nfeds;
  std::memcpy(&nfeds, raw, 4);
  size_t pos = 4;
  for (uint32_t i = 0; i < nfeds; ++i) {
    uint32_t fed, size;
    std::memcpy(&fed, raw + pos, 4);
    std::memcpy(&size, raw + pos + 4, 4);
    pos += 8;
    FEDRawData& d = buffers.FEDData(fed);
    d.resize(size);
    std::memcpy(d.data(), raw + pos, size);
    pos += size;
  }
  if (pos != len)
    return nullptr;

  // the FED loop of SiPixelRawToClusterCUDA::produce
  auto wordFed = std::make_unique<pixelgpudetails::SiPixelRawToClusterGPUKernel::WordFedAppender>();
  PixelFormatterErrors errors;
  unsigned int wordCounter = 0, fedCounter = 0;
  bool errorsInEvent = false;
  ErrorChecker checker;
  for (int fedId : c.fedIds) {
    if (fedId == 40)
      continue;
    fedCounter++;
    const FEDRawData& rawData = buffers.FEDData(fedId);
    int nWords = rawData.size() / sizeof(uint64_t);
    if (nWords == 0)
      continue;
    const uint64_t* trailer = reinterpret_cast<const uint64_t*>(rawData.data()) + (nWords - 1);
    if (not checker.checkCRC(errorsInEvent, fedId, trailer, errors))
      continue;
    const uint64_t* header = reinterpret_cast<const uint64_t*>(rawData.data());
    header--;
    bool moreHeaders = true;
    while (moreHeaders) {
      header++;
      moreHeaders = checker.checkHeader(errorsInEvent, fedId, header, errors);
    }
    bool moreTrailers = true;
    trailer++;
    while (moreTrailers) {
      trailer--;
      moreTrailers = checker.checkTrailer(errorsInEvent, fedId, nWords, trailer, errors);
    }
    const uint32_t* bw = (const uint32_t*)(header + 1);
    const uint32_t* ew = (const uint32_t*)(trailer);
    wordFed->initializeWordFed(fedId, wordCounter, bw, (ew - bw));
    wordCounter += (ew - bw);
  }

  auto ev = std::make_unique<ref_event>();
  Result& r = ev->r;
  r.nWords = wordCounter;
  r.fedCounter = fedCounter;
  r.words = take(wordFed->word(), wordCounter);
  r.fedIds = take(wordFed->fedId(), wordCounter / 2);

  pixelgpudetails::SiPixelRawToClusterGPUKernel algo;
  algo.makeClusters(true, c.cabling, c.modToUnp, &c.gain, *wordFed, std::move(errors), wordCounter, fedCounter, true,
                    true, false);
  auto products = algo.getResults();
  SiPixelDigisSoA digis = std::move(products.first);
  SiPixelClustersSoA clusters = std::move(products.second);
  SiPixelDigiErrorsSoA digiErrors = algo.getErrors();

  const uint32_t W = wordCounter;
  r.xx = take(digis.xx(), W);
  r.yy = take(digis.yy(), W);
  r.adc = take(digis.adc(), W);
  r.moduleInd = take(digis.moduleInd(), W);
  r.clus = take(digis.clus(), W);
  r.pdigi = take(digis.pdigi(), W);
  r.rawIdArr = take(digis.rawIdArr(), W);
  const auto* err = digiErrors.error();
  r.errors.resize(size_t(err->size()) * sizeof(PixelErrorCompact));
  if (err->size())
    std::memcpy(r.errors.data(), err->data(), r.errors.size());
  r.nModules = digis.nModules();
  r.nDigis = digis.nDigis();
  r.nClusters = clusters.nClusters();
  r.moduleStart = take(clusters.moduleStart(), gpuClustering::MaxNumModules + 1);
  r.clusInModule = take(clusters.clusInModule(), gpuClustering::MaxNumModules);
  r.moduleId = take(clusters.moduleId(), gpuClustering::MaxNumModules);
  r.clusModuleStart = take(clusters.clusModuleStart(), gpuClustering::MaxNumModules + 1);

  if (upto >= 3) {
    pixelgpudetails::PixelRecHitGPUKernel hitAlgo;
    TrackingRecHit2DCPU hits = hitAlgo.makeHits(digis, clusters, c.bs, &c.params);
    const uint32_t H = hits.nHits();
    r.nHits = H;
    const TrackingRecHit2DSOAView& v = *hits.view();
    for (uint32_t i = 0; i < H; ++i) {
      r.xl.push_back(v.xLocal(i));
      r.yl.push_back(v.yLocal(i));
      r.xerr.push_back(v.xerrLocal(i));
      r.yerr.push_back(v.yerrLocal(i));
      r.xg.push_back(v.xGlobal(i));
      r.yg.push_back(v.yGlobal(i));
      r.zg.push_back(v.zGlobal(i));
      r.rg.push_back(v.rGlobal(i));
      r.iphi.push_back(v.iphi(i));
      r.charge.push_back(v.charge(i));
      r.xsize.push_back(v.clusterSizeX(i));
      r.ysize.push_back(v.clusterSizeY(i));
      r.detInd.push_back(v.detectorIndex(i));
    }
    if (H) {
      r.hitsLayerStart = take(v.hitsLayerStart(), 11);
      const auto& hist = v.phiBinner();
      r.phiOff = take(hist.off, TrackingRecHit2DSOAView::Hist::totbins());
      r.phiBins = take(hist.bins, H);
    } else {
      r.hitsLayerStart.assign(11, 0);
      r.phiOff.assign(TrackingRecHit2DSOAView::Hist::totbins(), 0);
    }
    const auto& agc = v.averageGeometry();
    const float* a = reinterpret_cast<const float*>(&agc);
    r.avg.assign(a, a + sizeof(agc) / sizeof(float));
  }
  return ev.release();
}

void ref_event_destroy(ref_event* e) { delete e; }

const void* ref_field(const ref_event* e, const char* name, size_t* count, int* dtype) {
  const Result& r = e->r;
  std::string n(name);
#define F(NAME, VEC, CODE)  \
  if (n == NAME) {          \
    *count = r.VEC.size();  \
    *dtype = CODE;          \
    return r.VEC.data();    \
  }
  F("words", words, 4)
  F("fedIds", fedIds, 1)
  F("xx", xx, 2)
  F("yy", yy, 2)
  F("adc", adc, 2)
  F("moduleInd", moduleInd, 2)
  F("clus", clus, 5)
  F("pdigi", pdigi, 4)
  F("rawIdArr", rawIdArr, 4)
  F("errors", errors, 1)
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
  F("avgCorrected", avg, 6)
#undef F
  *count = 0;
  *dtype = 0;
  return nullptr;
}

uint32_t ref_scalar(const ref_event* e, const char* name) {
  const Result& r = e->r;
  std::string n(name);
  if (n == "nWords")
    return r.nWords;
  if (n == "nModules")
    return r.nModules;
  if (n == "nDigis")
    return r.nDigis;
  if (n == "nClusters")
    return r.nClusters;
  if (n == "nHits")
    return r.nHits;
  if (n == "fedCounter")
    return r.fedCounter;
  return 0xffffffffu;
}

// topology and arithmetic helpers of the reference, for direct unit checks
uint16_t ref_divu52(uint16_t n) { return phase1PixelTopology::divu52(n); }
int16_t ref_atan2s7(float y, float x) { return unsafe_atan2s<7>(y, x); }
uint16_t ref_localX(uint16_t px) { return phase1PixelTopology::localX(px); }
uint16_t ref_localY(uint16_t py) { return phase1PixelTopology::localY(py); }

}  // extern "C"

// ORACLE -- test infrastructure only (see oracle.h).
// C entry points used by the tests, the smoke check and the bench's CPU
// baseline: run the serial chain on one raw record and read its products by
// field name.
#include <atomic>
#include <map>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "oracle.h"

using namespace orc;

namespace orc {
  int g_mutation = kMutNone;
  int g_variant = 0;
}

extern "C" {

struct orc_conditions {
  Conditions c;
};

struct orc_event {
  Event ev;
  std::map<std::string, std::pair<std::vector<uint8_t>, int>> exported;  // materialised views
};

}  // extern "C"

namespace {
  template <typename T>
  void exportVec(orc_event* e, const std::string& name, const std::vector<T>& v, int code) {
    auto& slot = e->exported[name];
    slot.first.resize(v.size() * sizeof(T));
    if (!v.empty())
      std::memcpy(slot.first.data(), v.data(), slot.first.size());
    slot.second = code;
  }

  // stage C/D products as flat arrays
  void exportCA(orc_event* e) {
    const Event& ev = e->ev;
    const CAState& C = ev.ca;
    std::vector<uint16_t> inner, outer;
    std::vector<int32_t> did;
    std::vector<int16_t> lp;
    std::vector<uint16_t> used;
    std::vector<uint32_t> nbOff{0}, nbVal, trOff{0};
    std::vector<uint32_t> trVal;
    for (const Cell& c : C.cells) {
      inner.push_back(c.innerHit);
      outer.push_back(c.outerHit);
      did.push_back(c.doubletId);
      lp.push_back(c.layerPairId);
      used.push_back(c.used);
      const auto& nb = C.neighborStore[c.neighbors];
      for (int k = 0; k < nb.size; ++k)
        nbVal.push_back(nb.data[k]);
      nbOff.push_back(uint32_t(nbVal.size()));
      const auto& tr = C.trackStore[c.tracks];
      for (int k = 0; k < tr.size; ++k)
        trVal.push_back(tr.data[k]);
      trOff.push_back(uint32_t(trVal.size()));
    }
    exportVec(e, "cellInner", inner, 2);
    exportVec(e, "cellOuter", outer, 2);
    exportVec(e, "cellDoubletId", did, 5);
    exportVec(e, "cellLayerPair", lp, 3);
    exportVec(e, "cellUsed", used, 2);
    exportVec(e, "cellNeighborOff", nbOff, 4);
    exportVec(e, "cellNeighbors", nbVal, 4);
    exportVec(e, "cellTrackOff", trOff, 4);
    exportVec(e, "cellTracks", trVal, 4);
    std::vector<uint32_t> ohOff{0}, ohVal;
    for (uint32_t h = 0; h < ev.nHits && h < C.outerHitOfCell.size(); ++h) {
      const auto& v = C.outerHitOfCell[h];
      for (int k = 0; k < v.size; ++k)
        ohVal.push_back(v.data[k]);
      ohOff.push_back(uint32_t(ohVal.size()));
    }
    exportVec(e, "outerHitOff", ohOff, 4);
    exportVec(e, "outerHitCells", ohVal, 4);
    exportVec(e, "tupleOff", C.tupleOff, 4);
    std::vector<uint16_t> bins(C.tupleBins.begin(), C.tupleBins.begin() + std::min<size_t>(C.tupleBins.size(),
                                                                                              C.tupleOff.empty() ? 0 : C.tupleOff.back()));
    exportVec(e, "tupleBins", bins, 2);
    exportVec(e, "quality", C.quality, 1);
    std::vector<uint32_t> stats(C.stats.begin(), C.stats.end());
    exportVec(e, "stats", stats, 4);
    exportVec(e, "multOff", C.multOff, 4);
    exportVec(e, "multBins", C.multBins, 4);
    std::vector<uint16_t> det(C.detIndices.begin(), C.detIndices.begin() + bins.size());
    exportVec(e, "detIndices", det, 2);
    exportVec(e, "hitToTupleOff", C.hitToTupleOff, 4);
    exportVec(e, "hitToTupleBins", C.hitToTupleBins, 4);
    std::vector<float> st, cv;
    for (auto& a : C.state)
      st.insert(st.end(), a.begin(), a.end());
    for (auto& a : C.cov)
      cv.insert(cv.end(), a.begin(), a.end());
    exportVec(e, "state", st, 6);
    exportVec(e, "covariance", cv, 6);
    exportVec(e, "pt", C.pt, 6);
    exportVec(e, "eta", C.eta, 6);
    exportVec(e, "chi2", C.chi2, 6);
    const VertexState& V = ev.vtx;
    exportVec(e, "idv", V.idv, 3);
    exportVec(e, "zv", V.zv, 6);
    exportVec(e, "wv", V.wv, 6);
    exportVec(e, "chi2v", V.chi2, 6);
    exportVec(e, "ptv2", V.ptv2, 6);
    exportVec(e, "ndof", V.ndof, 5);
    exportVec(e, "sortInd", V.sortInd, 2);
  }
}  // namespace

extern "C" {

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

// fishbone passes of the chain (the reference runs the early one only)
int orc_conditions_set_fishbone(orc_conditions* h, int early, int late) {
  h->c.earlyFishbone = early != 0;
  h->c.lateFishbone = late != 0;
  return 0;
}

// helix fit of the chain: 0 broken line (the reference's default), 1 Riemann
int orc_conditions_set_riemann(orc_conditions* h, int riemann) {
  h->c.riemannFit = riemann != 0;
  return 0;
}

// field of the helix fits in 1/GeV (CAHitNtupletCUDA.cc:31) and the vertexer's track pt cut
// (PixelVertexProducerCUDA.cc:37)
int orc_conditions_set_bfield(orc_conditions* h, float bfield) {
  if (!(bfield > 0))
    return -1;
  h->c.bField = bfield;
  return 0;
}

int orc_conditions_set_vertex_ptmin(orc_conditions* h, float ptMin) {
  if (!(ptMin >= 0))
    return -1;
  h->c.vertexPtMin = ptMin;
  return 0;
}

// capacity profile of the ntuplet containers: 0 reference, 1 high occupancy
// vertex clustering algorithm of the chain: 0 density (the reference's default), 1 DBSCAN, 2 iterative
int orc_conditions_set_vertex_algo(orc_conditions* h, int algo) {
  if (algo < 0 || algo > 2)
    return -1;
  h->c.vertexAlgo = algo;
  return 0;
}

int orc_conditions_set_profile(orc_conditions* h, int profile) {
  if (profile == 0)
    h->c.caps = kCapsReference;
  else if (profile == 1)
    h->c.caps = kCapsHighOccupancy;
  else
    return -1;
  return 0;
}

// upto: 1 digis, 2 clusters, 3 rec-hits
orc_event* orc_process(const orc_conditions* h, const uint8_t* raw, size_t len, int upto) {
  auto ev = std::make_unique<orc_event>();
  if (!process(h->c, raw, len, upto, ev->ev))
    return nullptr;
  if (upto >= 4)
    exportCA(ev.get());
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
  auto it = e->exported.find(n);
  if (it != e->exported.end()) {
    static const size_t sizes[] = {0, 1, 2, 2, 4, 4, 4, 8};
    *dtype = it->second.second;
    *count = it->second.first.size() / sizes[*dtype];
    return it->second.first.data();
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
  if (n == "nCells")
    return ev.ca.nCells;
  if (n == "nTuples")
    return ev.ca.nTuples;
  if (n == "caStatus")
    return ev.ca.status;
  if (n == "nVertices")
    return ev.vtx.nvFinal;
  if (n == "vtxTracks")
    return ev.vtx.ntrks;
  if (n == "nVerticesFound")
    return ev.vtx.nFound;
  if (n == "nNeighborLists")
    return uint32_t(ev.ca.neighborStore.size());
  if (n == "nTrackLists")
    return uint32_t(ev.ca.trackStore.size());
  return 0xffffffffu;
}

int orc_fit_helix(int n, const double* hits, const float* ge, double B, float* state, float* cov, float* pt, float* eta,
                  float* chi2) {
  return fitHelix(n, hits, ge, B, state, cov, pt, eta, chi2);
}

void orc_invert_spd(const double* a, double* inv, int n) { invertSPDForTest(a, inv, n); }

// the broken-line fit's double-precision results (circle and line parameters and covariances,
// the two chi2) for the tests of its covariance and chi2
int orc_fit_helix_d(int n, const double* hits, const float* ge, double B, double* circle, double* ccov, double* line,
                    double* lcov, double* chi2c, double* chi2l) {
  return fitHelixD(n, hits, ge, B, circle, ccov, line, lcov, chi2c, chi2l);
}

// BrokenLine.h MultScatt: planar-angle variance of a segment of the given length
double orc_mult_scatt(double length, double B, double R, double slope) {
  return multScattForTest(length, B, R, slope);
}

// the track clustering of the vertex finder alone: iv (cluster per track, 9997 noise), nn
int orc_cluster_tracks(int algo, uint32_t nt, const float* zt, const float* ezt2, int32_t* iv, int32_t* nn) {
  if (algo < 0 || algo > 2)
    return -1;
  return int(clusterTracks(algo, nt, zt, ezt2, iv, nn));
}

// the broken-line circle fit's intermediate values (21 doubles: d, e, (phi, d, k) at the chord
// midpoint, their covariance, the scattering variance added at the first hit, S0, S1, R, slope)
int orc_circle_debug(int n, const double* hits, const float* ge, double B, const double* ff, double* out) {
  return circleDebug(n, hits, ge, B, ff, out);
}

// BrokenLine.h BL_Fast_fit: (X0, Y0, R, tan theta) of the hits
int orc_fast_fit(int n, const double* hits, double* ff) { return fastFitForTest(n, hits, ff); }

// BrokenLine.h TranslateKarimaki on (phi, d, k) and its 3x3 covariance, in place
void orc_translate_karimaki(double* par, double* cov, double x0, double y0) {
  translateKarimakiForTest(par, cov, x0, y0);
}

// planted mistakes for the tests that show a check catches them (oracle.h kMut*)
void orc_set_mutation(int m) { g_mutation = m; }

// departures from the reference's own expressions (oracle.h kVar* bits)
void orc_set_variant(int bits) { g_variant = bits; }

int orc_fit_helix_riemann(int n, const double* hits, const float* ge, double B, float* state, float* cov, float* pt,
                          float* eta, float* chi2) {
  return fitHelixRiemann(n, hits, ge, B, state, cov, pt, eta, chi2);
}

int orc_circle_fit_riemann(int n, const double* hits, const float* ge, double B, double* fast, double* par,
                           double* cov, int* q, float* chi2) {
  return circleFitRiemann(n, hits, ge, B, fast, par, cov, q, chi2);
}

double orc_rmath(int fn, double x, double y) { return rmathForTest(fn, x, y); }

void orc_eigen3(const double* a, int use_float, double* evals, double* evecs) {
  eigen3ForTest(a, use_float, evals, evecs);
}

// vertex finder on given track z, z error^2 and pt^2 (all tracks used)
int orc_vertices(uint32_t nt, const float* zt, const float* ezt2, const float* ptt2, int16_t* iv, float* zv, float* wv,
                 float* chi2, float* ptv2, int32_t* ndof, uint16_t* sortInd) {
  return verticesDirect(nt, zt, ezt2, ptt2, iv, zv, wv, chi2, ptv2, ndof, sortInd);
}

// the same with the clustering algorithm chosen: 0 density, 1 DBSCAN, 2 iterative
int orc_vertices_algo(int algo, uint32_t nt, const float* zt, const float* ezt2, const float* ptt2, int16_t* iv,
                      float* zv, float* wv, float* chi2, float* ptv2, int32_t* ndof, uint16_t* sortInd) {
  if (algo < 0 || algo > 2)
    return -1;
  return verticesDirect(nt, zt, ezt2, ptt2, iv, zv, wv, chi2, ptv2, ndof, sortInd, algo);
}

// circle through three points as the CA cuts see it: curvature and dca0 (distance of closest
// approach to the origin times the curvature)
void orc_circle(float x1, float y1, float x2, float y2, float x3, float y3, float* curvature, float* dca) {
  circleForTest(x1, y1, x2, y2, x3, y3, curvature, dca);
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

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
//   stage C/D (a11-a25): the reference's CA and vertexing sources need Eigen,
//     which is not available, so they cannot be compiled here; these parts are
//     pinned by brute force and invariants (doublets against an O(n^2) pair
//     search with the same cuts, tuples against an exhaustive walk of the cell
//     graph, fits against noiseless helices and matrix identities).  The fit's
//     numerical parity with the reference is unpinned.
#pragma once

#include <array>
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
    CACapacities caps = kCapsReference;
    int vertexAlgo = 0;  // 0 density, 1 DBSCAN, 2 iterative (gpuVertexFinder.h:46-60)
    bool earlyFishbone = true, lateFishbone = false;  // CAHitNtupletGeneratorOnGPU.cc:55-56
    bool riemannFit = false;                          // CAHitNtupletGeneratorOnGPU.cc:52 useRiemannFit
    float bField = 0.0114256972711507f;               // CAHitNtupletCUDA.cc:31, a float in makeTuples
    float vertexPtMin = 0.5f;                         // PixelVertexProducerCUDA.cc:37 m_ptMin
  };

  // ---- stage C/D fixed capacities (CAConstants.h, ZVertexSoA.h); the
  // run-time ones are in Event::caps (pxformats.h CACapacities)

  template <class T, int N>
  struct Vec {
    int size = 0;
    T data[N];
    int push(T v) {
      if (size < N) {
        data[size] = v;
        return size++;
      }
      return -1;
    }
  };

  // VecArray with a run-time capacity (OuterHitOfCell, CAConstants.h:61, whose
  // capacity maxCellsPerHit is a capacity of the profile here)
  struct CappedList {
    int size = 0;
    std::vector<uint32_t> data;
    int push(uint32_t v, uint32_t cap) {
      if (uint32_t(size) < cap) {
        data.push_back(v);
        return size++;
      }
      return -1;
    }
  };

  struct Cell {
    int32_t doubletId;
    int16_t layerPairId;
    uint16_t used;
    float innerZ, innerR;
    uint16_t innerHit, outerHit;
    int32_t neighbors;  // slot in the neighbour store (0: none)
    int32_t tracks;     // slot in the track store (0: none)
  };

  struct CAState {
    uint32_t status = 0;
    uint32_t nCells = 0;
    std::vector<Cell> cells;
    std::vector<CappedList> outerHitOfCell;
    std::vector<Vec<uint32_t, 36>> neighborStore;
    std::vector<Vec<TupleIndex, 48>> trackStore;
    uint32_t apcM = 0, apcN = 0, nTuples = 0;
    std::vector<uint32_t> tupleOff;   // [32769]
    std::vector<uint16_t> tupleBins;  // hit indices
    std::vector<uint8_t> quality;
    std::vector<uint32_t> multOff;  // [9]
    std::vector<TupleIndex> multBins;
    std::vector<uint16_t> detIndices;
    std::vector<uint32_t> hitToTupleOff;
    std::vector<TupleIndex> hitToTupleBins;
    std::vector<std::array<float, 5>> state;
    std::vector<std::array<float, 15>> cov;
    std::vector<float> pt, eta, chi2;
    // the event's contribution to the generator's counters (CAHitNtupletGeneratorKernels.h:14-26):
    // nEvents, nHits, nCells, nTuples, nFitTracks, nGoodTracks, nUsedHits, nDupHits,
    // nKilledCells, nEmptyCells, nZeroTrackCells
    std::array<uint32_t, 11> stats{};
  };

  struct VertexState {
    uint32_t status = 0;
    uint32_t ntrks = 0, nvFinal = 0;
    uint32_t nFound = 0;  // vertices found by the clustering, before splitVertices adds its own
    std::vector<int16_t> idv;
    std::vector<int32_t> ndof;
    std::vector<float> zv, wv, chi2, ptv2;
    std::vector<uint16_t> sortInd;
    std::vector<float> zt;
    std::vector<int32_t> iv;
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
    // a11-a22: CA ntuplets and fits; a23-a25: vertices
    CACapacities caps = kCapsReference;
    int vertexAlgo = 0;
    float vertexPtMin = 0.5f;
    bool earlyFishbone = true, lateFishbone = false;
    CAState ca;
    VertexState vtx;
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

  // a11-a17, a20-a22 (orc_stageC.cc)
  void buildDoublets(Event& ev);
  void connectCells(Event& ev);
  void fishbone(Event& ev, bool checkTrack = false);
  void fishboneCleaner(Event& ev);
  void findNtuplets(Event& ev);
  void earlyDuplicatesAndMultiplicity(Event& ev);
  void fillHitDetIndices(Event& ev);
  void classifyTracks(Event& ev);
  void fastDuplicateRemover(Event& ev);
  void hitToTupleAndTripletCleaner(Event& ev);
  void countStats(Event& ev);
  void circleForTest(float x1, float y1, float x2, float y2, float x3, float y3, float* curvature, float* dca);
  // a18-a19 (orc_fit.cc)
  void fitTracks(const Conditions& c, Event& ev);
  int fitHelix(int n, const double* hits, const float* ge, double B, float* state, float* cov, float* pt, float* eta,
               float* chi2);
  void invertSPDForTest(const double* a, double* inv, int n);
  int fitHelixD(int n, const double* hits, const float* ge, double B, double* circle, double* ccov, double* line,
                double* lcov, double* chi2c, double* chi2l);
  double multScattForTest(double length, double B, double R, double slope);
  int circleDebug(int n, const double* hits, const float* ge, double B, const double* ffIn, double* out);
  int fastFitForTest(int n, const double* hits, double* ff);
  void translateKarimakiForTest(double* par, double* cov, double x0, double y0);
  // a18/a19 alternative: the Riemann fit (orc_riemann.cc)
  void fitTracksRiemann(const Conditions& c, Event& ev);
  int fitHelixRiemann(int n, const double* hits, const float* ge, double B, float* state, float* cov, float* pt,
                      float* eta, float* chi2);
  int circleFitRiemann(int n, const double* hits, const float* ge, double B, double* fast, double* par, double* cov,
                       int* q, float* chi2);
  void eigen3ForTest(const double* a, int useFloat, double* evals, double* evecs);
  double rmathForTest(int fn, double x, double y);
  // a23-a25 (orc_stageD.cc)
  void findVertices(Event& ev);
  uint32_t clusterTracks(int algo, uint32_t nt, const float* zt, const float* ezt2, int32_t* iv, int32_t* nn);
  void clusterAndFit(VertexState& V, const std::vector<TupleIndex>& itrk, const std::vector<float>& zt,
                     const std::vector<float>& ezt2, const std::vector<float>& ptt2, int algo = 0);
  int verticesDirect(uint32_t nt, const float* zt, const float* ezt2, const float* ptt2, int16_t* iv, float* zv,
                     float* wv, float* chi2, float* ptv2, int32_t* ndof, uint16_t* sortInd, int algo = 0);

  // runs the chain up to and including the given stage: 1 = digis, 2 = clusters,
  // 3 = rec-hits, 4 = ntuplets + fits, 5 = vertices
  bool process(const Conditions& c, const uint8_t* raw, size_t len, int upto, Event& ev);

  // exposed for unit tests
  int16_t atan2s7(float y, float x);

  // Planted mistakes, for the tests that show a check catches them (0: none).  Each
  // changes one term of the restatement the way a plausible transcription error would.
  enum : int {
    kMutNone = 0,
    kMutXXI0 = 1,            // multiple scattering: material per cm doubled
    kMutChi2Ndf = 2,         // fit chi2 normalised by 2N-4 instead of 2N-5
    kMutVarBetaOneSide = 3,  // kink variance from the next segment only
    kMutNoFirstMSCircle = 4, // circle fit: first segment's scattering not added to phi
    kMutNoSlopeFactor = 5,   // circle fit: kink variances without the (1 + cot^2) factor
    kMutNoFirstMSLine = 6,   // line fit: first segment's scattering not added to the slope
    kMutLineWeight = 7,      // line fit: hit weight from the transverse instead of the z element
    kMutDensityTie = 8,      // density clustering: equal-density tie broken towards higher z
    kMutDensityNoChi2 = 9,   // density clustering: neighbour count without the chi2 cut
    kMutCuBorder = 10,       // circle fit: a term of the curvature border of C_u dropped
    kMutPhiJacobian = 11,    // circle fit: x and y of the first radius swapped in d(phi)/d(u0)
  };
  extern int g_mutation;
  // Departures from the reference's own expressions, for the tests that show what those
  // expressions do (bits, 0: the reference's arithmetic).
  enum : int {
    kVarHalfChargeReal = 1,   // the circle fit's (q / 2) Jacobian factor in real, not integer, arithmetic
    kVarPhiJacobianSign = 2,  // d(phi)/d(u0) with the sign of the derivative of atan2(e - d)
  };
  extern int g_variant;

}  // namespace orc

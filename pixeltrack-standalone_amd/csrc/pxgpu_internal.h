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
    uint32_t nMediumRuns;  // module runs of 65 to 256 digis, listed at the front of bigList
    uint32_t nBigRuns;     // module runs of more than 256 digis, listed from the back of bigList
    uint32_t pad[2];
  };

  struct alignas(8) HitAttr {
    int16_t iphi, ysize;
    uint16_t detInd, pad;
  };

  // What the doublet search reads of a candidate outer hit, in one 16-byte load
  // (the phi histogram's copy of the hit: z, r, iphi, ysize, detInd)
  struct alignas(16) PhiCand {
    float z, r;
    int16_t iphi, ysize;
    uint16_t detInd, pad;
  };

  // A cluster's pixel extent and integer charges, as the rec-hit position estimate
  // uses them (gpuPixelRecHits.h:73-140): the charge sum and the charges on its first
  // and last row and column.
  struct alignas(16) ClusterExt {
    uint16_t minRow, maxRow, minCol, maxCol;
    int32_t qsum, qfx, qlx, qfy, qly, pad;
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
    // [maxWords]: the clusters of a module run of at most 64 digis, found by the
    // clustering, at the run's first digi + cluster id
    ClusterExt* clExt;
    uint32_t* bigList;  // per event (stride kMaxModuleRanges): its module runs of more than 64 digis
    // clusters, per event
    uint32_t *moduleStart, *clusInModule, *moduleId, *clusModuleStart;  // stride kMaxModuleRanges
    // hits, per event (stride kMaxNumClusters)
    float *xl, *yl, *xerr, *yerr, *xg, *yg, *zg, *rg;
    float4* pos;  // (xg, yg, zg, rg) of each hit in one 16-byte load, for the CA kernels' gathers
    HitAttr* attr;  // iphi, ysize, detInd of each hit in one 8-byte load, for the doublet search
    int16_t* iphi;
    int32_t* charge;
    int16_t *xsize, *ysize;
    uint16_t* detInd;
    uint32_t* hitsLayerStart;  // stride 16
    uint32_t* phiOff;          // stride kPhiHistTotBins (padded to 1288)
    uint16_t* phiBins;         // stride kMaxNumClusters
    PhiCand* candPhi;          // hit phiBins[k] at k: a phi window's candidates in one contiguous range
    AverageGeometry* avg;      // per event
  };

  constexpr uint32_t kLayerStartStride = 16;
  constexpr uint32_t kPhiOffStride = 1288;

  // ---- ntuplets and vertices ---------------------------------------------------------
  constexpr int kNumPairs = 19;                              // layer pairs of the doublet search
  constexpr uint32_t kDoubletSlots = 6 * kMaxNumClusters;    // (pair, inner hit) slots: layer 0 is inner in 6 pairs
  constexpr uint32_t kHitStride = kMaxNumClusters + 64;      // per-event stride of per-hit offset arrays
  constexpr uint32_t kCaBlock = 256;                         // workgroup of the grid-stride CA kernels
  constexpr uint32_t kEventBlock = 1024;                     // workgroup of the per-event kernels
  constexpr uint32_t kListsInLdsMinEvents = 8;               // batches this large build lists per event in LDS
  // vertex clustering algorithms (gpuVertexFinder.h:46-60; density is the reference's default)
  enum : int { kVertexDensity = 0, kVertexDBSCAN = 1, kVertexIterative = 2 };
  constexpr uint32_t kDupPredCap = 6;                        // predecessors kept per cell for the duplicate remover

  // ---- XCD-aware workgroup order -------------------------------------------------------
  // The dispatcher hands workgroups to the chip's 8 XCDs round-robin by linear
  // id, and each XCD has its own L2.  Kernels on a (workgroups per event,
  // events) grid read an event's hits and cells from every one of its
  // workgroups; taken in dispatch order those workgroups would sit on all eight
  // XCDs and the event's data would be fetched into all eight L2s.  This
  // permutation of [0, gridDim.x * gridDim.y) gives consecutive logical ids --
  // the workgroups of one event -- one XCD: XCD k runs the k-th contiguous
  // share of the logical ids (the first n % 8 XCDs one workgroup more).  Every
  // logical workgroup still runs exactly once.
  constexpr uint32_t kXcds = 8;
  __host__ __device__ constexpr uint32_t xcdLogicalId(uint32_t pid, uint32_t n) {
    const uint32_t xcd = pid % kXcds, local = pid / kXcds;
    const uint32_t per = (n + kXcds - 1) / kXcds, tall = n % kXcds;
    return (tall == 0 || xcd < tall) ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
  }
  struct LogicalBlock {
    uint32_t x, y;
  };
  __device__ __forceinline__ LogicalBlock logicalBlock() {
    const uint32_t l = xcdLogicalId(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    return {l % gridDim.x, l / gridDim.x};
  }

  // Per-event counters of the ntuplet and vertex stages.
  struct CACounts {
    uint32_t nCells;       // cells stored (min(doublets found, maxDoublets))
    uint32_t nDoublets;    // doublets found
    uint32_t apcM, apcN;   // tuples and tuple hits found (the reference's atomic pair counter)
    uint32_t nTuples;      // tuples stored (min(apcM, maxTuples))
    uint32_t status;       // kCa* bits
    uint32_t nbLists;      // neighbour lists requested
    uint32_t trLists;      // track lists requested
    uint32_t nMulti;       // cells in two or more tuples
    uint32_t nMult;        // entries of the multiplicity container
    uint32_t nVertices;    // vertices after splitting (nvFinal)
    uint32_t vtxTracks;    // tracks given to the vertex finder
    uint32_t vstatus;      // 1: vertex container full
    uint32_t nSlots;       // (layer pair, inner hit) slots of the doublet search
    uint32_t nNeighbors;   // neighbour entries stored
    uint32_t nStart;       // cells the ntuplet walk starts from
    uint32_t nPaths;       // tuples whose path the counting walk recorded
  };

  // Device arrays of the ntuplet and vertex stages, all per event at e * stride
  // (strides follow the capacity profile chosen at stream creation).
  struct CAPool {
    CACapacities caps;
    CACounts* cnt;
    uint32_t* dOff;          // [kDoubletSlots + 64]: doublets per (pair, inner hit), then offsets
    uint64_t* dMask;         // [2 * kDoubletSlots]: accepted candidates of a slot's first 128
    uint16_t *cellInner, *cellOuter;  // [maxDoublets]
    uint8_t* cellPair;
    uint8_t* cellFlags;      // bit0 listed in its outer hit's list, bit1 killed, bit2 duplicate pass done
    uint8_t* cellUsed;       // [maxDoublets] 1: the cell has a neighbour or is one
    uint32_t* ohCount;       // [kHitStride] cells per outer hit, then fill cursors
    uint32_t* ohOff;         // [kHitStride]
    uint2* ohCells;          // [maxDoublets] cells by outer hit (fill order): x cell, y inner hit | outer hit << 16
    uint32_t* startCells;    // [maxDoublets] start cells of the ntuplet walk, ascending
    int32_t* nbSlot;         // [maxDoublets] neighbour list slot (-1 none, -2 store full)
    uint32_t* ohRank;        // [maxDoublets] position of the cell in its outer hit's list
    uint8_t* nbCount;        // [maxDoublets] neighbours stored (<= 36)
    uint32_t* nbPool;        // [maxCellLists * 36]
    uint64_t* startMask;     // [maxDoublets / 64 + 64] start cells of the ntuplet walk
    uint32_t *dfsT, *dfsH;   // [maxDoublets] tuples / hits per start cell, then offsets
    int32_t* trSlot;         // [maxDoublets] track list slot (-1 none, -2 store full)
    uint16_t* trCount;       // [maxDoublets] tracks of the cell (about 20 at PU200), then stored (<= 48)
    TupleIndex* trPool;      // [maxCellLists * 48]
    uint32_t* multiCells;    // [maxDoublets] cells with two or more tracks (any order)
    uint32_t* dupPreds;      // [maxCellLists * 8] per multiCells entry: its lower cells sharing a tuple (first 8)
    uint8_t* dupCount;       // [maxCellLists] their number, 255 if more than 8
    uint32_t* tupleOff;      // [maxTuples + 64]
    uint16_t* tupleBins;     // [5 * maxTuples]
    uint32_t* tupleCells;    // [4 * maxTuples] cells of a tuple (0xffffffff past the end)
    uint2* tuplePaths;       // [maxTuples] (start index, path) of the tuples the counting walk recorded
    uint8_t* quality;        // [maxTuples]
    uint32_t* multOff;       // [16]
    TupleIndex* multBins;    // [maxTuplesMult]
    uint16_t* detIndices;    // [5 * maxTuples]
    float* state;            // [5][maxTuples]: phi, tip, q/pt, cot(theta), zip
    float* cov;              // [15][maxTuples]: packed upper triangle, row by row
    float *pt, *eta, *chi2;  // [maxTuples]
    uint32_t* h2tOff;        // [kHitStride]
    TupleIndex* h2tBins;     // [maxHitToTuple]
    // vertices
    int16_t* idv;            // [maxTuples]
    int32_t* ndof;           // [maxTuples]: per-track neighbour counts, then per-vertex ndof
    float *zv, *wv, *vchi2, *ptv2;  // [kMaxVertices]
    uint16_t* sortInd;       // [kMaxVertices]
    TupleIndex* vItrk;       // [maxTuples] tuple index of each loaded track
    float *vZt, *vEzt2, *vPtt2;      // [maxTuples]
    int32_t *vIv, *vTmp;     // [maxTuples]
    float *vZs, *vEs;        // [maxTuples]: z and its error of the tracks in z-bin order
    int32_t* vNnS;           // [maxTuples]: neighbour counts in z-bin order
    uint8_t* vIzt;           // [maxTuples]
    uint32_t* vTrk;          // [maxTuples]: tracks by vertex, ascending within a vertex
    uint32_t* vOff;          // [kMaxVertices + 64]

    // per-event strides
    __host__ __device__ size_t cellBase(uint32_t e) const { return size_t(e) * caps.maxDoublets; }
    __host__ __device__ size_t tupleBase(uint32_t e) const { return size_t(e) * caps.maxTuples; }
    __host__ __device__ size_t offBase(uint32_t e) const { return size_t(e) * (caps.maxTuples + 64); }
  };

  // Brackets each kernel launch (HIP events on the launch stream when profiling).
  struct Timer {
    void (*begin)(void* ctx, const char* name) = nullptr;
    void (*end)(void* ctx) = nullptr;
    void* ctx = nullptr;
    void b(const char* name) const {
      if (begin)
        begin(ctx, name);
    }
    void e() const {
      if (end)
        end(ctx);
    }
  };

  // kernel launchers (stage C, D); `blocks` is the number of workgroups per event of the
  // grid-stride kernels
  struct CALaunch {
    const DevConditions* cond;
    const Pool* p;
    const CAPool* ca;
    uint32_t nEvents, blocks;
    hipStream_t s;
    const Timer* t;
    uint32_t doubletLanes;  // lanes per (layer pair, inner hit) slot of the doublet search: 1..64
    int vertexAlgo;         // vertex clustering: kVertexDensity, kVertexDBSCAN or kVertexIterative
    bool earlyFishbone;     // fishbone on the connected cells before the ntuplet search
    bool lateFishbone;      // fishbone on the cells in tuples after it, and the cleaner after classify
    bool riemannFit;        // Riemann fit in place of the broken-line fit
    unsigned long long* counters;  // doStats counters (11), null when off
    double bField;          // the fits' field in 1/GeV (a float widened, as makeTuples passes it)
    float vertexPtMin;      // tracks below this pt are left out of the vertex finder
    uint32_t connectStage;  // candidate cells connect stages per wave in LDS: 128, 256 or 512
    uint32_t fishboneWindow;  // list entries the fishbone stages per wave in LDS: 128, 256 or 384
  };
  void launchDoublets(const CALaunch& l);
  void launchConnect(const CALaunch& l);
  void launchFishbone(const CALaunch& l, bool checkTrack);
  void launchNtuplets(const CALaunch& l);
  void launchMultiplicity(const CALaunch& l);
  void launchFit(const CALaunch& l);
  void launchRiemannFit(const CALaunch& l);
  void launchClassify(const CALaunch& l);
  void launchHitToTuple(const CALaunch& l);
  void launchVertices(const CALaunch& l);

  // kernel launchers (stage A, B); nBlocks = total decode workgroups of the batch
  void launchDecode(const DevConditions& c, const Pool& p, uint32_t nEvents, uint32_t maxWords,
                    hipStream_t s);
  void launchModuleStarts(const Pool& p, uint32_t nEvents, hipStream_t s);
  void launchClusterize(const Pool& p, uint32_t nEvents, hipStream_t s);
  void launchHits(const DevConditions& c, const Pool& p, uint32_t nEvents, hipStream_t s);
  void launchPhiBinner(const DevConditions& c, const Pool& p, uint32_t nEvents, hipStream_t s);

}  // namespace pxg

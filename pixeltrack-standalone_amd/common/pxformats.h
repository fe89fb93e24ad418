// Binary layouts and detector constants shared by the event generator, the GPU
// library, the host plugins and the oracle.  Only data layouts and constants
// live here: every component implements its own arithmetic, so that a parity
// test between two components compares two independent implementations.
//
// Layouts follow the reference's on-disk conditions formats byte for byte:
//   cabling map   CondFormats/SiPixelFedCablingMapGPU.h:4-24
//   gains         CondFormats/SiPixelGainForHLTonGPU.h:160-211
//   CPE params    CondFormats/pixelCPEforGPU.h:20-54, DataFormats/SOARotation.h
//   beam spot     DataFormats/BeamSpotPOD.h:12-19
//   error record  DataFormats/PixelErrors.h:12-17
// Topology constants: Geometry/phase1PixelTopology.h:9-51.
#pragma once

#include <cstddef>
#include <cstdint>

namespace pxf {

  // ---- pixel module topology (phase1PixelTopology.h:9-36) -------------------
  constexpr uint32_t kRowsPerRoc = 80;
  constexpr uint32_t kColsPerRoc = 52;
  constexpr uint32_t kRowsPerModule = 2 * kRowsPerRoc;  // 160
  constexpr uint32_t kColsPerModule = 8 * kColsPerRoc;  // 416
  constexpr uint32_t kLastRow = kRowsPerModule - 1;
  constexpr uint32_t kLastCol = kColsPerModule - 1;
  constexpr int kXOffset = -81;   // local-x origin, in 100 um units
  constexpr int kYOffset = -216;  // local-y origin, in 150 um units

  constexpr uint32_t kNumModules = 1856;
  constexpr uint32_t kNumLayers = 10;
  constexpr uint32_t kLayerStart[kNumLayers + 1] = {0, 96, 320, 672, 1184, 1296, 1408, 1520, 1632, 1744, 1856};
  constexpr uint32_t kModulesInBarrel = 1184;
  constexpr uint32_t kLaddersInBarrel = kModulesInBarrel / 8;  // 148
  // largest power of two dividing every layer start (phase1PixelTopology.h:64-81)
  constexpr uint32_t kModuleStride = 16;
  constexpr uint32_t kLayerIndexSize = kNumModules / kModuleStride;  // 116

  // ---- capacities of the per-event products --------------------------------
  constexpr uint32_t kMaxFed = 150;
  constexpr uint32_t kMaxLink = 48;
  constexpr uint32_t kMaxRoc = 8;
  constexpr uint32_t kCablingSize = kMaxFed * kMaxLink * kMaxRoc;  // 57600
  constexpr uint32_t kMaxFedWords = kMaxFed * 2000;                // 300000
  constexpr uint32_t kMaxNumModules = 2000;
  constexpr uint32_t kMaxPixInModule = 4000;
  constexpr uint32_t kMaxClustersPerModule = 1024;
  constexpr uint32_t kMaxNumClusters = 48 * 1024;  // hits per event
  constexpr uint32_t kMaxHitsInIter = 160;
  constexpr uint16_t kInvId = 9999;
  constexpr uint32_t kPhiBins = 128;
  constexpr uint32_t kPhiHistTotBins = kNumLayers * kPhiBins + 1;  // 1281

  // Capacities of the ntuplet containers.  The reference profile is the
  // serial program's (CAConstants.h:21-64, CAHitNtupletGeneratorOnGPU.cc:48,
  // PixelTrackHeterogeneous.h:62); the high-occupancy profile scales them for
  // PU200-like events (about 1.2M cells, 122k tuples and up to 160 cells
  // sharing an outer hit with this geometry).  Tuple indices are stored as 32
  // bit (the reference's tindex_type is 16 bit, CAConstants.h:46, which caps
  // it at 65535 tuples; the values are the same below that).
  struct CACapacities {
    uint32_t maxDoublets;      // cells
    uint32_t maxTuples;        // tuple container bins (TrackSoA stride)
    uint32_t maxTuplesMult;    // tuple multiplicity entries
    uint32_t maxHitToTuple;    // hit -> tuple entries
    uint32_t maxCellLists;     // cell neighbour lists / cell track lists (each store)
    uint32_t maxCellsPerHit;   // cells kept per outer hit (CAConstants.h maxCellsPerHit)
  };
  using TupleIndex = uint32_t;
  constexpr TupleIndex kNoTuple = 0xffffffffu;  // "none" of the duplicate removers
  // Per-event status bits of the ntuplet stage (a set bit names a container
  // that reached its capacity; what was stored is still well defined).
  enum : uint32_t {
    kCaCellOverflow = 1u << 0,           // more doublets than maxDoublets
    kCaOuterHitOverflow = 1u << 1,       // a hit is the outer hit of more than 128 cells
    kCaNeighborOverflow = 1u << 2,       // a cell with more than 36 outer neighbours
    kCaNeighborStoreOverflow = 1u << 3,  // neighbour lists store full
    kCaTrackOverflow = 1u << 4,          // a cell in more than 48 tuples
    kCaTrackStoreOverflow = 1u << 5,     // track lists store full
    kCaTupleOverflow = 1u << 6,          // more tuples than maxTuples
    kCaDepthOverflow = 1u << 7,          // a chain longer than 4 cells
    kCaHitToTupleOverflow = 1u << 8,     // hit -> tuple container full
    kCaMultiplicityOverflow = 1u << 9,   // tuple multiplicity container full
  };
  constexpr uint32_t kCaMaxNeighbors = 36;      // CAConstants.h CellNeighbors
  constexpr uint32_t kCaMaxCellTracks = 48;     // CAConstants.h CellTracks
  constexpr uint32_t kCaMaxCellsPerHitCap = 1024;  // largest maxCellsPerHit of any profile
  constexpr uint32_t kMaxVertices = 1024;       // ZVertexSoA.h MAXVTX
  enum : uint8_t { kQualityBad = 0, kQualityDup = 1, kQualityLoose = 2 };

  constexpr CACapacities kCapsReference{458752, 32 * 1024, 24 * 1024, 4 * 24 * 1024, 64 * 1024, 128};
  constexpr CACapacities kCapsHighOccupancy{4 * 458752, 192 * 1024, 192 * 1024, 4 * 192 * 1024, 256 * 1024, 256};

  // ---- conditions ------------------------------------------------------------
  struct alignas(128) SiPixelFedCablingMapGPU {
    alignas(128) uint32_t fed[kCablingSize];
    alignas(128) uint32_t link[kCablingSize];
    alignas(128) uint32_t roc[kCablingSize];
    alignas(128) uint32_t RawId[kCablingSize];
    alignas(128) uint32_t rocInDet[kCablingSize];
    alignas(128) uint32_t moduleId[kCablingSize];
    alignas(128) uint8_t badRocs[kCablingSize];
    alignas(128) uint32_t size;
  };
  static_assert(sizeof(SiPixelFedCablingMapGPU) == 1440128, "cabling map layout");

  struct GainRangeAndCols {
    uint32_t first;   // byte offset of the module's first column block
    uint32_t second;  // one past its last byte
    int32_t nCols;
  };
  struct GainDecoding {
    uint8_t gain;
    uint8_t ped;
  };
  struct SiPixelGainForHLTonGPU {
    uint64_t v_pedestals;  // a pointer in the reference; meaningless on disk
    GainRangeAndCols rangeAndCols[kMaxNumModules];
    float minPed_, maxPed_, minGain_, maxGain_;
    float pedPrecision, gainPrecision;
    uint32_t numberOfRowsAveragedOver_;
    uint32_t nBinsToUseForEncoding_;
    uint32_t deadFlag_;
    uint32_t noisyFlag_;
  };
  static_assert(sizeof(SiPixelGainForHLTonGPU) == 24048, "gain layout");

  struct CPECommonParams {
    float theThicknessB, theThicknessE, thePitchX, thePitchY;
  };
  struct SOAFrameF {
    float px, py, pz;
    float R11, R12, R13, R21, R22, R23, R31, R32, R33;  // rows = local x,y,z axes in global coords
  };
  struct CPEDetParams {
    bool isBarrel;
    bool isPosZ;
    uint16_t layer;
    uint16_t index;
    uint32_t rawId;
    float shiftX, shiftY;
    float chargeWidthX, chargeWidthY;
    float x0, y0, z0;  // the beam-line origin in the module's local frame
    float sx[3], sy[3];
    SOAFrameF frame;
  };
  static_assert(sizeof(CPEDetParams) == 112, "DetParams layout");
  struct CPELayerGeometry {
    uint32_t layerStart[kNumLayers + 1];
    uint8_t layer[kLayerIndexSize];
  };
  struct AverageGeometry {
    float ladderZ[kLaddersInBarrel];
    float ladderX[kLaddersInBarrel];
    float ladderY[kLaddersInBarrel];
    float ladderR[kLaddersInBarrel];
    float ladderMinZ[kLaddersInBarrel];
    float ladderMaxZ[kLaddersInBarrel];
    float endCapZ[2];
  };

  struct BeamSpotPOD {
    float x, y, z;
    float sigmaZ;
    float beamWidthX, beamWidthY;
    float dxdz, dydz;
    float emittanceX, emittanceY;
    float betaStar;
  };

  struct PixelErrorCompact {
    uint32_t rawId;
    uint32_t word;
    uint8_t errorType;
    uint8_t fedId;
  };
  static_assert(sizeof(PixelErrorCompact) == 12, "error record layout");

  // ---- raw word fields (SiPixelRawToClusterGPUKernel.h:35-72) ----------------
  constexpr uint32_t kAdcShift = 0, kAdcMask = 0xFF;
  constexpr uint32_t kPxidShift = 8, kPxidMask = 0xFF;
  constexpr uint32_t kDcolShift = 16, kDcolMask = 0x1F;
  constexpr uint32_t kRocShift = 21, kRocMask = 0x1F;
  constexpr uint32_t kLinkShift = 26, kLinkMask = 0x3F;
  constexpr uint32_t kRowShiftL1 = 8, kRowMaskL1 = 0x7F;
  constexpr uint32_t kColShiftL1 = 15, kColMaskL1 = 0x3F;

  // ---- FED framing (fed_header.h, fed_trailer.h) ---------------------------------
  constexpr uint32_t kFirstPixelFed = 1200;

}  // namespace pxf

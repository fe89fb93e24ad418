// Event and conditions products of the pixel chain, named after the
// reference's (CUDADataFormats/*, CondFormats/*, DataFormats/*).  Event
// products are views of device memory in the EDM stream's pool plus the
// host-visible counts; they stay valid for the event they belong to (the next
// event of the same stream reuses the pool).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pxgpu.h"
#include "../common/pxformats.h"

// ---- conditions (host copies read from the data directory) -----------------------------
struct SiPixelFedIds {
  std::vector<uint32_t> ids;  // fedIds.bin
};

struct SiPixelFedCablingMapGPUWrapper {  // cablingMap.bin: the map and the modules to unpack
  std::unique_ptr<pxf::SiPixelFedCablingMapGPU> map;
  std::vector<uint8_t> modToUnp;
};

struct SiPixelGainCalibrationForHLTGPU {  // gain.bin: the header and the gain/pedestal payload
  pxf::SiPixelGainForHLTonGPU gain;
  std::vector<uint8_t> data;
};

struct PixelCPEFast {  // cpefast.bin
  pxf::CPECommonParams common;
  std::vector<pxf::CPEDetParams> det;
  pxf::AverageGeometry average;
  pxf::CPELayerGeometry layer;
};

// The device with the conditions uploaded (made once from the four products above).
class PxgDevice {
public:
  explicit PxgDevice(pxg_context* ctx) : ctx_(ctx) {}
  ~PxgDevice() { pxg_finalize(ctx_); }
  PxgDevice(PxgDevice const&) = delete;
  PxgDevice& operator=(PxgDevice const&) = delete;
  pxg_context* get() const { return ctx_; }

private:
  pxg_context* ctx_;
};

// ---- event products ------------------------------------------------------------------
struct FEDRawDataCollection {  // one raw.bin record: nfeds, then {fedId, size, bytes} per FED
  std::vector<uint8_t> record;
};

struct BeamSpotPOD {
  pxg_beamspot bs;
};

// The stream the products live on is carried along, so that every producer of
// an event enqueues on the same HIP stream and device pool.
struct SiPixelDigisSoA {
  pxg_stream* stream;
  pxg_digis view;
  uint32_t nModules() const { return view.nModules; }
  uint32_t nDigis() const { return view.nDigis; }
};

struct SiPixelClustersSoA {
  pxg_stream* stream;
  pxg_clusters view;
  uint32_t nClusters() const { return view.nClusters; }
};

struct SiPixelDigiErrorsSoA {
  pxg_stream* stream;
  pxg_errors view;
  int status;  // capacity bits of the raw-to-cluster step
};

struct TrackingRecHit2DCPU {
  pxg_stream* stream;
  pxg_hits view;
  uint32_t nHits() const { return view.nHits; }
};

struct PixelTrackHeterogeneous {
  pxg_stream* stream;
  pxg_tracks view;
  int status;
  uint32_t nTracks() const { return view.nTuples; }
};

struct ZVertexHeterogeneous {
  pxg_stream* stream;
  pxg_vertices view;
  uint32_t nvFinal() const { return view.nvFinal; }
};

// ---- expected counts read with the raw data (DataFormats/DigiClusterCount.h, TrackCount.h,
// VertexCount.h; bin/Source.cc:58-68 reads them from digicluster.bin, tracks.bin, vertices.bin)
class DigiClusterCount {
public:
  DigiClusterCount(uint32_t nm, uint32_t nd, uint32_t nc) : modules_(nm), digis_(nd), clusters_(nc) {}
  uint32_t nModules() const { return modules_; }
  uint32_t nDigis() const { return digis_; }
  uint32_t nClusters() const { return clusters_; }

private:
  uint32_t modules_, digis_, clusters_;
};

class TrackCount {
public:
  explicit TrackCount(uint32_t n) : tracks_(n) {}
  uint32_t nTracks() const { return tracks_; }

private:
  uint32_t tracks_;
};

class VertexCount {
public:
  explicit VertexCount(uint32_t n) : vertices_(n) {}
  uint32_t nVertices() const { return vertices_; }

private:
  uint32_t vertices_;
};

// ---- host copies made by the transfer modules (the CUDA program's *SoAFromCUDA products) ----
struct SiPixelDigisSoAHost {  // SiPixelDigisSoAFromCUDA.cc: adc, clus, pdigi, rawIdArr
  uint32_t nModules = 0, nDigis = 0;
  std::vector<uint16_t> adc;
  std::vector<int32_t> clus;
  std::vector<uint32_t> pdigi, rawIdArr;
};

struct PixelTrackSoAHost {  // TrackSoAT (PixelTrackHeterogeneous.h) for the stored tuples
  uint32_t nTracks = 0;
  std::vector<uint32_t> hitOff;  // [nTracks + 1]
  std::vector<uint16_t> hitIndices;
  std::vector<uint8_t> quality;
  std::vector<float> chi2, eta, pt;
  std::vector<float> state;  // [5][nTracks]: phi, tip, q/pt, cot(theta), zip
  int nHits(uint32_t i) const { return int(hitOff[i + 1] - hitOff[i]); }
  float phi(uint32_t i) const { return state[i]; }
  float tip(uint32_t i) const { return state[nTracks + i]; }
  float zip(uint32_t i) const { return state[4 * nTracks + i]; }
};

struct ZVertexSoAHost {  // ZVertexSoA for the final vertices
  uint32_t nvFinal = 0;
  std::vector<float> zv, wv, chi2, ptv2;
  std::vector<int32_t> ndof;
  std::vector<uint16_t> sortInd;
};

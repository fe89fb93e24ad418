/* C ABI of the MI355X (gfx950) pixel reconstruction library, libpxgpu.so.
 *
 * This is the boundary the host plugins call (SiPixelRawToClusterCUDA,
 * SiPixelRecHitCUDA, ...): plain pointers, sizes and status codes, no C++ or
 * torch types.  Products are views of device memory owned by a pxg_stream; a
 * view stays valid until the next call that runs events on the same stream
 * (one event in flight per EDM stream, as bin/StreamSchedule.cc:61-93 runs
 * them).
 *
 * Status codes: 0 ok; > 0 a PXG_STATUS_* bit set (a capacity was exceeded;
 * the product is truncated exactly as the reference truncates it, never
 * silently); < 0 a HIP or usage error, described by pxg_last_error().
 * No entry point throws.
 */
#ifndef PXGPU_H
#define PXGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pxg_context pxg_context;
typedef struct pxg_stream pxg_stream;

enum {
  PXG_STATUS_MODULE_OVERFLOW = 1,  /* > 2000 module runs in an event */
  PXG_STATUS_PIXEL_OVERFLOW = 2,   /* a module with > 4000 pixels */
  PXG_STATUS_CLUSTER_OVERFLOW = 4, /* a module with > 1024 clusters */
  PXG_STATUS_HIT_OVERFLOW = 8,     /* > 49152 clusters in an event */
  PXG_ERR_HIP = -1,
  PXG_ERR_ARG = -2,
  PXG_ERR_CAPACITY = -3,
};

/* Conditions in the reference's in-memory layouts (see pixeltrack-standalone_amd/common/pxformats.h).
 * Replaces the ES products SiPixelFedCablingMapGPUWrapper, SiPixelGainCalibrationForHLTGPU and
 * PixelCPEFast (plugin-SiPixelClusterizer/SiPixelFedCablingMapGPUWrapperESProducer.cc:20-41,
 * SiPixelGainCalibrationForHLTGPUESProducer.cc:20-30, CondFormats/PixelCPEFast.cc:17-28). */
typedef struct {
  const void* cabling;            /* SiPixelFedCablingMapGPU */
  const uint8_t* mod_to_unp;      /* 57600 bytes */
  const void* gain;               /* SiPixelGainForHLTonGPU */
  const uint8_t* gain_data;
  uint32_t gain_data_bytes;
  const void* cpe_common;         /* pixelCPEforGPU::CommonParams */
  const void* cpe_det;            /* pixelCPEforGPU::DetParams[n_det] */
  uint32_t n_det;
  const void* cpe_layer;          /* pixelCPEforGPU::LayerGeometry */
  const void* cpe_average;        /* phase1PixelTopology::AverageGeometry */
} pxg_conditions;

typedef struct {
  float x, y, z, sigmaZ, beamWidthX, beamWidthY, dxdz, dydz, emittanceX, emittanceY, betaStar;
} pxg_beamspot; /* DataFormats/BeamSpotPOD.h:12-19 */

/* SiPixelDigisSoA (CUDADataFormats/SiPixelDigisSoA.h:7-91) */
typedef struct {
  uint16_t *xx, *yy, *adc, *moduleInd;
  int32_t* clus;
  uint32_t *pdigi, *rawIdArr;
  uint32_t nModules, nDigis;
} pxg_digis;

/* SiPixelClustersSoA (CUDADataFormats/SiPixelClustersSoA.h:7-71) */
typedef struct {
  uint32_t *moduleStart, *clusInModule, *moduleId, *clusModuleStart;
  uint32_t nClusters;
} pxg_clusters;

/* SiPixelDigiErrorsSoA (CUDADataFormats/SiPixelDigiErrorsSoA.h): PixelErrorCompact records, word order */
typedef struct {
  void* data; /* PixelErrorCompact[size], 12 bytes each */
  uint32_t size;
} pxg_errors;

/* TrackingRecHit2DSOAView (CUDADataFormats/TrackingRecHit2DSOAView.h:13-101) */
typedef struct {
  float *xl, *yl, *xerr, *yerr, *xg, *yg, *zg, *rg;
  int16_t* iphi;
  int32_t* charge;
  int16_t *xsize, *ysize;
  uint16_t* detInd;
  uint32_t* hitsModuleStart; /* = clusters.clusModuleStart */
  uint32_t* hitsLayerStart;  /* [11] */
  uint32_t* phiOff;          /* phi binner offsets [1281] */
  uint16_t* phiBins;         /* phi binner contents [nHits] */
  float* averageGeometry;    /* AverageGeometry corrected for the beam spot */
  uint32_t nHits;
} pxg_hits;

/* Context: device + conditions uploaded once, immutable afterwards (pxg_init is the
 * ESProducer -> device transfer the CUDA program does in its ES products). */
int pxg_init(const pxg_conditions* cond, pxg_context** out);
void pxg_finalize(pxg_context* ctx);
const char* pxg_last_error(void);
int pxg_device_name(char* buf, size_t n);

/* Stream: one hipStream and a device-resident pool sized for max_events events of
 * max_words raw words in total (one per EDM stream; the pool replaces the caching
 * allocator of src/cuda/CUDACore). */
int pxg_stream_create(pxg_context* ctx, uint32_t max_events, uint32_t max_words, pxg_stream** out);
void pxg_stream_destroy(pxg_stream* s);
int pxg_sync(pxg_stream* s);

/* ---- plain per-event path (one event, one launch sequence) ------------------------- */

/* Raw words -> digis and errors (RawToDigi_kernel + calibDigis + countModules:
 * SiPixelRawToClusterGPUKernel.cc:342-462, gpuCalibPixel.h:22-66, gpuClustering.h:19-37).
 * words/fedIds are host arrays as built by the FED loop of SiPixelRawToClusterCUDA::produce
 * (SiPixelRawToClusterCUDA.cc:87-138): fedIds holds fedId-1200 for every two words. */
int pxg_raw_to_digi(pxg_stream* s, const uint32_t* words, const uint8_t* fedIds, uint32_t nWords,
                    pxg_digis* digis, pxg_errors* errors);

/* Raw words -> digis, clusters, errors (SiPixelRawToClusterGPUKernel::makeClusters,
 * SiPixelRawToClusterGPUKernel.cc:509-600). */
int pxg_raw_to_cluster(pxg_stream* s, const uint32_t* words, const uint8_t* fedIds, uint32_t nWords,
                       pxg_digis* digis, pxg_clusters* clusters, pxg_errors* errors);

/* Digis + clusters -> rec-hits (PixelRecHitGPUKernel::makeHits, PixelRecHits.cc:33-53); the
 * inputs must be the products of the last pxg_raw_to_cluster on the same stream. */
int pxg_rechits(pxg_stream* s, const pxg_digis* digis, const pxg_clusters* clusters, const pxg_beamspot* bs,
                pxg_hits* hits);

/* ---- resident / batched path --------------------------------------------------------- */

typedef struct {
  const uint32_t* words;
  const uint8_t* fedIds;
  uint32_t nWords;
} pxg_raw_view;

/* Copies n events' words into the stream's device pool (256-word aligned slots) and
 * records their beam spot; afterwards the events are resident in HBM. */
int pxg_load_events(pxg_stream* s, uint32_t n, const pxg_raw_view* raws, const pxg_beamspot* bs);

/* Enqueues the chain on resident events [first, first+count) as ONE launch per stage
 * (count = 1 is the plain per-event launch sequence).  upto: 1 digis, 2 clusters,
 * 3 rec-hits.  Asynchronous; pxg_fetch_counts synchronises. */
int pxg_run(pxg_stream* s, uint32_t first, uint32_t count, int upto);

/* Copies the per-event counters of [first, first+count) to the host and synchronises;
 * returns the OR of their status bits. */
int pxg_fetch_counts(pxg_stream* s, uint32_t first, uint32_t count);

/* Product views of resident event e (after pxg_fetch_counts covering e). */
int pxg_event_products(pxg_stream* s, uint32_t e, pxg_digis* digis, pxg_clusters* clusters, pxg_errors* errors,
                       pxg_hits* hits);

/* Host copies of resident event e's counters (after pxg_fetch_counts covering e):
 * out[0] module runs, out[1] clusters (= hits), out[2] error records, out[3] status bits. */
int pxg_event_counts(pxg_stream* s, uint32_t e, uint32_t out[4]);

/* Device -> host copy of product memory (validation, --transfer). */
int pxg_copy_to_host(pxg_stream* s, void* dst, const void* src, size_t bytes);

/* ---- profiling ------------------------------------------------------------------------ */

/* When enabled, every kernel launch of the stream is bracketed by HIP events on the
 * stream it runs on; pxg_profile_read sums the durations per kernel. */
int pxg_profile_enable(pxg_stream* s, int on);
/* Fills up to max entries: kernel name, number of launches, total milliseconds.  Returns
 * the number of kernels; resets the accumulators. */
int pxg_profile_read(pxg_stream* s, const char** names, uint32_t* calls, double* total_ms, int max);

#ifdef __cplusplus
}
#endif
#endif

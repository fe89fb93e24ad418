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
  /* ntuplet stage: the container bits of pxformats.h (kCa*) shifted by 8 */
  PXG_STATUS_CA_SHIFT = 8,
  /* vertex stage: more than 1024 vertices */
  PXG_STATUS_VERTEX_OVERFLOW = 1 << 20,
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

/* PixelTrackHeterogeneous / TrackSoAT (CUDADataFormats/PixelTrackHeterogeneous.h:13-70):
 * tuples as an offsets/contents pair (hitIndices), per-tuple fit results at index t,
 * state and covariance as [component][stride] (TrajectoryStateSoA.h:8-60). */
typedef struct {
  uint32_t* hitOff;      /* [stride + 1] */
  uint16_t* hitIndices;  /* [5 * stride] */
  uint16_t* detIndices;  /* [5 * stride] */
  uint8_t* quality;      /* trackQuality: 0 bad, 1 dup, 2 loose */
  float *chi2, *eta, *pt;
  float* state;          /* [5][stride]: phi, tip, q/pt, cot(theta), zip */
  float* covariance;     /* [15][stride]: packed upper triangle, row by row */
  uint32_t stride;       /* tuple capacity */
  uint32_t nTuples;
} pxg_tracks;

/* ZVertexSoA (CUDADataFormats/ZVertexSoA.h:10-26) */
typedef struct {
  int16_t* idv;          /* [stride]: vertex of each track, -1 none */
  float *zv, *wv, *chi2, *ptv2;  /* [1024] */
  int32_t* ndof;         /* [stride]; the first nvFinal entries are per vertex */
  uint16_t* sortInd;     /* [1024]: vertices by increasing sum pt^2 */
  uint32_t nvFinal;
} pxg_vertices;

/* Internal containers of the ntuplet stage, exposed for parity checks
 * (GPUCACell, CAConstants.h containers).  Lists are (slot, count) pairs into a pool. */
typedef struct {
  uint16_t *cellInner, *cellOuter;
  uint8_t* cellPair;
  uint8_t* cellFlags;    /* bit0 listed in its outer hit's list, bit1 killed by the fishbone */
  uint8_t* cellUsed;     /* one byte per cell: 1 if the cell has a neighbour or is one */
  uint32_t* outerHitOff;
  uint32_t* outerHitCells;  /* cells by outer hit in fill order, as u32 pairs (cell, inner hit | outer hit << 16) */
  int32_t* neighborSlot; uint8_t* neighborCount; uint32_t* neighborPool;  /* 36 per slot */
  int32_t* trackSlot; uint16_t* trackCount; uint32_t* trackPool;          /* 48 tuple indices per slot */
  uint32_t* multOff; uint32_t* multBins;   /* tuple multiplicity (tuple indices) */
  uint32_t *hitToTupleOff; uint32_t* hitToTupleBins;
  uint32_t counts[20];   /* nCells, nDoublets, apcM, apcN, nTuples, status, nbLists, trLists,
                            nMulti, nMult, nVertices, vtxTracks, vstatus, nSlots, nNeighbors, nStart,
                            nPaths, 0, 0, 0 */
} pxg_ca_internals;

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
/* Same, choosing the capacities of the ntuplet containers: PXG_PROFILE_REFERENCE (the serial
 * program's, CAConstants.h:21-64) or PXG_PROFILE_HIGH_OCCUPANCY (scaled for PU200-like events). */
enum { PXG_PROFILE_REFERENCE = 0, PXG_PROFILE_HIGH_OCCUPANCY = 1 };
int pxg_stream_create_profile(pxg_context* ctx, uint32_t max_events, uint32_t max_words, int profile,
                              pxg_stream** out);
/* Device bytes a stream's pool of max_events events and max_words raw words takes. */
size_t pxg_pool_bytes(uint32_t max_events, uint32_t max_words, int profile);
/* Capacities of a profile's ntuplet containers (the values of CAConstants.h:18-39 for
 * PXG_PROFILE_REFERENCE).  Returns 0, or PXG_ERR_ARG for an unknown profile. */
typedef struct {
  uint32_t max_doublets, max_tuples, max_tuples_mult, max_hit_to_tuple, max_cell_lists, max_cells_per_hit;
} pxg_capacities;
int pxg_profile_capacities(int profile, pxg_capacities* out);
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

/* No counterpart in the reference: the same products as pxg_raw_to_cluster for one event,
 * for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops and harnesses that forbid hipMalloc).  It allocates nothing: the inputs and outputs
 * are the caller's device arrays, and the kernels' scratch (event descriptor and counters,
 * module-boundary and error masks, error records at their word index, the clusteriser's
 * work list: PXG_NOALLOC_SCRATCH_BYTES) is a static device array that each GPU holds once
 * until the process ends.  words (n_words u32) and fed_ids (n_words / 2 u8) are device
 * arrays; digis->xx, yy, adc, moduleInd, clus, pdigi, rawIdArr hold n_words entries each,
 * clusters->moduleStart, clusInModule, moduleId, clusModuleStart 2001 each, errors->data
 * n_words PixelErrorCompact records; the counts (nModules, nDigis, nClusters, size) are
 * written back.  hip_stream is the hipStream_t to run on (0: the null stream); the call
 * returns after that stream has finished.  n_words above PXG_NOALLOC_MAX_WORDS is rejected
 * with PXG_ERR_CAPACITY.  Calls that overlap on one GPU share its scratch array and race:
 * serialise them.  Returns the status of pxg_raw_to_cluster. */
#define PXG_NOALLOC_MAX_WORDS 300000u /* the words of one event at most (150 FEDs x 2000) */
#define PXG_NOALLOC_SCRATCH_BYTES 3684096u
int pxg_raw_to_cluster_noalloc(pxg_context* ctx, void* hip_stream, const uint32_t* words, const uint8_t* fed_ids,
                               uint32_t n_words, pxg_digis* digis, pxg_clusters* clusters, pxg_errors* errors);

/* Digis + clusters -> rec-hits (PixelRecHitGPUKernel::makeHits, PixelRecHits.cc:33-53); the
 * inputs must be the products of the last pxg_raw_to_cluster on the same stream. */
int pxg_rechits(pxg_stream* s, const pxg_digis* digis, const pxg_clusters* clusters, const pxg_beamspot* bs,
                pxg_hits* hits);

/* Rec-hits -> ntuplets with broken-line fits and quality (CAHitNtupletGeneratorOnGPU::makeTuples,
 * CAHitNtupletGeneratorOnGPU.cc:96-126; CAHitNtupletCUDA::produce, CAHitNtupletCUDA.cc:38-47); the
 * hits must be the products of the last pxg_rechits on the same stream.  bfield is the field of the
 * helix fits in 1/GeV that produce passes to makeTuples (0.0114256972711507, CAHitNtupletCUDA.cc:31);
 * it stays the stream's setting (pxg_set_bfield) for later batched runs. */
int pxg_ntuplets(pxg_stream* s, const pxg_hits* hits, float bfield, pxg_tracks* tracks);

/* Tracks -> vertices (PixelVertexProducerCUDA::produce, PixelVertexProducerCUDA.cc:26-40;
 * gpuVertexFinder::Producer::make, gpuVertexFinderImpl.h:84-111); the tracks must be the
 * products of the last pxg_ntuplets on the same stream.  ptMin is the producer's track pt cut
 * (m_ptMin = 0.5 GeV, PixelVertexProducerCUDA.cc:37, applied in loadTracks,
 * gpuVertexFinderImpl.h:33); it stays the stream's setting (pxg_set_vertex_ptmin). */
int pxg_find_vertices(pxg_stream* s, const pxg_tracks* tracks, float ptMin, pxg_vertices* vertices);

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
 * 3 rec-hits, 4 ntuplets and fits, 5 vertices.  Asynchronous; pxg_fetch_counts synchronises. */
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

/* Track and vertex views of resident event e (after pxg_fetch_counts covering e). */
int pxg_event_tracks(pxg_stream* s, uint32_t e, pxg_tracks* tracks, pxg_vertices* vertices);

/* Views of the ntuplet stage's internal containers of resident event e. */
int pxg_event_ca_internals(pxg_stream* s, uint32_t e, pxg_ca_internals* out);

/* Device -> host copy of product memory (validation, --transfer). */
int pxg_copy_to_host(pxg_stream* s, void* dst, const void* src, size_t bytes);

/* ---- serving loop (continuous batching) ---------------------------------------------- */

/* No counterpart in the reference (the framework runs one event per EDM stream).  Events are
 * submitted one by one from any host thread; a dispatcher thread admits waiting events into a
 * batch whenever one of n_streams slots is free, as long as the batch's events (<= max_batch)
 * and raw words fit the slot's device pool, and runs the chain on the batch as one launch per
 * stage; slots run concurrently.  The pools are sized from hbm_budget (bytes over all slots;
 * 0: max_batch events of 300000 words each).  Completed events carry their products packed in
 * one host buffer (layout: 16 u32 header | tupleOff[nT+1] u32 | tupleBins u16 | quality u8 |
 * state[5][nT] f32 | idv[nT] i16 | zv[nV] f32 | sortInd[nV] u16, parts 4-byte aligned, nBins
 * tuple hits reserved as 5 nT). */
typedef struct pxg_server pxg_server;
typedef struct {
  uint32_t max_batch;      /* events per batch (one launch per stage) */
  uint32_t n_streams;      /* batches in flight */
  uint64_t hbm_budget;     /* device bytes for all slots; 0 = no limit */
  int profile;             /* PXG_PROFILE_* */
  int upto;                /* last stage, 1..5 */
  uint32_t doublet_lanes;  /* launch shape (0: default) */
  uint32_t grid_target;
  int transfer;            /* 1: copy each event's products to the host (pxg_server_products);
                              0: only counters (the products stay on the device and are reused) */
} pxg_server_config;

typedef struct {
  uint64_t id;
  int status;              /* PXG_STATUS_* bits of the event */
  uint64_t t_submit_ns, t_launch_ns, t_done_ns;  /* steady clock */
  uint32_t batch_size;     /* events in the batch that carried it */
  uint32_t n_words, n_hits, n_tuples, n_vertices;
} pxg_server_result;

int pxg_server_create(pxg_context* ctx, const pxg_server_config* cfg, pxg_server** out);
void pxg_server_destroy(pxg_server* sv); /* finishes the events already submitted */
/* Copies the caller's words and FED ids into one of the server's pinned chunks and queues the
 * event (the caller may reuse its buffers on return; waits for a free chunk when all hold
 * queued events). */
int pxg_server_submit(pxg_server* sv, const uint32_t* words, const uint8_t* fedIds, uint32_t nWords,
                      const pxg_beamspot* bs, uint64_t* id);
/* Queues the event without a host copy: words and fedIds must be mapped pinned host memory
 * (pxg_host_alloc), which the device reads while it loads the batch; the caller leaves them
 * unchanged until the event is polled.  PXG_ERR_ARG for other memory. */
int pxg_server_submit_pinned(pxg_server* sv, const uint32_t* words, const uint8_t* fedIds, uint32_t nWords,
                             const pxg_beamspot* bs, uint64_t* id);
/* Mapped pinned host memory (readable by the device) for pxg_server_submit_pinned. */
int pxg_host_alloc(size_t bytes, void** out);
void pxg_host_free(void* p);
/* Completed events (up to max), waiting up to wait_ms for the first; returns the number. */
int pxg_server_poll(pxg_server* sv, pxg_server_result* out, int max, int wait_ms);
/* Packed products of a polled event, valid until pxg_server_release. */
int pxg_server_products(pxg_server* sv, uint64_t id, const void** packed, uint32_t* bytes);
int pxg_server_release(pxg_server* sv, uint64_t id);
int pxg_server_info(pxg_server* sv, uint32_t* max_batch_words, uint64_t* batches, uint64_t* events);

/* ---- tuning ---------------------------------------------------------------------------- */

/* Launch shape of the stream's kernels (products do not depend on it): lanes that scan one
 * (layer pair, inner hit) slot of the doublet search together (power of two, 1..64), and
 * the number of workgroups of a grid-stride launch, shared by the events of the launch. */
int pxg_set_tuning(pxg_stream* s, uint32_t doublet_lanes, uint32_t grid_target);

/* Vertex clustering algorithm of the stream's vertex finder: the switches of the reference's
 * gpuVertexFinder::Producer (plugin-PixelVertexFinding/gpuVertexFinder.h:46-60, dispatched in
 * gpuVertexFinderImpl.h:96-102); PixelVertexProducerCUDA.cc:26-31 configures density, the default. */
#define PXG_VERTEX_DENSITY 0
#define PXG_VERTEX_DBSCAN 1
#define PXG_VERTEX_ITERATIVE 2
int pxg_set_vertex_algorithm(pxg_stream* s, int algo);
/* Fishbone passes of the stream's ntuplet stage: the reference's earlyFishbone / lateFishbone
 * switches (CAHitNtupletGeneratorKernels.cc:109-112 before the ntuplet search on the connected
 * cells, :133-136 after it on the cells in tuples, with kernel_fishboneCleaner after the
 * classification at :161-164; CAHitNtupletGeneratorOnGPU.cc:55-56 sets early on, late off). */
int pxg_set_fishbone(pxg_stream* s, int early, int late);
/* Helix fit of the stream's ntuplet stage: the reference's useRiemannFit switch
 * (CAHitNtupletGeneratorOnGPU.cc:131-135 and :168-172, RiemannFitOnGPU.cc:3-113; default off,
 * CAHitNtupletGeneratorOnGPU.cc:52): nonzero fits with the Riemann fit, zero with the broken line. */
int pxg_set_riemann_fit(pxg_stream* s, int riemann);
/* Field of the stream's helix fits in 1/GeV, as CAHitNtupletCUDA::produce passes it to makeTuples
 * (CAHitNtupletCUDA.cc:31, default 0.0114256972711507; HelixFitOnGPU keeps it as a float,
 * HelixFitOnGPU.h:43).  Positive and finite, else PXG_ERR_ARG. */
int pxg_set_bfield(pxg_stream* s, float bfield);
/* Track pt cut of the stream's vertex finder (PixelVertexProducerCUDA's m_ptMin,
 * PixelVertexProducerCUDA.cc:37, default 0.5 GeV).  Not negative and finite, else PXG_ERR_ARG. */
int pxg_set_vertex_ptmin(pxg_stream* s, float ptMin);

/* The ntuplet generator's statistics: the reference's doStats switch (default off,
 * CAHitNtupletGeneratorOnGPU.cc:58) and its Counters (CAHitNtupletGeneratorKernels.h:14-26),
 * accumulated over every event the stream processes while on (kernel_checkOverflows,
 * kernel_doStatsForTracks, kernel_doStatsForHitInTracks in CAHitNtupletGeneratorKernelsImpl.h).
 * pxg_print_counters prints them as kernel_printCounters does (:577-608). */
enum {
  PXG_COUNTER_EVENTS = 0,
  PXG_COUNTER_HITS,
  PXG_COUNTER_CELLS,
  PXG_COUNTER_TUPLES,
  PXG_COUNTER_FIT_TRACKS,
  PXG_COUNTER_GOOD_TRACKS,
  PXG_COUNTER_USED_HITS,
  PXG_COUNTER_DUP_HITS,
  PXG_COUNTER_KILLED_CELLS,
  PXG_COUNTER_EMPTY_CELLS,
  PXG_COUNTER_ZERO_TRACK_CELLS,
  PXG_N_COUNTERS
};
int pxg_set_stats(pxg_stream* s, int on);
int pxg_get_counters(pxg_stream* s, uint64_t* out); /* PXG_N_COUNTERS values */
int pxg_reset_counters(pxg_stream* s);
int pxg_print_counters(pxg_stream* s);
int pxg_get_tuning(pxg_stream* s, uint32_t* doublet_lanes, uint32_t* grid_target);
/* With graphs on, pxg_run captures the stage sequence of each (first, count, upto) it is
 * called with as a HIP graph once and replays it afterwards.  Loading events and changing a
 * setting drop the captured graphs; a profiled stream runs eagerly.  No counterpart in the
 * reference, which launches the kernels of a stage one by one. */
int pxg_set_graphs(pxg_stream* s, int on);
/* The launch shape in full (products do not depend on it; no counterpart in the reference). */
typedef struct {
  uint32_t doublet_lanes;        /* lanes per doublet-search slot: power of two, 1..64 */
  uint32_t grid_target;          /* workgroups of a grid-stride launch, shared by its events (>= 4) */
  uint32_t max_blocks_per_event; /* at most this many of them per event (4..4096) */
  uint32_t connect_stage;        /* candidate cells connect stages per wavefront in LDS: 128, 256, 512 */
  uint32_t fishbone_window;      /* list entries the fishbone stages per wavefront: 128, 256, 384 */
} pxg_launch_shape;
int pxg_set_launch_shape(pxg_stream* s, const pxg_launch_shape* shape);
int pxg_get_launch_shape(pxg_stream* s, pxg_launch_shape* shape);
/* Times the chain on resident events [first, first+count) over launch shapes (median of 3
 * timed passes each, after a warm-up pass): doublet lanes x grid target, then, around the
 * best, connect's staging, the fishbone window and the workgroups per event; sets the fastest
 * on the stream and reports it with its time (ms of one pass).  No counterpart in the reference. */
int pxg_autotune_shape(pxg_stream* s, uint32_t first, uint32_t count, int upto, pxg_launch_shape* best,
                       float* best_ms);
/* The same, reporting only the lanes and the grid target. */
int pxg_autotune(pxg_stream* s, uint32_t first, uint32_t count, int upto, uint32_t* best_lanes,
                 uint32_t* best_grid, float* best_ms);

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

// C ABI of libpxgpu.so (include/pxgpu.h): context, per-stream device pools,
// launch sequences and product views.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/pxgpu.h"
#include "pxgpu_internal.h"
#include "pxgpu_stream.h"

using namespace pxg;

static_assert(sizeof(pxg_beamspot) == sizeof(BeamSpotPOD), "beam spot layout");

namespace {
  thread_local std::string g_lastError;
}  // namespace

namespace pxg::detail {
  int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_lastError = buf;
    return code;
  }
}  // namespace pxg::detail

using pxg::detail::fail;

namespace {

  template <typename T>
  hipError_t devAlloc(T*& p, size_t n, std::vector<void*>& owned) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T) + 256);
    if (e == hipSuccess) {
      owned.push_back(q);
      p = static_cast<T*>(q);
    }
    return e;
  }

  uint32_t alignUp(uint32_t v, uint32_t a) { return (v + a - 1) / a * a; }
}  // namespace

namespace pxg::detail {
  // every device array of a pool, with its element count (shared by the
  // allocation and by the memory budget of the serving loop)
  template <typename F>
  void forEachArray(Pool& p, CAPool& ca, uint32_t maxEvents, uint32_t W, F&& f) {
    const size_t E = maxEvents, H = size_t(maxEvents) * kMaxNumClusters, M = size_t(maxEvents) * kMaxModuleRanges;
    f(p.words, W);
    f(p.fedIds, W / 2);
    f(p.events, E);
    f(p.counts, E);
    f(p.xx, W);
    f(p.yy, W);
    f(p.adc, W);
    f(p.moduleInd, W);
    f(p.clus, W);
    f(p.pdigi, W);
    f(p.rawIdArr, W);
    f(p.boundaryMask, 2 * (W / 64));
    f(p.errorsSparse, W);
    f(p.errors, W);
    f(p.clExt, W);
    f(p.bigList, M);
    f(p.moduleStart, M);
    f(p.clusInModule, M);
    f(p.moduleId, M);
    f(p.clusModuleStart, M);
    f(p.xl, H);
    f(p.yl, H);
    f(p.xerr, H);
    f(p.yerr, H);
    f(p.xg, H);
    f(p.yg, H);
    f(p.zg, H);
    f(p.rg, H);
    f(p.pos, H);
    f(p.attr, H);
    f(p.candPhi, H);
    f(p.iphi, H);
    f(p.charge, H);
    f(p.xsize, H);
    f(p.ysize, H);
    f(p.detInd, H);
    f(p.hitsLayerStart, E * kLayerStartStride);
    f(p.phiOff, E * kPhiOffStride);
    f(p.phiBins, H);
    f(p.avg, E);
    const size_t D = E * ca.caps.maxDoublets, T = E * ca.caps.maxTuples, HS = E * kHitStride;
    const size_t L = E * ca.caps.maxCellLists;
    f(ca.cnt, E);
    f(ca.dOff, E * (kDoubletSlots + 64));
    f(ca.dMask, E * 2 * kDoubletSlots);
    f(ca.cellInner, D);
    f(ca.cellOuter, D);
    f(ca.cellPair, D);
    f(ca.cellFlags, D);
    f(ca.cellUsed, D);
    f(ca.ohCount, HS);
    f(ca.ohOff, HS);
    f(ca.ohCells, D);
    f(ca.startCells, D);
    f(ca.nbSlot, D);
    f(ca.ohRank, D);
    f(ca.nbCount, D);
    f(ca.nbPool, L * kCaMaxNeighbors);
    f(ca.startMask, E * (ca.caps.maxDoublets / 64 + 64));
    f(ca.dfsT, D);
    f(ca.dfsH, D);
    f(ca.trSlot, D);
    f(ca.trCount, D);
    f(ca.trPool, L * kCaMaxCellTracks);
    f(ca.multiCells, D);
    f(ca.dupPreds, L * kDupPredCap);
    f(ca.dupCount, L);
    f(ca.tupleOff, E * (ca.caps.maxTuples + 64));
    f(ca.tupleBins, 5 * T);
    f(ca.tupleCells, 4 * T);
    f(ca.tuplePaths, T);
    f(ca.quality, T);
    f(ca.multOff, 16 * E);
    f(ca.multBins, E * ca.caps.maxTuplesMult);
    f(ca.detIndices, 5 * T);
    f(ca.state, 5 * T);
    f(ca.cov, 15 * T);
    f(ca.pt, T);
    f(ca.eta, T);
    f(ca.chi2, T);
    f(ca.h2tOff, HS);
    f(ca.h2tBins, E * ca.caps.maxHitToTuple);
    f(ca.idv, T);
    f(ca.ndof, T);
    f(ca.zv, E * kMaxVertices);
    f(ca.wv, E * kMaxVertices);
    f(ca.vchi2, E * kMaxVertices);
    f(ca.ptv2, E * kMaxVertices);
    f(ca.sortInd, E * kMaxVertices);
    f(ca.vItrk, T);
    f(ca.vZt, T);
    f(ca.vEzt2, T);
    f(ca.vPtt2, T);
    f(ca.vIv, T);
    f(ca.vTmp, T);
    f(ca.vZs, T);
    f(ca.vEs, T);
    f(ca.vNnS, T);
    f(ca.vIzt, T);
    f(ca.vTrk, T);
    f(ca.vOff, E * (kMaxVertices + 64));
  }

  size_t poolBytes(uint32_t maxEvents, uint32_t maxWords, const CACapacities& caps) {
    Pool p{};
    CAPool ca{};
    ca.caps = caps;
    const uint32_t W = alignUp(maxWords, kWordAlign) + kWordAlign * maxEvents;
    size_t total = 0;
    forEachArray(p, ca, maxEvents, W, [&](auto* ptr, size_t n) { total += n * sizeof(*ptr) + 256; });
    return total;
  }

  // per-event arrays of the pool shifted to event `first`
  Pool viewFrom(const Pool& p, uint32_t first) {
    Pool v = p;
    v.events = p.events + first;
    v.counts = p.counts + first;
    const size_t m = size_t(first) * kMaxModuleRanges;
    v.moduleStart = p.moduleStart + m;
    v.clusInModule = p.clusInModule + m;
    v.moduleId = p.moduleId + m;
    v.clusModuleStart = p.clusModuleStart + m;
    v.bigList = p.bigList + m;
    const size_t h = size_t(first) * kMaxNumClusters;
    v.xl = p.xl + h;
    v.yl = p.yl + h;
    v.xerr = p.xerr + h;
    v.yerr = p.yerr + h;
    v.xg = p.xg + h;
    v.yg = p.yg + h;
    v.zg = p.zg + h;
    v.rg = p.rg + h;
    v.iphi = p.iphi + h;
    v.charge = p.charge + h;
    v.xsize = p.xsize + h;
    v.ysize = p.ysize + h;
    v.detInd = p.detInd + h;
    v.phiBins = p.phiBins + h;
    v.pos = p.pos + h;
    v.attr = p.attr + h;
    v.candPhi = p.candPhi + h;
    v.hitsLayerStart = p.hitsLayerStart + size_t(first) * kLayerStartStride;
    v.phiOff = p.phiOff + size_t(first) * kPhiOffStride;
    v.avg = p.avg + first;
    return v;
  }

  // ntuplet arrays of the pool shifted to event `first`
  CAPool caViewFrom(const CAPool& c, uint32_t first) {
    CAPool v = c;
    const size_t E = first, D = size_t(first) * c.caps.maxDoublets, T = size_t(first) * c.caps.maxTuples;
    const size_t H = size_t(first) * kHitStride, L = size_t(first) * c.caps.maxCellLists;
    v.cnt += E;
    v.dOff += size_t(first) * (kDoubletSlots + 64);
    v.dMask += size_t(first) * 2 * kDoubletSlots;
    v.cellInner += D;
    v.cellOuter += D;
    v.cellPair += D;
    v.cellFlags += D;
    v.cellUsed += D;
    v.ohCount += H;
    v.ohOff += H;
    v.ohCells += D;
    v.startCells += D;
    v.nbSlot += D;
    v.ohRank += D;
    v.nbCount += D;
    v.nbPool += L * kCaMaxNeighbors;
    v.startMask += size_t(first) * (c.caps.maxDoublets / 64 + 64);
    v.dfsT += D;
    v.dfsH += D;
    v.trSlot += D;
    v.trCount += D;
    v.trPool += L * kCaMaxCellTracks;
    v.multiCells += D;
    v.dupPreds += L * kDupPredCap;
    v.dupCount += L;
    v.tupleOff += c.offBase(first);
    v.tupleBins += 5 * T;
    v.tupleCells += 4 * T;
    v.tuplePaths += T;
    v.quality += T;
    v.multOff += 16 * E;
    v.multBins += size_t(first) * c.caps.maxTuplesMult;
    v.detIndices += 5 * T;
    v.state += 5 * T;
    v.cov += 15 * T;
    v.pt += T;
    v.eta += T;
    v.chi2 += T;
    v.h2tOff += H;
    v.h2tBins += size_t(first) * c.caps.maxHitToTuple;
    v.idv += T;
    v.ndof += T;
    v.zv += E * kMaxVertices;
    v.wv += E * kMaxVertices;
    v.vchi2 += E * kMaxVertices;
    v.ptv2 += E * kMaxVertices;
    v.sortInd += E * kMaxVertices;
    v.vItrk += T;
    v.vZt += T;
    v.vEzt2 += T;
    v.vPtt2 += T;
    v.vIv += T;
    v.vTmp += T;
    v.vZs += T;
    v.vEs += T;
    v.vNnS += T;
    v.vIzt += T;
    v.vTrk += T;
    v.vOff += E * (kMaxVertices + 64);
    return v;
  }

  void timerBegin(void* ctx, const char* name) {
    auto* s = static_cast<pxg_stream*>(ctx);
    ProfileMark m{name, nullptr, nullptr};
    if (hipEventCreate(&m.start) != hipSuccess || hipEventCreate(&m.stop) != hipSuccess)
      return;
    (void)hipEventRecord(m.start, s->stream);
    s->marks.push_back(m);
  }

  void timerEnd(void* ctx) {
    auto* s = static_cast<pxg_stream*>(ctx);
    if (!s->marks.empty())
      (void)hipEventRecord(s->marks.back().stop, s->stream);
  }

  template <typename F>
  int launchTimed(pxg_stream* s, const char* name, F&& f) {
    if (!s->profiling) {
      f();
    } else {
      ProfileMark m{name, nullptr, nullptr};
      PXG_HIP(hipEventCreate(&m.start));
      PXG_HIP(hipEventCreate(&m.stop));
      PXG_HIP(hipEventRecord(m.start, s->stream));
      f();
      PXG_HIP(hipEventRecord(m.stop, s->stream));
      s->marks.push_back(m);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return fail(PXG_ERR_HIP, "launch of %s failed: %s", name, hipGetErrorString(e));
    return 0;
  }

  void markStages(pxg_stream* s, uint32_t first, uint32_t count, int from, int upto) {
    for (uint32_t e = first; e < first + count && e < s->stagesRun.size(); ++e)
      s->stagesRun[e] = uint8_t(from <= 1 ? upto : std::max<int>(s->stagesRun[e], upto));
  }

  int runStages(pxg_stream* s, uint32_t first, uint32_t count, int from, int upto) {
    if (count == 0 || first + count > s->nLoaded)
      return fail(PXG_ERR_ARG, "events [%u, %u) not loaded (%u resident)", first, first + count, s->nLoaded);
    markStages(s, first, count, from, upto);
    const Pool v = viewFrom(s->pool, first);
    const DevConditions& c = s->ctx->cond;
    uint32_t maxWords = 0;  // the longest event of the launch sets the decode grid's width
    for (uint32_t e = first; e < first + count; ++e)
      maxWords = std::max(maxWords, s->hostEvents[e].nWords);
    int rc = 0;
    if (from <= 1 && upto >= 1) {
      if ((rc = launchTimed(s, "decode", [&] { launchDecode(c, v, count, maxWords, s->stream); })))
        return rc;
      if ((rc = launchTimed(s, "moduleStarts", [&] { launchModuleStarts(v, count, s->stream); })))
        return rc;
    }
    if (from <= 2 && upto >= 2) {
      if ((rc = launchTimed(s, "clusterize", [&] { launchClusterize(v, count, s->stream); })))
        return rc;
    }
    if (from <= 3 && upto >= 3) {
      if ((rc = launchTimed(s, "hits", [&] { launchHits(c, v, count, s->stream); })))
        return rc;
      if ((rc = launchTimed(s, "phiBinner", [&] { launchPhiBinner(c, v, count, s->stream); })))
        return rc;
    }
    const CAPool ca = caViewFrom(s->ca, first);
    Timer timer;
    if (s->profiling)
      timer = {timerBegin, timerEnd, s};
    const CALaunch l{&c,        &v,     &ca, count, std::max(4u, std::min(s->maxBlocksPerEvent, s->gridTarget / count)),
                     s->stream, &timer, s->doubletLanes, s->vertexAlgo, s->earlyFishbone, s->lateFishbone,
                     s->riemannFit, s->doStats ? s->counters : nullptr, double(s->bField), s->vertexPtMin,
                     s->connectStage, s->fishboneWindow};
    if (from <= 4 && upto >= 4) {
      launchDoublets(l);
      launchConnect(l);
      if (l.earlyFishbone)
        launchFishbone(l, false);
      launchNtuplets(l);
      launchMultiplicity(l);
      if (l.lateFishbone)
        launchFishbone(l, true);
      if (l.riemannFit)
        launchRiemannFit(l);
      else
        launchFit(l);
      launchClassify(l);
      launchHitToTuple(l);
    }
    if (from <= 5 && upto >= 5)
      launchVertices(l);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return fail(PXG_ERR_HIP, "launch of the ntuplet or vertex stage failed: %s", hipGetErrorString(e));
    return 0;
  }

  void fillTracks(pxg_stream* s, uint32_t e, pxg_tracks* t, pxg_vertices* vx) {
    const CAPool ca = caViewFrom(s->ca, e);
    const CACounts& c = s->hostCA[e];
    if (t) {
      *t = {ca.tupleOff, ca.tupleBins, ca.detIndices, ca.quality, ca.chi2, ca.eta, ca.pt, ca.state, ca.cov,
            ca.caps.maxTuples, c.nTuples};
    }
    if (vx)
      *vx = {ca.idv, ca.zv, ca.wv, ca.vchi2, ca.ptv2, ca.ndof, ca.sortInd, c.nVertices};
  }

  void fillProducts(pxg_stream* s, uint32_t e, pxg_digis* d, pxg_clusters* cl, pxg_errors* er, pxg_hits* h) {
    const Pool& p = s->pool;
    const EventDesc& ev = s->hostEvents[e];
    const EventCounts& cnt = s->hostCounts[e];
    const size_t w = ev.wordOffset;
    const size_t m = size_t(e) * kMaxModuleRanges;
    const size_t ho = size_t(e) * kMaxNumClusters;
    if (d) {
      *d = {p.xx + w, p.yy + w, p.adc + w, p.moduleInd + w, p.clus + w, p.pdigi + w, p.rawIdArr + w, cnt.nModules,
            ev.nWords};
    }
    if (cl)
      *cl = {p.moduleStart + m, p.clusInModule + m, p.moduleId + m, p.clusModuleStart + m, cnt.nClusters};
    if (er)
      *er = {p.errors + w, cnt.nErrors};
    if (h) {
      *h = {p.xl + ho,
            p.yl + ho,
            p.xerr + ho,
            p.yerr + ho,
            p.xg + ho,
            p.yg + ho,
            p.zg + ho,
            p.rg + ho,
            p.iphi + ho,
            p.charge + ho,
            p.xsize + ho,
            p.ysize + ho,
            p.detInd + ho,
            p.clusModuleStart + m,
            p.hitsLayerStart + size_t(e) * kLayerStartStride,
            p.phiOff + size_t(e) * kPhiOffStride,
            p.phiBins + ho,
            reinterpret_cast<float*>(p.avg + e),
            cnt.nClusters};
    }
  }
}  // namespace pxg::detail

using namespace pxg::detail;

void pxg::detail::dropGraphs(pxg_stream* s) {
  for (auto& g : s->graphs)
    (void)hipGraphExecDestroy(g.second);
  s->graphs.clear();
}

namespace {
  // Validates a batch and writes its descriptors (256-word aligned offsets) to the
  // stream's pinned descriptor array.
  int describeEvents(pxg_stream* s, uint32_t n, const pxg_raw_view* raws, const pxg_beamspot* bs,
                     uint32_t nBeamspots) {
    if (!s || !raws || !bs || n == 0 || (nBeamspots != 1 && nBeamspots != n))
      return fail(PXG_ERR_ARG, "pxg_load_events: bad arguments");
    dropGraphs(s);
    if (n > s->pool.maxEvents)
      return fail(PXG_ERR_CAPACITY, "pxg_load_events: %u events > capacity %u", n, s->pool.maxEvents);
    uint32_t off = 0, blocks = 0;
    for (uint32_t e = 0; e < n; ++e) {
      if (raws[e].nWords % 2)
        return fail(PXG_ERR_ARG, "event %u: odd number of words %u", e, raws[e].nWords);
      if (raws[e].nWords > kMaxFedWords)
        return fail(PXG_ERR_CAPACITY, "event %u: %u words > %u", e, raws[e].nWords, kMaxFedWords);
      if (uint64_t(off) + raws[e].nWords > s->wordCapacity)
        return fail(PXG_ERR_CAPACITY, "pxg_load_events: %u words exceed the pool", off + raws[e].nWords);
      EventDesc& d = s->hostEvents[e];
      d.wordOffset = off;
      d.nWords = raws[e].nWords;
      d.blockOffset = blocks;
      d.pad = 0;
      std::memcpy(&d.bs, bs + (nBeamspots == 1 ? 0 : e), sizeof(BeamSpotPOD));
      blocks += (raws[e].nWords + kDecodeBlock - 1) / kDecodeBlock;
      off = alignUp(off + raws[e].nWords, kWordAlign);
    }
    return 0;
  }

  void markLoaded(pxg_stream* s, uint32_t n) {
    s->nLoaded = n;
    std::fill(s->stagesRun.begin(), s->stagesRun.end(), uint8_t(0));
  }

  // event blockIdx.y of the table: its words and FED ids from host memory into the pool
  __global__ void __launch_bounds__(256) gatherKernel(const GatherSrc* table, uint32_t* words, uint8_t* fedIds) {
    const GatherSrc g = table[blockIdx.y];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g.nWords; i += stride)
      words[g.wordOffset + i] = g.words[i];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g.nWords / 2; i += stride)
      fedIds[g.wordOffset / 2 + i] = g.fedIds[i];
  }
}  // namespace

int pxg::detail::loadEvents(pxg_stream* s, uint32_t n, const pxg_raw_view* raws, const pxg_beamspot* bs,
                             uint32_t nBeamspots, bool sync) {
  int rc = describeEvents(s, n, raws, bs, nBeamspots);
  if (rc)
    return rc;
  for (uint32_t e = 0; e < n; ++e) {
    const EventDesc& d = s->hostEvents[e];
    if (raws[e].nWords) {
      PXG_HIP(hipMemcpyAsync(s->pool.words + d.wordOffset, raws[e].words, size_t(raws[e].nWords) * 4,
                             hipMemcpyHostToDevice, s->stream));
      PXG_HIP(hipMemcpyAsync(s->pool.fedIds + d.wordOffset / 2, raws[e].fedIds, raws[e].nWords / 2,
                             hipMemcpyHostToDevice, s->stream));
    }
  }
  PXG_HIP(hipMemcpyAsync(s->pool.events, s->hostEvents, n * sizeof(EventDesc), hipMemcpyHostToDevice, s->stream));
  if (sync)
    PXG_HIP(hipStreamSynchronize(s->stream));
  markLoaded(s, n);
  return 0;
}

int pxg::detail::loadEventsGather(pxg_stream* s, uint32_t n, const pxg_raw_view* raws, const pxg_beamspot* bs,
                                   GatherSrc* table, const GatherSrc* tableDev) {
  int rc = describeEvents(s, n, raws, bs, n);
  if (rc)
    return rc;
  uint32_t maxWords = 0;
  for (uint32_t e = 0; e < n; ++e) {
    table[e] = {raws[e].words, raws[e].fedIds, raws[e].nWords, s->hostEvents[e].wordOffset};
    maxWords = std::max(maxWords, raws[e].nWords);
  }
  PXG_HIP(hipMemcpyAsync(s->pool.events, s->hostEvents, n * sizeof(EventDesc), hipMemcpyHostToDevice, s->stream));
  // about 16 words per thread for the largest event
  const uint32_t blocksPerEvent = std::max(1u, std::min(256u, (maxWords + 4095) / 4096));
  gatherKernel<<<dim3(blocksPerEvent, n), 256, 0, s->stream>>>(tableDev, s->pool.words, s->pool.fedIds);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(PXG_ERR_HIP, "launch of the event gather failed: %s", hipGetErrorString(e));
  markLoaded(s, n);
  return 0;
}

extern "C" {

const char* pxg_last_error(void) { return g_lastError.c_str(); }

int pxg_device_name(char* buf, size_t n) {
  int dev = 0;
  PXG_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  PXG_HIP(hipGetDeviceProperties(&prop, dev));
  std::snprintf(buf, n, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return 0;
}

int pxg_init(const pxg_conditions* c, pxg_context** out) {
  if (!c || !out || !c->cabling || !c->mod_to_unp || !c->gain || !c->gain_data || !c->cpe_common || !c->cpe_det ||
      !c->cpe_layer || !c->cpe_average)
    return fail(PXG_ERR_ARG, "pxg_init: missing conditions");
  if (c->n_det != kNumModules)
    return fail(PXG_ERR_ARG, "pxg_init: %u CPE det params, expected %u", c->n_det, kNumModules);
  auto* ctx = new pxg_context;
  auto up = [&](const void* src, size_t bytes, void*& dst) -> hipError_t {
    hipError_t e = hipMalloc(&dst, bytes);
    if (e != hipSuccess)
      return e;
    ctx->owned.push_back(dst);
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  };
  void *cab, *m2u, *gain, *gdata, *com, *det, *lay, *avg;
  hipError_t e = hipGetDevice(&ctx->device);
  if (e == hipSuccess)
    e = up(c->cabling, sizeof(SiPixelFedCablingMapGPU), cab);
  if (e == hipSuccess)
    e = up(c->mod_to_unp, kCablingSize, m2u);
  if (e == hipSuccess)
    e = up(c->gain, sizeof(SiPixelGainForHLTonGPU), gain);
  if (e == hipSuccess)
    e = up(c->gain_data, c->gain_data_bytes, gdata);
  if (e == hipSuccess)
    e = up(c->cpe_common, sizeof(CPECommonParams), com);
  if (e == hipSuccess)
    e = up(c->cpe_det, sizeof(CPEDetParams) * c->n_det, det);
  if (e == hipSuccess)
    e = up(c->cpe_layer, sizeof(CPELayerGeometry), lay);
  if (e == hipSuccess)
    e = up(c->cpe_average, sizeof(AverageGeometry), avg);
  if (e != hipSuccess) {
    pxg_finalize(ctx);
    return fail(PXG_ERR_HIP, "pxg_init: %s", hipGetErrorString(e));
  }
  ctx->cond = {static_cast<const SiPixelFedCablingMapGPU*>(cab), static_cast<const uint8_t*>(m2u),
               static_cast<const SiPixelGainForHLTonGPU*>(gain), static_cast<const GainDecoding*>(gdata),
               c->gain_data_bytes,
               static_cast<const CPECommonParams*>(com),      static_cast<const CPEDetParams*>(det),
               static_cast<const CPELayerGeometry*>(lay),     static_cast<const AverageGeometry*>(avg)};
  const int rc = noallocSetup(ctx->device);
  if (rc) {
    pxg_finalize(ctx);
    return rc;
  }
  *out = ctx;
  return 0;
}

void pxg_finalize(pxg_context* ctx) {
  if (!ctx)
    return;
  for (void* p : ctx->owned)
    (void)hipFree(p);
  delete ctx;
}

int pxg_stream_create(pxg_context* ctx, uint32_t maxEvents, uint32_t maxWords, pxg_stream** out) {
  return pxg_stream_create_profile(ctx, maxEvents, maxWords, PXG_PROFILE_REFERENCE, out);
}

int pxg_stream_create_profile(pxg_context* ctx, uint32_t maxEvents, uint32_t maxWords, int profile,
                              pxg_stream** out) {
  if (!ctx || !out || maxEvents == 0 || maxWords == 0 || maxEvents > 65535)
    return fail(PXG_ERR_ARG, "pxg_stream_create: bad arguments");
  if (profile != PXG_PROFILE_REFERENCE && profile != PXG_PROFILE_HIGH_OCCUPANCY)
    return fail(PXG_ERR_ARG, "pxg_stream_create: unknown capacity profile %d", profile);
  auto* s = new pxg_stream;
  s->ctx = ctx;
  const uint32_t W = alignUp(maxWords, kWordAlign) + kWordAlign * maxEvents;
  s->wordCapacity = W;
  Pool& p = s->pool;
  p.maxEvents = maxEvents;
  p.maxWords = W;
  const size_t E = maxEvents, H = size_t(maxEvents) * kMaxNumClusters, M = size_t(maxEvents) * kMaxModuleRanges;
  auto& o = s->owned;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  CAPool& ca = s->ca;
  ca.caps = profile == PXG_PROFILE_HIGH_OCCUPANCY ? kCapsHighOccupancy : kCapsReference;
  forEachArray(p, ca, maxEvents, W, [&](auto*& ptr, size_t n) {
    if (e == hipSuccess)
      e = devAlloc(ptr, n, o);
  });
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void**>(&s->hostEvents), E * sizeof(EventDesc), hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void**>(&s->hostCounts), E * sizeof(EventCounts), hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void**>(&s->hostCA), E * sizeof(CACounts), hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipMemsetAsync(p.counts, 0, E * sizeof(EventCounts), s->stream);
  if (e == hipSuccess)
    e = hipMemsetAsync(ca.cnt, 0, E * sizeof(CACounts), s->stream);
  if (e == hipSuccess)
    e = devAlloc(s->counters, PXG_N_COUNTERS, o);
  if (e == hipSuccess)
    e = hipMemsetAsync(s->counters, 0, PXG_N_COUNTERS * sizeof(unsigned long long), s->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    pxg_stream_destroy(s);
    return fail(PXG_ERR_HIP, "pxg_stream_create: %s", hipGetErrorString(e));
  }
  std::memset(s->hostCounts, 0, E * sizeof(EventCounts));
  std::memset(s->hostCA, 0, E * sizeof(CACounts));
  s->stagesRun.assign(E, 0);
  *out = s;
  return 0;
}

void pxg_stream_destroy(pxg_stream* s) {
  if (!s)
    return;
  if (s->stream)
    (void)hipStreamSynchronize(s->stream);
  dropGraphs(s);
  for (auto& m : s->marks) {
    (void)hipEventDestroy(m.start);
    (void)hipEventDestroy(m.stop);
  }
  for (void* p : s->owned)
    (void)hipFree(p);
  if (s->hostEvents)
    (void)hipHostFree(s->hostEvents);
  if (s->hostCounts)
    (void)hipHostFree(s->hostCounts);
  if (s->hostCA)
    (void)hipHostFree(s->hostCA);
  if (s->stream)
    (void)hipStreamDestroy(s->stream);
  delete s;
}

int pxg_sync(pxg_stream* s) {
  if (!s)
    return fail(PXG_ERR_ARG, "pxg_sync: null stream");
  PXG_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

int pxg_load_events(pxg_stream* s, uint32_t n, const pxg_raw_view* raws, const pxg_beamspot* bs) {
  return loadEvents(s, n, raws, bs, 1, true);
}

int pxg_run(pxg_stream* s, uint32_t first, uint32_t count, int upto) {
  if (!s || upto < 1 || upto > 5)
    return fail(PXG_ERR_ARG, "pxg_run: bad arguments");
  if (!s->useGraphs || s->profiling)
    return runStages(s, first, count, 1, upto);
  if (count == 0 || first + count > s->nLoaded)
    return fail(PXG_ERR_ARG, "events [%u, %u) not loaded (%u resident)", first, first + count, s->nLoaded);
  const auto key = std::make_tuple(first, count, upto);
  auto it = s->graphs.find(key);
  if (it == s->graphs.end()) {
    // the sequence is kernels only (counters are cleared by kernels, not memsets)
    PXG_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int rc = runStages(s, first, count, 1, upto);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (rc) {
      if (g)
        (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess)
      return fail(PXG_ERR_HIP, "pxg_run: stage sequence capture failed: %s", hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
      return fail(PXG_ERR_HIP, "pxg_run: graph instantiation failed: %s", hipGetErrorString(e));
    it = s->graphs.emplace(key, x).first;
  }
  PXG_HIP(hipGraphLaunch(it->second, s->stream));
  markStages(s, first, count, 1, upto);
  return 0;
}

int pxg_set_graphs(pxg_stream* s, int on) {
  if (!s)
    return fail(PXG_ERR_ARG, "pxg_set_graphs: null stream");
  s->useGraphs = on != 0;
  dropGraphs(s);
  return 0;
}

int pxg_fetch_counts(pxg_stream* s, uint32_t first, uint32_t count) {
  if (!s || first + count > s->nLoaded)
    return fail(PXG_ERR_ARG, "pxg_fetch_counts: bad range");
  PXG_HIP(hipMemcpyAsync(s->hostCounts + first, s->pool.counts + first, count * sizeof(EventCounts),
                         hipMemcpyDeviceToHost, s->stream));
  PXG_HIP(hipMemcpyAsync(s->hostCA + first, s->ca.cnt + first, count * sizeof(CACounts), hipMemcpyDeviceToHost,
                         s->stream));
  PXG_HIP(hipStreamSynchronize(s->stream));
  uint32_t st = 0;
  for (uint32_t e = first; e < first + count; ++e) {
    st |= s->hostCounts[e].status;
    if (s->stagesRun[e] >= 4)
      st |= s->hostCA[e].status << PXG_STATUS_CA_SHIFT;
    if (s->stagesRun[e] >= 5 && s->hostCA[e].vstatus)
      st |= uint32_t(PXG_STATUS_VERTEX_OVERFLOW);
  }
  return int(st);
}

int pxg_event_products(pxg_stream* s, uint32_t e, pxg_digis* d, pxg_clusters* c, pxg_errors* er, pxg_hits* h) {
  if (!s || e >= s->nLoaded)
    return fail(PXG_ERR_ARG, "pxg_event_products: event %u not resident", e);
  fillProducts(s, e, d, c, er, h);
  return 0;
}

int pxg_event_counts(pxg_stream* s, uint32_t e, uint32_t out[4]) {
  if (!s || !out || e >= s->nLoaded)
    return fail(PXG_ERR_ARG, "pxg_event_counts: event %u not resident", e);
  const EventCounts& c = s->hostCounts[e];
  out[0] = c.nModules;
  out[1] = c.nClusters;
  out[2] = c.nErrors;
  out[3] = c.status;
  return 0;
}

int pxg_event_tracks(pxg_stream* s, uint32_t e, pxg_tracks* t, pxg_vertices* v) {
  if (!s || e >= s->nLoaded)
    return fail(PXG_ERR_ARG, "pxg_event_tracks: event %u not resident", e);
  fillTracks(s, e, t, v);
  return 0;
}

int pxg_event_ca_internals(pxg_stream* s, uint32_t e, pxg_ca_internals* o) {
  if (!s || !o || e >= s->nLoaded)
    return fail(PXG_ERR_ARG, "pxg_event_ca_internals: event %u not resident", e);
  const CAPool ca = caViewFrom(s->ca, e);
  *o = {};
  o->cellInner = ca.cellInner;
  o->cellOuter = ca.cellOuter;
  o->cellPair = ca.cellPair;
  o->cellFlags = ca.cellFlags;
  o->cellUsed = ca.cellUsed;
  o->outerHitOff = ca.ohOff;
  o->outerHitCells = reinterpret_cast<uint32_t*>(ca.ohCells);
  o->neighborSlot = ca.nbSlot;
  o->neighborCount = ca.nbCount;
  o->neighborPool = ca.nbPool;
  o->trackSlot = ca.trSlot;
  o->trackCount = ca.trCount;
  o->trackPool = ca.trPool;
  o->multOff = ca.multOff;
  o->multBins = ca.multBins;
  o->hitToTupleOff = ca.h2tOff;
  o->hitToTupleBins = ca.h2tBins;
  static_assert(sizeof(CACounts) <= sizeof(o->counts), "counts");
  std::memset(o->counts, 0, sizeof(o->counts));
  std::memcpy(o->counts, &s->hostCA[e], sizeof(CACounts));
  return 0;
}

int pxg_copy_to_host(pxg_stream* s, void* dst, const void* src, size_t bytes) {
  if (!s || (!dst && bytes) || (!src && bytes))
    return fail(PXG_ERR_ARG, "pxg_copy_to_host: bad arguments");
  if (bytes) {
    PXG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream));
    PXG_HIP(hipStreamSynchronize(s->stream));
  }
  return 0;
}

// ---- plain per-event entry points ---------------------------------------------------
static int plainEvent(pxg_stream* s, const uint32_t* words, const uint8_t* fedIds, uint32_t nWords, int upto) {
  if (!s || (nWords && (!words || !fedIds)))
    return fail(PXG_ERR_ARG, "null input");
  pxg_raw_view v{words, fedIds, nWords};
  pxg_beamspot bs{};
  int rc = pxg_load_events(s, 1, &v, &bs);
  if (rc)
    return rc;
  rc = runStages(s, 0, 1, 1, upto);
  if (rc)
    return rc;
  return pxg_fetch_counts(s, 0, 1);
}

int pxg_raw_to_digi(pxg_stream* s, const uint32_t* words, const uint8_t* fedIds, uint32_t nWords, pxg_digis* d,
                    pxg_errors* er) {
  int rc = plainEvent(s, words, fedIds, nWords, 1);
  if (rc < 0)
    return rc;
  fillProducts(s, 0, d, nullptr, er, nullptr);
  return rc;
}

int pxg_raw_to_cluster(pxg_stream* s, const uint32_t* words, const uint8_t* fedIds, uint32_t nWords, pxg_digis* d,
                       pxg_clusters* c, pxg_errors* er) {
  int rc = plainEvent(s, words, fedIds, nWords, 2);
  if (rc < 0)
    return rc;
  fillProducts(s, 0, d, c, er, nullptr);
  return rc;
}

int pxg_rechits(pxg_stream* s, const pxg_digis* d, const pxg_clusters* c, const pxg_beamspot* bs, pxg_hits* h) {
  if (!s || !d || !c || !bs || !h)
    return fail(PXG_ERR_ARG, "pxg_rechits: null argument");
  if (s->nLoaded < 1 || d->xx != s->pool.xx + s->hostEvents[0].wordOffset || c->moduleStart != s->pool.moduleStart)
    return fail(PXG_ERR_ARG, "pxg_rechits: digis/clusters are not the last products of this stream");
  std::memcpy(&s->hostEvents[0].bs, bs, sizeof(BeamSpotPOD));
  PXG_HIP(hipMemcpyAsync(s->pool.events, s->hostEvents, sizeof(EventDesc), hipMemcpyHostToDevice, s->stream));
  int rc = runStages(s, 0, 1, 3, 3);
  if (rc)
    return rc;
  rc = pxg_fetch_counts(s, 0, 1);
  if (rc < 0)
    return rc;
  fillProducts(s, 0, nullptr, nullptr, nullptr, h);
  return rc;
}

int pxg_ntuplets(pxg_stream* s, const pxg_hits* h, float bfield, pxg_tracks* t) {
  if (!s || !h || !t)
    return fail(PXG_ERR_ARG, "pxg_ntuplets: null argument");
  if (s->nLoaded < 1 || h->xg != s->pool.xg)
    return fail(PXG_ERR_ARG, "pxg_ntuplets: hits are not the last products of this stream");
  int rc = pxg_set_bfield(s, bfield);
  if (rc)
    return rc;
  rc = runStages(s, 0, 1, 4, 4);
  if (rc)
    return rc;
  rc = pxg_fetch_counts(s, 0, 1);
  if (rc < 0)
    return rc;
  fillTracks(s, 0, t, nullptr);
  return rc;
}

int pxg_find_vertices(pxg_stream* s, const pxg_tracks* t, float ptMin, pxg_vertices* v) {
  if (!s || !t || !v)
    return fail(PXG_ERR_ARG, "pxg_find_vertices: null argument");
  if (s->nLoaded < 1 || t->hitOff != s->ca.tupleOff)
    return fail(PXG_ERR_ARG, "pxg_find_vertices: tracks are not the last products of this stream");
  int rc = pxg_set_vertex_ptmin(s, ptMin);
  if (rc)
    return rc;
  rc = runStages(s, 0, 1, 5, 5);
  if (rc)
    return rc;
  rc = pxg_fetch_counts(s, 0, 1);
  if (rc < 0)
    return rc;
  fillTracks(s, 0, nullptr, v);
  return rc;
}

size_t pxg_pool_bytes(uint32_t maxEvents, uint32_t maxWords, int profile) {
  return poolBytes(maxEvents, maxWords, profile == PXG_PROFILE_HIGH_OCCUPANCY ? kCapsHighOccupancy : kCapsReference);
}

int pxg_profile_capacities(int profile, pxg_capacities* out) {
  if (!out || (profile != PXG_PROFILE_REFERENCE && profile != PXG_PROFILE_HIGH_OCCUPANCY))
    return fail(PXG_ERR_ARG, "pxg_profile_capacities: unknown profile or null output");
  const CACapacities& c = profile == PXG_PROFILE_HIGH_OCCUPANCY ? kCapsHighOccupancy : kCapsReference;
  *out = {c.maxDoublets, c.maxTuples, c.maxTuplesMult, c.maxHitToTuple, c.maxCellLists, c.maxCellsPerHit};
  return 0;
}

// ---- tuning ---------------------------------------------------------------------------
int pxg_set_tuning(pxg_stream* s, uint32_t doubletLanes, uint32_t gridTarget) {
  if (!s || doubletLanes == 0 || doubletLanes > 64 || (doubletLanes & (doubletLanes - 1)) || gridTarget < 4)
    return fail(PXG_ERR_ARG, "pxg_set_tuning: lanes must be a power of two in [1, 64], grid target >= 4");
  s->doubletLanes = doubletLanes;
  s->gridTarget = gridTarget;
  dropGraphs(s);
  return 0;
}

int pxg_set_vertex_algorithm(pxg_stream* s, int algo) {
  if (!s)
    return fail(PXG_ERR_ARG, "pxg_set_vertex_algorithm: null stream");
  if (algo != PXG_VERTEX_DENSITY && algo != PXG_VERTEX_DBSCAN && algo != PXG_VERTEX_ITERATIVE)
    return fail(PXG_ERR_ARG, "pxg_set_vertex_algorithm: unknown algorithm %d", algo);
  static_assert(PXG_VERTEX_DENSITY == kVertexDensity && PXG_VERTEX_DBSCAN == kVertexDBSCAN &&
                PXG_VERTEX_ITERATIVE == kVertexIterative);
  s->vertexAlgo = algo;
  dropGraphs(s);
  return 0;
}

int pxg_set_bfield(pxg_stream* s, float bfield) {
  if (!s || !(bfield > 0.f) || !(bfield < 1e30f))
    return fail(PXG_ERR_ARG, "pxg_set_bfield: null stream or field not positive and finite");
  if (bfield != s->bField) {
    s->bField = bfield;
    dropGraphs(s);
  }
  return 0;
}

int pxg_set_vertex_ptmin(pxg_stream* s, float ptMin) {
  if (!s || !(ptMin >= 0.f) || !(ptMin < 1e30f))
    return fail(PXG_ERR_ARG, "pxg_set_vertex_ptmin: null stream or pt cut negative or not finite");
  if (ptMin != s->vertexPtMin) {
    s->vertexPtMin = ptMin;
    dropGraphs(s);
  }
  return 0;
}

int pxg_set_fishbone(pxg_stream* s, int early, int late) {
  if (!s)
    return fail(PXG_ERR_ARG, "pxg_set_fishbone: null stream");
  s->earlyFishbone = early != 0;
  s->lateFishbone = late != 0;
  dropGraphs(s);
  return 0;
}

int pxg_set_riemann_fit(pxg_stream* s, int riemann) {
  if (!s)
    return fail(PXG_ERR_ARG, "pxg_set_riemann_fit: null stream");
  s->riemannFit = riemann != 0;
  dropGraphs(s);
  return 0;
}

int pxg_set_stats(pxg_stream* s, int on) {
  if (!s)
    return fail(PXG_ERR_ARG, "pxg_set_stats: null stream");
  s->doStats = on != 0;
  dropGraphs(s);
  return 0;
}

int pxg_get_counters(pxg_stream* s, uint64_t* out) {
  if (!s || !out)
    return fail(PXG_ERR_ARG, "pxg_get_counters: null argument");
  static_assert(sizeof(unsigned long long) == sizeof(uint64_t));
  PXG_HIP(hipMemcpyAsync(out, s->counters, PXG_N_COUNTERS * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
  PXG_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

int pxg_reset_counters(pxg_stream* s) {
  if (!s)
    return fail(PXG_ERR_ARG, "pxg_reset_counters: null stream");
  PXG_HIP(hipMemsetAsync(s->counters, 0, PXG_N_COUNTERS * sizeof(uint64_t), s->stream));
  PXG_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

int pxg_print_counters(pxg_stream* s) {
  uint64_t c[PXG_N_COUNTERS];
  if (int rc = pxg_get_counters(s, c))
    return rc;
  // kernel_printCounters (CAHitNtupletGeneratorKernelsImpl.h:577-608), including its order of
  // the raw nGoodTracks and nFitTracks columns
  std::printf(
      "||Counters | nEvents | nHits | nCells | nTuples | nFitTacks  |  nGoodTracks | nUsedHits | nDupHits | "
      "nKilledCells | nEmptyCells | nZeroTrackCells ||\n");
  std::printf("Counters Raw %llu %llu %llu %llu %llu %llu %llu %llu %llu %llu %llu\n", (unsigned long long)c[0],
              (unsigned long long)c[1], (unsigned long long)c[2], (unsigned long long)c[3], (unsigned long long)c[5],
              (unsigned long long)c[4], (unsigned long long)c[6], (unsigned long long)c[7], (unsigned long long)c[8],
              (unsigned long long)c[9], (unsigned long long)c[10]);
  const double ne = double(c[0]), nc = double(c[2]);
  std::printf("Counters Norm %llu ||  %.1f|  %.1f|  %.1f|  %.1f|  %.1f|  %.1f|  %.1f|  %.1f|  %.3f|  %.3f||\n",
              (unsigned long long)c[0], c[1] / ne, c[2] / ne, c[3] / ne, c[4] / ne, c[5] / ne, c[6] / ne, c[7] / ne,
              c[8] / ne, c[9] / nc, c[10] / nc);
  std::fflush(stdout);
  return 0;
}

int pxg_get_tuning(pxg_stream* s, uint32_t* doubletLanes, uint32_t* gridTarget) {
  if (!s || !doubletLanes || !gridTarget)
    return fail(PXG_ERR_ARG, "pxg_get_tuning: null argument");
  *doubletLanes = s->doubletLanes;
  *gridTarget = s->gridTarget;
  return 0;
}

int pxg_set_launch_shape(pxg_stream* s, const pxg_launch_shape* shape) {
  if (!s || !shape)
    return fail(PXG_ERR_ARG, "pxg_set_launch_shape: null argument");
  const pxg_launch_shape& h = *shape;
  if (h.doublet_lanes == 0 || h.doublet_lanes > 64 || (h.doublet_lanes & (h.doublet_lanes - 1)) ||
      h.grid_target < 4 || h.max_blocks_per_event < 4 || h.max_blocks_per_event > 4096 ||
      (h.connect_stage != 128 && h.connect_stage != 256 && h.connect_stage != 512) ||
      (h.fishbone_window != 128 && h.fishbone_window != 256 && h.fishbone_window != 384))
    return fail(PXG_ERR_ARG, "pxg_set_launch_shape: lanes a power of two in [1, 64], grid target >= 4, "
                             "4..4096 workgroups per event, connect stage 128/256/512, fishbone window 128/256/384");
  s->doubletLanes = h.doublet_lanes;
  s->gridTarget = h.grid_target;
  s->maxBlocksPerEvent = h.max_blocks_per_event;
  s->connectStage = h.connect_stage;
  s->fishboneWindow = h.fishbone_window;
  dropGraphs(s);
  return 0;
}

int pxg_get_launch_shape(pxg_stream* s, pxg_launch_shape* shape) {
  if (!s || !shape)
    return fail(PXG_ERR_ARG, "pxg_get_launch_shape: null argument");
  *shape = {s->doubletLanes, s->gridTarget, s->maxBlocksPerEvent, s->connectStage, s->fishboneWindow};
  return 0;
}

namespace {
  // median of `passes` timed runs of the chain (one warm-up run first)
  int timeShape(pxg_stream* s, uint32_t first, uint32_t count, int upto, int passes, hipEvent_t a, hipEvent_t b,
                float& ms) {
    int rc = runStages(s, first, count, 1, upto);
    float t[8];
    for (int k = 0; !rc && k < passes; ++k) {
      (void)hipEventRecord(a, s->stream);
      rc = runStages(s, first, count, 1, upto);
      (void)hipEventRecord(b, s->stream);
      if (!rc && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&t[k], a, b) != hipSuccess))
        rc = fail(PXG_ERR_HIP, "pxg_autotune: timing failed");
    }
    if (rc)
      return rc;
    std::sort(t, t + passes);
    ms = t[passes / 2];
    return 0;
  }
}  // namespace

int pxg_autotune_shape(pxg_stream* s, uint32_t first, uint32_t count, int upto, pxg_launch_shape* best,
                       float* bestMs) {
  if (!s || upto < 1 || upto > 5 || count == 0 || first + count > s->nLoaded)
    return fail(PXG_ERR_ARG, "pxg_autotune: bad arguments");
  constexpr int kPasses = 3;
  static const uint32_t lanes[] = {2, 4, 8, 16, 32, 64};
  static const uint32_t grids[] = {2048, 4096, 8192, 16384};
  static const uint32_t stages[] = {128, 256, 512};
  static const uint32_t windows[] = {128, 256, 384};
  static const uint32_t perEvent[] = {256, 512, 1024};
  dropGraphs(s);  // captured with the shape it replaces
  const bool wasProfiling = s->profiling;
  s->profiling = false;
  hipEvent_t a, b;
  PXG_HIP(hipEventCreate(&a));
  PXG_HIP(hipEventCreate(&b));
  float bestT = 1e30f;
  pxg_launch_shape cur{}, keep{};
  (void)pxg_get_launch_shape(s, &cur);
  keep = cur;
  int rc = 0;
  auto trial = [&](const pxg_launch_shape& h) {
    if (rc)
      return;
    rc = pxg_set_launch_shape(s, &h);
    float ms = 0;
    if (!rc)
      rc = timeShape(s, first, count, upto, kPasses, a, b, ms);
    if (!rc && ms < bestT) {
      bestT = ms;
      keep = h;
    }
  };
  // coordinate descent: doublet lanes x grid target, then connect's staging, the
  // fishbone's window and the workgroups per event, each around the best so far
  for (uint32_t g : grids)
    for (uint32_t ln : lanes) {
      pxg_launch_shape h = cur;
      h.doublet_lanes = ln;
      h.grid_target = g;
      trial(h);
    }
  for (uint32_t v : stages) {
    pxg_launch_shape h = keep;
    h.connect_stage = v;
    trial(h);
  }
  for (uint32_t v : windows) {
    pxg_launch_shape h = keep;
    h.fishbone_window = v;
    trial(h);
  }
  for (uint32_t v : perEvent) {
    pxg_launch_shape h = keep;
    h.max_blocks_per_event = v;
    trial(h);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  s->profiling = wasProfiling;
  if (rc)
    return rc;
  rc = pxg_set_launch_shape(s, &keep);
  if (rc)
    return rc;
  if (best)
    *best = keep;
  if (bestMs)
    *bestMs = bestT;
  return 0;
}

int pxg_autotune(pxg_stream* s, uint32_t first, uint32_t count, int upto, uint32_t* bestLanes,
                 uint32_t* bestGrid, float* bestMs) {
  pxg_launch_shape h{};
  const int rc = pxg_autotune_shape(s, first, count, upto, &h, bestMs);
  if (rc)
    return rc;
  if (bestLanes)
    *bestLanes = h.doublet_lanes;
  if (bestGrid)
    *bestGrid = h.grid_target;
  return 0;
}

// ---- profiling ------------------------------------------------------------------------
int pxg_profile_enable(pxg_stream* s, int on) {
  if (!s)
    return fail(PXG_ERR_ARG, "null stream");
  s->profiling = on != 0;
  return 0;
}

int pxg_profile_read(pxg_stream* s, const char** names, uint32_t* calls, double* totalMs, int max) {
  if (!s)
    return fail(PXG_ERR_ARG, "null stream");
  PXG_HIP(hipStreamSynchronize(s->stream));
  for (auto& m : s->marks) {
    float ms = 0;
    PXG_HIP(hipEventElapsedTime(&ms, m.start, m.stop));
    auto& a = s->profile[m.name];
    a.first += 1;
    a.second += ms;
    (void)hipEventDestroy(m.start);
    (void)hipEventDestroy(m.stop);
  }
  s->marks.clear();
  int i = 0;
  static thread_local std::vector<std::string> keep;
  keep.clear();
  for (auto& [name, v] : s->profile)
    keep.push_back(name);
  for (auto& [name, v] : s->profile) {
    if (i < max) {
      names[i] = keep[i].c_str();
      calls[i] = v.first;
      totalMs[i] = v.second;
    }
    ++i;
  }
  s->profile.clear();
  return i;
}

}  // extern "C"

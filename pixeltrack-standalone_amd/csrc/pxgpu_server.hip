// Serving loop of libpxgpu.so: continuous batching of events that arrive one
// by one (include/pxgpu.h, pxg_server_*).
//
// Requests enter a queue.  A dispatcher thread owns `nStreams` slots, each a
// pxg_stream with a device pool for one batch.  Whenever a slot is free and
// requests wait, it admits the oldest waiting requests into a batch while the
// batch's event count and raw words fit the slot's pool (the pool sizes come
// from the HBM budget), and enqueues the load and the whole chain as one launch
// per stage.  The load is one kernel that reads every request's words from
// mapped pinned host memory: the library's chunks (pxg_server_submit copies the
// caller's words there) or the caller's own pinned buffers
// (pxg_server_submit_pinned, no host copy at all).  Slots run concurrently on their own HIP streams, so a
// batch formed while others are on the device overlaps with them.  When a
// batch's counters are back, the per-event products are packed on the device
// into one buffer and copied to the host in one transfer; the requests are
// then complete, with their submit and completion times.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pxgpu_stream.h"

using namespace pxg;
using pxg::detail::fail;

namespace {
  uint64_t nowNs() {
    return uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(
                        std::chrono::steady_clock::now().time_since_epoch())
                        .count());
  }

  // Packed result of one event (byte offsets relative to the event's base):
  // header[16] u32 | tupleOff[nT + 1] u32 | tupleBins[nBins] u16 | quality[nT] u8 |
  // state[5][nT] f32 | idv[nT] i16 | zv[nV] f32 | sortInd[nV] u16, each part 4-byte aligned.
  struct PackLayout {
    uint32_t base, nT, nBinsMax, nV;
    uint32_t oOff, oBins, oQual, oState, oIdv, oZv, oSort, size;
  };

  uint32_t align4(uint32_t v) { return (v + 3) & ~3u; }

  PackLayout layoutOf(uint32_t base, uint32_t nT, uint32_t nV) {
    PackLayout l{};
    l.base = base;
    l.nT = nT;
    l.nBinsMax = 5 * nT;
    l.nV = nV;
    uint32_t o = 16 * 4;
    l.oOff = o;
    o += align4(4 * (nT + 1));
    l.oBins = o;
    o += align4(2 * l.nBinsMax);
    l.oQual = o;
    o += align4(nT);
    l.oState = o;
    o += 4 * 5 * nT;
    l.oIdv = o;
    o += align4(2 * nT);
    l.oZv = o;
    o += 4 * nV;
    l.oSort = o;
    o += align4(2 * nV);
    l.size = align4(o);
    return l;
  }

  // one workgroup per event of the batch
  __global__ void __launch_bounds__(256) packKernel(Pool p, CAPool ca, const PackLayout* layouts, uint8_t* out) {
    const uint32_t e = blockIdx.x;
    const PackLayout L = layouts[e];
    uint8_t* o = out + L.base;
    const size_t tb = ca.tupleBase(e);
    const uint32_t M = ca.caps.maxTuples;
    const uint32_t* off = ca.tupleOff + ca.offBase(e);
    if (threadIdx.x == 0) {
      uint32_t* h = reinterpret_cast<uint32_t*>(o);
      const EventCounts c = p.counts[e];
      const CACounts k = ca.cnt[e];
      const uint32_t vals[16] = {p.events[e].nWords, c.nModules, c.nClusters, c.status, k.status, k.nTuples,
                                 k.nVertices, off[L.nT], k.nCells, k.apcM, k.vstatus, 0, 0, 0, 0, 0};
      for (int i = 0; i < 16; ++i)
        h[i] = vals[i];
    }
    uint32_t* pOff = reinterpret_cast<uint32_t*>(o + L.oOff);
    for (uint32_t t = threadIdx.x; t <= L.nT; t += blockDim.x)
      pOff[t] = off[t];
    const uint32_t nBins = min(off[L.nT], L.nBinsMax);
    uint16_t* pBins = reinterpret_cast<uint16_t*>(o + L.oBins);
    for (uint32_t i = threadIdx.x; i < nBins; i += blockDim.x)
      pBins[i] = ca.tupleBins[5 * tb + i];
    for (uint32_t t = threadIdx.x; t < L.nT; t += blockDim.x) {
      o[L.oQual + t] = ca.quality[tb + t];
      reinterpret_cast<int16_t*>(o + L.oIdv)[t] = ca.idv[tb + t];
    }
    float* pState = reinterpret_cast<float*>(o + L.oState);
    for (uint32_t i = threadIdx.x; i < 5 * L.nT; i += blockDim.x) {
      const uint32_t k = i / L.nT, t = i % L.nT;
      pState[i] = ca.state[5 * tb + size_t(k) * M + t];
    }
    for (uint32_t v = threadIdx.x; v < L.nV; v += blockDim.x) {
      reinterpret_cast<float*>(o + L.oZv)[v] = ca.zv[size_t(e) * kMaxVertices + v];
      reinterpret_cast<uint16_t*>(o + L.oSort)[v] = ca.sortInd[size_t(e) * kMaxVertices + v];
    }
  }

  // Mapped pinned host chunks holding one event's words and FED ids, which the
  // load kernel reads directly (sized for the largest event).
  constexpr size_t kChunkWords = kMaxFedWords;
  constexpr size_t kChunkBytes = kChunkWords * 4 + kChunkWords / 2 + 256;
  constexpr unsigned kMappedPinned = hipHostMallocMapped | hipHostMallocPortable;

  struct Chunk {
    uint8_t* host = nullptr;       // words at host, FED ids at host + 4 * kChunkWords
    const uint8_t* dev = nullptr;  // the same memory as the device addresses it
  };

  struct Request {
    uint64_t id;
    Chunk chunk;                         // the library's chunk, or none for caller buffers
    const uint32_t* words = nullptr;     // device addresses of the words and FED ids
    const uint8_t* fedIds = nullptr;
    uint32_t nWords = 0;
    pxg_beamspot bs;
    uint32_t counts[4] = {0, 0, 0, 0};  // words, hits, tuples, vertices
    uint64_t tSubmit = 0, tLaunch = 0, tDone = 0;
    int status = 0;
    uint32_t batchSize = 0;
    std::vector<uint8_t> packed;  // the event's packed products
  };

  enum class SlotState { Idle, Running, Packing };

  struct Slot {
    pxg_stream* s = nullptr;
    SlotState state = SlotState::Idle;
    std::vector<std::unique_ptr<Request>> batch;
    PackLayout* hLayout = nullptr;
    PackLayout* dLayout = nullptr;
    uint8_t* dPack = nullptr;
    uint8_t* hPack = nullptr;
    size_t packCapacity = 0;
    uint32_t packBytes = 0;
    hipEvent_t evRun = nullptr, evPack = nullptr;
    pxg::detail::GatherSrc* gather = nullptr;           // mapped pinned, one entry per event
    const pxg::detail::GatherSrc* gatherDev = nullptr;  // its device address
  };
}  // namespace

struct pxg_server {
  pxg_server_config cfg{};
  pxg_context* ctx = nullptr;
  std::vector<Chunk> freeChunks, allChunks;  // guarded by mu
  uint32_t maxBatchWords = 0;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::unique_ptr<Request>> waiting;
  std::deque<std::unique_ptr<Request>> done;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> kept;  // polled, products readable
  uint64_t nextId = 1;
  std::atomic<bool> stop{false};
  std::atomic<int> error{0};
  std::string errorText;
  std::thread thread;
  uint64_t batches = 0, batchedEvents = 0;

  int launch(Slot& sl);
  int pack(Slot& sl);
  void finish(Slot& sl);
  void loop();
};

int pxg_server::launch(Slot& sl) {
  const uint32_t n = uint32_t(sl.batch.size());
  std::vector<pxg_raw_view> views(n);
  std::vector<pxg_beamspot> beamspots(n);  // each request's own beam spot
  const uint64_t t = nowNs();
  for (uint32_t e = 0; e < n; ++e) {
    Request& r = *sl.batch[e];
    views[e] = {r.words, r.fedIds, r.nWords};
    beamspots[e] = r.bs;
    r.tLaunch = t;
    r.batchSize = n;
  }
  int rc = pxg::detail::loadEventsGather(sl.s, n, views.data(), beamspots.data(), sl.gather, sl.gatherDev);
  if (!rc)
    rc = pxg::detail::runStages(sl.s, 0, n, 1, cfg.upto);
  if (rc)
    return rc;
  pxg_stream* s = sl.s;
  if (hipMemcpyAsync(s->hostCounts, s->pool.counts, n * sizeof(EventCounts), hipMemcpyDeviceToHost, s->stream) ||
      hipMemcpyAsync(s->hostCA, s->ca.cnt, n * sizeof(CACounts), hipMemcpyDeviceToHost, s->stream) ||
      hipEventRecord(sl.evRun, s->stream))
    return fail(PXG_ERR_HIP, "server: enqueue of a batch failed");
  sl.state = SlotState::Running;
  return 0;
}

int pxg_server::pack(Slot& sl) {
  pxg_stream* s = sl.s;
  const uint32_t n = uint32_t(sl.batch.size());
  uint32_t base = 0;
  for (uint32_t e = 0; e < n; ++e) {
    const uint32_t nT = cfg.upto >= 4 ? s->hostCA[e].nTuples : 0;
    const uint32_t nV = cfg.upto >= 5 ? std::min(s->hostCA[e].nVertices, kMaxVertices) : 0;
    sl.hLayout[e] = layoutOf(base, nT, nV);
    base += sl.hLayout[e].size;
  }
  if (base > sl.packCapacity)
    return fail(PXG_ERR_CAPACITY, "server: packed products of a batch exceed %zu bytes", sl.packCapacity);
  sl.packBytes = base;
  if (hipMemcpyAsync(sl.dLayout, sl.hLayout, n * sizeof(PackLayout), hipMemcpyHostToDevice, s->stream))
    return fail(PXG_ERR_HIP, "server: layout upload failed");
  const CAPool ca = pxg::detail::caViewFrom(s->ca, 0);
  const Pool v = pxg::detail::viewFrom(s->pool, 0);
  packKernel<<<n, 256, 0, s->stream>>>(v, ca, sl.dLayout, sl.dPack);
  if (hipGetLastError() || hipMemcpyAsync(sl.hPack, sl.dPack, base, hipMemcpyDeviceToHost, s->stream) ||
      hipEventRecord(sl.evPack, s->stream))
    return fail(PXG_ERR_HIP, "server: packing failed");
  sl.state = SlotState::Packing;
  return 0;
}

void pxg_server::finish(Slot& sl) {
  const uint64_t t = nowNs();
  pxg_stream* s = sl.s;
  // products copied out of the pinned buffer before taking the lock
  if (cfg.transfer)
    for (uint32_t e = 0; e < sl.batch.size(); ++e) {
      const PackLayout& L = sl.hLayout[e];
      sl.batch[e]->packed.assign(sl.hPack + L.base, sl.hPack + L.base + L.size);
    }
  std::lock_guard<std::mutex> g(mu);
  for (uint32_t e = 0; e < sl.batch.size(); ++e) {
    auto& r = sl.batch[e];
    r->counts[0] = r->nWords;
    r->counts[1] = s->hostCounts[e].nClusters;
    r->counts[2] = cfg.upto >= 4 ? s->hostCA[e].nTuples : 0;
    r->counts[3] = cfg.upto >= 5 ? s->hostCA[e].nVertices : 0;
    if (r->chunk.host)
      freeChunks.push_back(r->chunk);
    r->chunk = {};
    r->words = nullptr;
    r->fedIds = nullptr;
    // counters of the stages the server runs only (the others hold an earlier batch's)
    const uint32_t caBits = cfg.upto >= 4 ? s->hostCA[e].status << PXG_STATUS_CA_SHIFT : 0u;
    const uint32_t vBits = cfg.upto >= 5 && s->hostCA[e].vstatus ? uint32_t(PXG_STATUS_VERTEX_OVERFLOW) : 0u;
    r->status = int(s->hostCounts[e].status | caBits | vBits);
    r->tDone = t;
    done.push_back(std::move(r));
  }
  batches += 1;
  batchedEvents += sl.batch.size();
  sl.batch.clear();
  sl.state = SlotState::Idle;
  cv.notify_all();
}

void pxg_server::loop() {
  (void)hipSetDevice(ctx->device);
  while (true) {
    bool progressed = false;
    for (Slot& sl : slots) {
      int rc = 0;
      if (sl.state == SlotState::Running && hipEventQuery(sl.evRun) == hipSuccess) {
        if (cfg.transfer)
          rc = pack(sl);
        else
          finish(sl);
        progressed = true;
      } else if (sl.state == SlotState::Packing && hipEventQuery(sl.evPack) == hipSuccess) {
        finish(sl);
        progressed = true;
      }
      if (!rc && sl.state == SlotState::Idle) {
        // admit the oldest waiting requests that fit the slot's pool
        std::unique_lock<std::mutex> g(mu);
        uint32_t words = 0;
        while (!waiting.empty() && sl.batch.size() < cfg.max_batch) {
          const uint32_t w = (waiting.front()->nWords + 255) / 256 * 256;
          if (!sl.batch.empty() && words + w > maxBatchWords)
            break;
          words += w;
          sl.batch.push_back(std::move(waiting.front()));
          waiting.pop_front();
        }
        g.unlock();
        if (!sl.batch.empty()) {
          rc = launch(sl);
          progressed = true;
        }
      }
      if (rc) {
        std::lock_guard<std::mutex> g(mu);
        error = rc;
        errorText = pxg_last_error();
        stop = true;
        cv.notify_all();
        return;
      }
    }
    if (!progressed) {
      std::unique_lock<std::mutex> g(mu);
      bool busy = false;
      for (Slot& sl : slots)
        busy |= sl.state != SlotState::Idle;
      if (stop && !busy && waiting.empty())
        return;
      // wait for a request, or poll the device again shortly
      cv.wait_for(g, std::chrono::microseconds(busy ? 20 : 200));
    }
  }
}

extern "C" {

int pxg_server_create(pxg_context* ctx, const pxg_server_config* cfg, pxg_server** out) {
  if (!ctx || !cfg || !out || cfg->max_batch == 0 || cfg->n_streams == 0 || cfg->upto < 1 || cfg->upto > 5)
    return fail(PXG_ERR_ARG, "pxg_server_create: bad arguments");
  if (cfg->profile != PXG_PROFILE_REFERENCE && cfg->profile != PXG_PROFILE_HIGH_OCCUPANCY)
    return fail(PXG_ERR_ARG, "pxg_server_create: unknown capacity profile %d", cfg->profile);
  const CACapacities caps = cfg->profile == PXG_PROFILE_HIGH_OCCUPANCY ? kCapsHighOccupancy : kCapsReference;
  // pool of one slot: B events and W raw words; W from the HBM budget
  const uint32_t B = cfg->max_batch;
  const size_t packPerEvent = layoutOf(0, caps.maxTuples, kMaxVertices).size;
  const size_t fixed = pxg::detail::poolBytes(B, 0, caps) + B * packPerEvent;
  const double perWord = double(pxg::detail::poolBytes(B, 1u << 20, caps) - pxg::detail::poolBytes(B, 0, caps)) /
                         double(1u << 20);
  uint64_t W = uint64_t(B) * kMaxFedWords;
  if (cfg->hbm_budget) {
    const double perSlot = double(cfg->hbm_budget) / cfg->n_streams;
    if (perSlot < double(fixed) + perWord * kMaxFedWords)
      return fail(PXG_ERR_CAPACITY, "pxg_server_create: budget %llu B cannot hold %u slots of %u events",
                  (unsigned long long)cfg->hbm_budget, cfg->n_streams, B);
    W = std::min<uint64_t>(W, uint64_t((perSlot - double(fixed)) / perWord));
  }
  auto* sv = new pxg_server;
  sv->cfg = *cfg;
  sv->ctx = ctx;
  sv->maxBatchWords = uint32_t(W);
  sv->slots.resize(cfg->n_streams);
  for (Slot& sl : sv->slots) {
    int rc = pxg_stream_create_profile(ctx, B, sv->maxBatchWords, cfg->profile, &sl.s);
    if (!rc && cfg->doublet_lanes)
      rc = pxg_set_tuning(sl.s, cfg->doublet_lanes, cfg->grid_target ? cfg->grid_target : 2048);
    sl.packCapacity = B * packPerEvent;
    hipError_t e = hipSuccess;
    if (!rc) {
      e = hipHostMalloc(reinterpret_cast<void**>(&sl.hLayout), B * sizeof(PackLayout), hipHostMallocDefault);
      if (!e)
        e = hipHostMalloc(reinterpret_cast<void**>(&sl.hPack), sl.packCapacity, hipHostMallocDefault);
      if (!e)
        e = hipMalloc(&sl.dLayout, B * sizeof(PackLayout));
      if (!e)
        e = hipMalloc(&sl.dPack, sl.packCapacity);
      if (!e)
        e = hipEventCreateWithFlags(&sl.evRun, hipEventDisableTiming);
      if (!e)
        e = hipEventCreateWithFlags(&sl.evPack, hipEventDisableTiming);
      if (!e)
        e = hipHostMalloc(reinterpret_cast<void**>(&sl.gather), B * sizeof(pxg::detail::GatherSrc), kMappedPinned);
      if (!e)
        e = hipHostGetDevicePointer(reinterpret_cast<void**>(const_cast<pxg::detail::GatherSrc**>(&sl.gatherDev)),
                                    sl.gather, 0);
    }
    if (rc || e) {
      pxg_server_destroy(sv);
      return rc ? rc : fail(PXG_ERR_HIP, "pxg_server_create: %s", hipGetErrorString(e));
    }
  }
  // pinned chunks for the events of all slots, a batch's worth waiting, and a margin
  const size_t nChunks = size_t(cfg->n_streams + 1) * cfg->max_batch + 256;
  for (size_t k = 0; k < nChunks; ++k) {
    Chunk c;
    void* dev = nullptr;
    if (hipHostMalloc(reinterpret_cast<void**>(&c.host), kChunkBytes, kMappedPinned) != hipSuccess ||
        hipHostGetDevicePointer(&dev, c.host, 0) != hipSuccess) {
      if (c.host)
        (void)hipHostFree(c.host);
      pxg_server_destroy(sv);
      return fail(PXG_ERR_HIP, "pxg_server_create: pinned host memory for %zu events unavailable", nChunks);
    }
    c.dev = static_cast<const uint8_t*>(dev);
    sv->allChunks.push_back(c);
    sv->freeChunks.push_back(c);
  }
  sv->thread = std::thread([sv] { sv->loop(); });
  *out = sv;
  return 0;
}

void pxg_server_destroy(pxg_server* sv) {
  if (!sv)
    return;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop = true;
    sv->cv.notify_all();
  }
  if (sv->thread.joinable())
    sv->thread.join();
  for (Slot& sl : sv->slots) {
    if (sl.s)
      (void)hipStreamSynchronize(sl.s->stream);
    if (sl.evRun)
      (void)hipEventDestroy(sl.evRun);
    if (sl.evPack)
      (void)hipEventDestroy(sl.evPack);
    if (sl.dLayout)
      (void)hipFree(sl.dLayout);
    if (sl.dPack)
      (void)hipFree(sl.dPack);
    if (sl.hLayout)
      (void)hipHostFree(sl.hLayout);
    if (sl.hPack)
      (void)hipHostFree(sl.hPack);
    if (sl.gather)
      (void)hipHostFree(sl.gather);
    pxg_stream_destroy(sl.s);
  }
  for (const Chunk& c : sv->allChunks)
    (void)hipHostFree(c.host);
  delete sv;
}

int pxg_server_info(pxg_server* sv, uint32_t* max_batch_words, uint64_t* batches, uint64_t* events) {
  if (!sv)
    return fail(PXG_ERR_ARG, "pxg_server_info: null server");
  std::lock_guard<std::mutex> g(sv->mu);
  if (max_batch_words)
    *max_batch_words = sv->maxBatchWords;
  if (batches)
    *batches = sv->batches;
  if (events)
    *events = sv->batchedEvents;
  return 0;
}

int pxg_server_submit(pxg_server* sv, const uint32_t* words, const uint8_t* fedIds, uint32_t nWords,
                      const pxg_beamspot* bs, uint64_t* id) {
  if (!sv || !bs || !id || (nWords && (!words || !fedIds)) || nWords % 2)
    return fail(PXG_ERR_ARG, "pxg_server_submit: bad arguments");
  if (nWords > kChunkWords || nWords > sv->maxBatchWords)
    return fail(PXG_ERR_CAPACITY, "pxg_server_submit: %u words exceed the capacity", nWords);
  if (sv->error)
    return fail(sv->error, "pxg_server_submit: server stopped: %s", sv->errorText.c_str());
  auto r = std::make_unique<Request>();
  r->tSubmit = nowNs();
  {
    // a free pinned chunk; when all are taken the caller waits (back-pressure)
    std::unique_lock<std::mutex> g(sv->mu);
    sv->cv.wait(g, [&] { return !sv->freeChunks.empty() || sv->error; });
    if (sv->error)
      return fail(sv->error, "pxg_server_submit: server stopped: %s", sv->errorText.c_str());
    r->chunk = sv->freeChunks.back();
    sv->freeChunks.pop_back();
  }
  std::memcpy(r->chunk.host, words, size_t(nWords) * 4);
  std::memcpy(r->chunk.host + 4 * kChunkWords, fedIds, nWords / 2);
  r->words = reinterpret_cast<const uint32_t*>(r->chunk.dev);
  r->fedIds = r->chunk.dev + 4 * kChunkWords;
  r->nWords = nWords;
  r->bs = *bs;
  std::lock_guard<std::mutex> g(sv->mu);
  r->id = sv->nextId++;
  *id = r->id;
  sv->waiting.push_back(std::move(r));
  sv->cv.notify_all();
  return 0;
}

int pxg_server_submit_pinned(pxg_server* sv, const uint32_t* words, const uint8_t* fedIds, uint32_t nWords,
                             const pxg_beamspot* bs, uint64_t* id) {
  if (!sv || !bs || !id || (nWords && (!words || !fedIds)) || nWords % 2)
    return fail(PXG_ERR_ARG, "pxg_server_submit_pinned: bad arguments");
  if (nWords > kChunkWords || nWords > sv->maxBatchWords)
    return fail(PXG_ERR_CAPACITY, "pxg_server_submit_pinned: %u words exceed the capacity", nWords);
  if (sv->error)
    return fail(sv->error, "pxg_server_submit_pinned: server stopped: %s", sv->errorText.c_str());
  auto r = std::make_unique<Request>();
  r->tSubmit = nowNs();
  if (nWords) {
    void *dw = nullptr, *df = nullptr;
    if (hipHostGetDevicePointer(&dw, const_cast<uint32_t*>(words), 0) != hipSuccess ||
        hipHostGetDevicePointer(&df, const_cast<uint8_t*>(fedIds), 0) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PXG_ERR_ARG, "pxg_server_submit_pinned: the words or FED ids are not mapped pinned memory "
                               "(pxg_host_alloc)");
    }
    r->words = static_cast<const uint32_t*>(dw);
    r->fedIds = static_cast<const uint8_t*>(df);
  }
  r->nWords = nWords;
  r->bs = *bs;
  std::lock_guard<std::mutex> g(sv->mu);
  r->id = sv->nextId++;
  *id = r->id;
  sv->waiting.push_back(std::move(r));
  sv->cv.notify_all();
  return 0;
}

int pxg_host_alloc(size_t bytes, void** out) {
  if (!out || bytes == 0)
    return fail(PXG_ERR_ARG, "pxg_host_alloc: bad arguments");
  *out = nullptr;
  const hipError_t e = hipHostMalloc(out, bytes, kMappedPinned);
  if (e != hipSuccess)
    return fail(PXG_ERR_HIP, "pxg_host_alloc(%zu): %s", bytes, hipGetErrorString(e));
  return 0;
}

void pxg_host_free(void* p) {
  if (p)
    (void)hipHostFree(p);
}

int pxg_server_poll(pxg_server* sv, pxg_server_result* out, int max, int wait_ms) {
  if (!sv || !out || max <= 0)
    return fail(PXG_ERR_ARG, "pxg_server_poll: bad arguments");
  std::unique_lock<std::mutex> g(sv->mu);
  if (sv->done.empty() && wait_ms > 0)
    sv->cv.wait_for(g, std::chrono::milliseconds(wait_ms), [&] { return !sv->done.empty() || sv->error; });
  if (sv->error && sv->done.empty())
    return fail(sv->error, "pxg_server_poll: server stopped: %s", sv->errorText.c_str());
  int n = 0;
  while (n < max && !sv->done.empty()) {
    auto r = std::move(sv->done.front());
    sv->done.pop_front();
    out[n] = {r->id,        r->status,   r->tSubmit,  r->tLaunch,  r->tDone,
              r->batchSize, r->counts[0], r->counts[1], r->counts[2], r->counts[3]};
    sv->kept[r->id] = std::move(r);
    ++n;
  }
  return n;
}

int pxg_server_products(pxg_server* sv, uint64_t id, const void** packed, uint32_t* bytes) {
  if (!sv || !packed || !bytes)
    return fail(PXG_ERR_ARG, "pxg_server_products: bad arguments");
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->kept.find(id);
  if (it == sv->kept.end())
    return fail(PXG_ERR_ARG, "pxg_server_products: request %llu not completed or released", (unsigned long long)id);
  if (!sv->cfg.transfer)
    return fail(PXG_ERR_ARG, "pxg_server_products: the server was created without transfer");
  *packed = it->second->packed.data();
  *bytes = uint32_t(it->second->packed.size());
  return 0;
}

int pxg_server_release(pxg_server* sv, uint64_t id) {
  if (!sv)
    return fail(PXG_ERR_ARG, "pxg_server_release: null server");
  std::lock_guard<std::mutex> g(sv->mu);
  sv->kept.erase(id);
  return 0;
}

}  // extern "C"

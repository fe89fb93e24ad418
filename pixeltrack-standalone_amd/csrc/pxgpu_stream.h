// Host-side state of libpxgpu.so shared by the C ABI (pxgpu_api.hip) and the
// serving loop (pxgpu_server.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pxgpu.h"
#include "pxgpu_internal.h"

struct pxg_context {
  int device = 0;
  pxg::DevConditions cond{};
  std::vector<void*> owned;
};

struct ProfileMark {
  const char* name;
  hipEvent_t start, stop;
};

struct pxg_stream {
  pxg_context* ctx = nullptr;
  hipStream_t stream = nullptr;
  pxg::Pool pool{};
  pxg::CAPool ca{};
  pxg::CACounts* hostCA = nullptr;  // pinned
  std::vector<void*> owned;
  pxg::EventDesc* hostEvents = nullptr;    // pinned
  pxg::EventCounts* hostCounts = nullptr;  // pinned
  uint32_t nLoaded = 0;
  uint32_t wordCapacity = 0;
  // per resident event: the last stage run on it since it was loaded (0 none), so that
  // counters of stages not run (left over from earlier events) are never reported
  std::vector<uint8_t> stagesRun;
  bool profiling = false;
  uint32_t doubletLanes = 8;   // tuning: lanes per doublet slot
  uint32_t gridTarget = 2048;  // tuning: workgroups per grid-stride launch (divided among the events)
  uint32_t maxBlocksPerEvent = 512;  // tuning: cap of the workgroups per event of such a launch
  uint32_t connectStage = 256;       // tuning: connect's staged candidates per wave (128, 256, 512)
  uint32_t fishboneWindow = 256;     // tuning: the fishbone's staged list entries per wave (128, 256, 384)
  int vertexAlgo = 0;          // vertex clustering algorithm (PXG_VERTEX_*)
  bool earlyFishbone = true;   // the reference's defaults: early fishbone on, late off
  bool lateFishbone = false;
  bool riemannFit = false;     // the reference's default: broken-line fit
  bool doStats = false;        // the reference's default: no counters
  float bField = 0.0114256972711507f;  // CAHitNtupletCUDA.cc:31
  float vertexPtMin = 0.5f;            // PixelVertexProducerCUDA.cc:37
  unsigned long long* counters = nullptr;  // device, PXG_N_COUNTERS
  // pxg_run's stage sequences captured as HIP graphs, by (first, count, upto); their
  // grid sizes come from the loaded events and the settings, so loading events or
  // changing a setting drops them
  bool useGraphs = false;
  std::map<std::tuple<uint32_t, uint32_t, int>, hipGraphExec_t> graphs;
  std::vector<ProfileMark> marks;
  std::map<std::string, std::pair<uint32_t, double>> profile;
};


namespace pxg::detail {
  // records a message for pxg_last_error and returns code
  int fail(int code, const char* fmt, ...);
  // per-event arrays shifted to event `first`
  Pool viewFrom(const Pool& p, uint32_t first);
  CAPool caViewFrom(const CAPool& c, uint32_t first);
  // enqueues stages [from, upto] on resident events [first, first + count)
  int runStages(pxg_stream* s, uint32_t first, uint32_t count, int from, int upto);
  // records that stages [from, upto] ran on events [first, first + count)
  void markStages(pxg_stream* s, uint32_t first, uint32_t count, int from, int upto);
  // copies n events into the pool; host arrays must stay valid until the copies
  // complete (pinned and kept alive by the caller when sync is false).  bs holds
  // nBeamspots entries: 1 (every event gets bs[0]) or n (event e gets bs[e]).
  int loadEvents(pxg_stream* s, uint32_t n, const pxg_raw_view* raws, const pxg_beamspot* bs, uint32_t nBeamspots,
                 bool sync);
  // one event's raw words for loadEventsGather, in host memory the device can read
  // (mapped pinned memory, addresses as the device sees them)
  struct GatherSrc {
    const uint32_t* words;
    const uint8_t* fedIds;
    uint32_t nWords, wordOffset;
  };
  // like loadEvents, but one kernel reads every event's words and FED ids from
  // device-accessible host memory (raws hold device addresses), instead of two
  // copies per event; table: n entries of mapped pinned memory, tableDev its
  // device address.  Asynchronous: the sources must stay unchanged until the
  // stream has passed the load.
  int loadEventsGather(pxg_stream* s, uint32_t n, const pxg_raw_view* raws, const pxg_beamspot* bs,
                       GatherSrc* table, const GatherSrc* tableDev);
  // bytes of device memory a pool of maxEvents events and maxWords words takes
  size_t poolBytes(uint32_t maxEvents, uint32_t maxWords, const CACapacities& caps);
  // destroys the stream's captured stage sequences
  void dropGraphs(pxg_stream* s);
  // looks up the device's address of the no-alloc entry point's scratch array (once per
  // device; the device must be current)
  int noallocSetup(int device);
}  // namespace pxg::detail

// returns PXG_ERR_HIP (with the message for pxg_last_error) from the enclosing function
#define PXG_HIP(call)                                                                        \
  do {                                                                                       \
    hipError_t err_ = (call);                                                                \
    if (err_ != hipSuccess)                                                                  \
      return pxg::detail::fail(PXG_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(err_), __FILE__, \
                               __LINE__);                                                    \
  } while (0)

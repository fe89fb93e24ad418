// The modules of the pixel chain job (bin/main.cc:107-112 of the serial
// program): four conditions producers reading the data directory, the beam
// spot copy, and the four event producers, each calling the device library
// through its C ABI.  A module instance belongs to one EDM stream; the
// raw-to-cluster producer owns that stream's device pool and the products
// carry it to the later producers of the same event.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <stdexcept>
#include <string>

#include "DataFormats.h"
#include "Framework.h"

extern "C" int pxh_unpack_raw_record(const uint8_t* raw, size_t len, const uint32_t* fedList, uint32_t nFedList,
                                     uint32_t* words, uint8_t* fedIds, uint32_t capacity, uint32_t* nWords);

namespace {
  void check(int rc, char const* what) {
    if (rc < 0)
      throw std::runtime_error(std::string(what) + ": " + pxg_last_error());
  }

  struct Reader {
    std::ifstream in;
    std::string path;
    explicit Reader(std::string const& dir, char const* file) : in(dir + "/" + file, std::ios::binary), path(dir + "/" + file) {
      if (!in)
        throw std::runtime_error("cannot open " + path);
    }
    template <typename T>
    void read(T* dst, size_t n = 1) {
      in.read(reinterpret_cast<char*>(dst), std::streamsize(sizeof(T) * n));
      if (!in)
        throw std::runtime_error("truncated " + path);
    }
    uint32_t u32() {
      uint32_t v;
      read(&v);
      return v;
    }
  };
}  // namespace

// ---- conditions ------------------------------------------------------------------------

// BeamSpotESProducer.cc:19-26: beamspot.bin holds one BeamSpotPOD
class BeamSpotESProducer : public edm::ESProducer {
public:
  explicit BeamSpotESProducer(std::string const& dir) : dir_(dir) {}
  void produce(edm::EventSetup& es) override {
    Reader r(dir_, "beamspot.bin");
    auto bs = std::make_unique<BeamSpotPOD>();
    r.read(&bs->bs);
    es.put(std::move(bs));
  }

private:
  std::string dir_;
};
DEFINE_FWK_EVENTSETUP_MODULE(BeamSpotESProducer);

// SiPixelFedCablingMapGPUWrapperESProducer.cc:20-41: fedIds.bin (count + ids) and
// cablingMap.bin (the map, then count + modules-to-unpack flags)
class SiPixelFedCablingMapGPUWrapperESProducer : public edm::ESProducer {
public:
  explicit SiPixelFedCablingMapGPUWrapperESProducer(std::string const& dir) : dir_(dir) {}
  void produce(edm::EventSetup& es) override {
    {
      Reader r(dir_, "fedIds.bin");
      auto ids = std::make_unique<SiPixelFedIds>();
      ids->ids.resize(r.u32());
      r.read(ids->ids.data(), ids->ids.size());
      es.put(std::move(ids));
    }
    Reader r(dir_, "cablingMap.bin");
    auto w = std::make_unique<SiPixelFedCablingMapGPUWrapper>();
    w->map = std::make_unique<pxf::SiPixelFedCablingMapGPU>();
    r.read(w->map.get());
    w->modToUnp.resize(r.u32());
    r.read(w->modToUnp.data(), w->modToUnp.size());
    es.put(std::move(w));
  }

private:
  std::string dir_;
};
DEFINE_FWK_EVENTSETUP_MODULE(SiPixelFedCablingMapGPUWrapperESProducer);

// SiPixelGainCalibrationForHLTGPUESProducer.cc:20-30: gain.bin (header, then count + payload)
class SiPixelGainCalibrationForHLTGPUESProducer : public edm::ESProducer {
public:
  explicit SiPixelGainCalibrationForHLTGPUESProducer(std::string const& dir) : dir_(dir) {}
  void produce(edm::EventSetup& es) override {
    Reader r(dir_, "gain.bin");
    auto g = std::make_unique<SiPixelGainCalibrationForHLTGPU>();
    r.read(&g->gain);
    g->data.resize(r.u32());
    r.read(g->data.data(), g->data.size());
    es.put(std::move(g));
  }

private:
  std::string dir_;
};
DEFINE_FWK_EVENTSETUP_MODULE(SiPixelGainCalibrationForHLTGPUESProducer);

// PixelCPEFast.cc:17-28: cpefast.bin (common params, count + det params, average geometry,
// layer geometry)
class PixelCPEFastESProducer : public edm::ESProducer {
public:
  explicit PixelCPEFastESProducer(std::string const& dir) : dir_(dir) {}
  void produce(edm::EventSetup& es) override {
    Reader r(dir_, "cpefast.bin");
    auto c = std::make_unique<PixelCPEFast>();
    r.read(&c->common);
    c->det.resize(r.u32());
    r.read(c->det.data(), c->det.size());
    r.read(&c->average);
    r.read(&c->layer);
    es.put(std::move(c));
  }

private:
  std::string dir_;
};
DEFINE_FWK_EVENTSETUP_MODULE(PixelCPEFastESProducer);

// Uploads the conditions to the device once (the ES products' device copies of the
// reference are made per device on first use; here at the start of the job).
class PxgDeviceESProducer : public edm::ESProducer {
public:
  explicit PxgDeviceESProducer(std::string const&) {}
  void produce(edm::EventSetup& es) override {
    auto const& cab = es.get<SiPixelFedCablingMapGPUWrapper>();
    auto const& gain = es.get<SiPixelGainCalibrationForHLTGPU>();
    auto const& cpe = es.get<PixelCPEFast>();
    pxg_conditions c{cab.map.get(),     cab.modToUnp.data(), &gain.gain,   gain.data.data(),
                     uint32_t(gain.data.size()), &cpe.common, cpe.det.data(), uint32_t(cpe.det.size()),
                     &cpe.layer,        &cpe.average};
    pxg_context* ctx = nullptr;
    check(pxg_init(&c, &ctx), "pxg_init");
    es.put(std::make_unique<PxgDevice>(ctx));
  }
};
DEFINE_FWK_EVENTSETUP_MODULE(PxgDeviceESProducer);

// ---- event producers ---------------------------------------------------------------------

// BeamSpotToPOD.cc: the conditions beam spot as an event product
class BeamSpotToPOD : public edm::EDProducer {
public:
  explicit BeamSpotToPOD(edm::ProductRegistry& reg) : put_(reg.produces<BeamSpotPOD>()) {}
  void produce(edm::Event& ev, edm::EventSetup const& es) override { ev.emplace(put_, es.get<BeamSpotPOD>()); }

private:
  edm::EDPutTokenT<BeamSpotPOD> put_;
};
DEFINE_FWK_MODULE(BeamSpotToPOD);

// SiPixelRawToClusterCUDA.cc:32-149: FED loop on the host, then digis, clusters and errors
class SiPixelRawToClusterCUDA : public edm::EDProducer {
public:
  explicit SiPixelRawToClusterCUDA(edm::ProductRegistry& reg)
      : raw_(reg.consumes<FEDRawDataCollection>()),
        digis_(reg.produces<SiPixelDigisSoA>()),
        clusters_(reg.produces<SiPixelClustersSoA>()),
        errors_(reg.produces<SiPixelDigiErrorsSoA>()),
        words_(pxf::kMaxFedWords),
        fedIds_(pxf::kMaxFedWords / 2) {}
  ~SiPixelRawToClusterCUDA() override { pxg_stream_destroy(stream_); }

  void produce(edm::Event& ev, edm::EventSetup const& es) override {
    if (!stream_)
      check(pxg_stream_create(es.get<PxgDevice>().get(), 1, pxf::kMaxFedWords, &stream_), "pxg_stream_create");
    auto const& fedList = es.get<SiPixelFedIds>().ids;
    auto const& raw = ev.get(raw_).record;
    uint32_t nWords = 0;
    if (pxh_unpack_raw_record(raw.data(), raw.size(), fedList.data(), uint32_t(fedList.size()), words_.data(),
                              fedIds_.data(), uint32_t(words_.size()), &nWords) != 0)
      throw std::runtime_error("SiPixelRawToClusterCUDA: malformed raw data");
    pxg_digis d;
    pxg_clusters c;
    pxg_errors e;
    const int st = pxg_raw_to_cluster(stream_, words_.data(), fedIds_.data(), nWords, &d, &c, &e);
    check(st, "pxg_raw_to_cluster");
    ev.emplace(digis_, SiPixelDigisSoA{stream_, d});
    ev.emplace(clusters_, SiPixelClustersSoA{stream_, c});
    ev.emplace(errors_, SiPixelDigiErrorsSoA{stream_, e, st});
  }

private:
  edm::EDGetTokenT<FEDRawDataCollection> raw_;
  edm::EDPutTokenT<SiPixelDigisSoA> digis_;
  edm::EDPutTokenT<SiPixelClustersSoA> clusters_;
  edm::EDPutTokenT<SiPixelDigiErrorsSoA> errors_;
  pxg_stream* stream_ = nullptr;
  std::vector<uint32_t> words_;
  std::vector<uint8_t> fedIds_;
};
DEFINE_FWK_MODULE(SiPixelRawToClusterCUDA);

// SiPixelRecHitCUDA.cc:22-52
class SiPixelRecHitCUDA : public edm::EDProducer {
public:
  explicit SiPixelRecHitCUDA(edm::ProductRegistry& reg)
      : bs_(reg.consumes<BeamSpotPOD>()),
        digis_(reg.consumes<SiPixelDigisSoA>()),
        clusters_(reg.consumes<SiPixelClustersSoA>()),
        hits_(reg.produces<TrackingRecHit2DCPU>()) {}

  void produce(edm::Event& ev, edm::EventSetup const&) override {
    auto const& d = ev.get(digis_);
    auto const& c = ev.get(clusters_);
    pxg_hits h;
    check(pxg_rechits(d.stream, &d.view, &c.view, &ev.get(bs_).bs, &h), "pxg_rechits");
    ev.emplace(hits_, TrackingRecHit2DCPU{d.stream, h});
  }

private:
  edm::EDGetTokenT<BeamSpotPOD> bs_;
  edm::EDGetTokenT<SiPixelDigisSoA> digis_;
  edm::EDGetTokenT<SiPixelClustersSoA> clusters_;
  edm::EDPutTokenT<TrackingRecHit2DCPU> hits_;
};
DEFINE_FWK_MODULE(SiPixelRecHitCUDA);

// CAHitNtupletCUDA.cc:19-47
class CAHitNtupletCUDA : public edm::EDProducer {
public:
  explicit CAHitNtupletCUDA(edm::ProductRegistry& reg)
      : hits_(reg.consumes<TrackingRecHit2DCPU>()), tracks_(reg.produces<PixelTrackHeterogeneous>()) {}

  void produce(edm::Event& ev, edm::EventSetup const&) override {
    const float bf = 0.0114256972711507f;  // 1/fieldInGeV, CAHitNtupletCUDA.cc:31
    auto const& h = ev.get(hits_);
    pxg_tracks t;
    const int st = pxg_ntuplets(h.stream, &h.view, bf, &t);
    check(st, "pxg_ntuplets");
    ev.emplace(tracks_, PixelTrackHeterogeneous{h.stream, t, st});
  }

private:
  edm::EDGetTokenT<TrackingRecHit2DCPU> hits_;
  edm::EDPutTokenT<PixelTrackHeterogeneous> tracks_;
};
DEFINE_FWK_MODULE(CAHitNtupletCUDA);

// PixelVertexProducerCUDA.cc:17-40
class PixelVertexProducerCUDA : public edm::EDProducer {
public:
  explicit PixelVertexProducerCUDA(edm::ProductRegistry& reg)
      : tracks_(reg.consumes<PixelTrackHeterogeneous>()), vertices_(reg.produces<ZVertexHeterogeneous>()) {}

  void produce(edm::Event& ev, edm::EventSetup const&) override {
    auto const& t = ev.get(tracks_);
    pxg_vertices v;
    check(pxg_find_vertices(t.stream, &t.view, ptMin_, &v), "pxg_find_vertices");
    ev.emplace(vertices_, ZVertexHeterogeneous{t.stream, v});
  }

private:
  const float ptMin_ = 0.5f;  // m_ptMin, PixelVertexProducerCUDA.cc:37
  edm::EDGetTokenT<PixelTrackHeterogeneous> tracks_;
  edm::EDPutTokenT<ZVertexHeterogeneous> vertices_;
};
DEFINE_FWK_MODULE(PixelVertexProducerCUDA);

// Per-event counts (modules, digis, clusters, tracks of loose quality, vertices) for
// regression between runs; plays the role of CountValidator without expected values.
class CountLogger : public edm::EDProducer {
public:
  explicit CountLogger(edm::ProductRegistry& reg)
      : digis_(reg.consumes<SiPixelDigisSoA>()),
        clusters_(reg.consumes<SiPixelClustersSoA>()),
        tracks_(reg.consumes<PixelTrackHeterogeneous>()),
        vertices_(reg.consumes<ZVertexHeterogeneous>()) {}

  void produce(edm::Event& ev, edm::EventSetup const&) override {
    auto const& d = ev.get(digis_);
    auto const& t = ev.get(tracks_);
    std::vector<uint8_t> q(t.view.nTuples);
    if (!q.empty())
      check(pxg_copy_to_host(t.stream, q.data(), t.view.quality, q.size()), "pxg_copy_to_host");
    uint32_t loose = 0;
    for (uint8_t x : q)
      loose += x == pxf::kQualityLoose;
    char line[160];
    std::snprintf(line, sizeof(line), "%d %u %u %u %u %u %u\n", ev.eventID(), d.nModules(), d.nDigis(),
                  ev.get(clusters_).nClusters(), t.view.nTuples, loose, ev.get(vertices_).nvFinal());
    std::lock_guard<std::mutex> g(mutex());
    lines().emplace_back(ev.eventID(), line);
  }

  static std::mutex& mutex() {
    static std::mutex m;
    return m;
  }
  static std::vector<std::pair<int, std::string>>& lines() {
    static std::vector<std::pair<int, std::string>> l;
    return l;
  }

private:
  edm::EDGetTokenT<SiPixelDigisSoA> digis_;
  edm::EDGetTokenT<SiPixelClustersSoA> clusters_;
  edm::EDGetTokenT<PixelTrackHeterogeneous> tracks_;
  edm::EDGetTokenT<ZVertexHeterogeneous> vertices_;
};
DEFINE_FWK_MODULE(CountLogger);

namespace edm {
  // sorted per-event count lines written by CountLogger
  std::vector<std::string> countLines() {
    std::lock_guard<std::mutex> g(CountLogger::mutex());
    auto l = CountLogger::lines();
    std::sort(l.begin(), l.end());
    std::vector<std::string> out;
    for (auto& x : l)
      out.push_back(x.second);
    return out;
  }
}  // namespace edm

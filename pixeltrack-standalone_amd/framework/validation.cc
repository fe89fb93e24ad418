// Transfer and validation modules of the job: the device->host copies the
// CUDA program makes with --transfer (PixelTrackSoAFromCUDA.cc,
// PixelVertexSoAFromCUDA.cc, SiPixelDigisSoAFromCUDA.cc), the per-event count
// check against the expected counts shipped with the data
// (plugin-Validation/CountValidator.cc) and the histograms of the products
// (plugin-Validation/HistoValidator.cc, SimpleAtomicHisto.h).
#include <atomic>
#include <cmath>
#include <fstream>
#include <iostream>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "DataFormats.h"
#include "Framework.h"

namespace {
  void checkRc(int rc, char const* what) {
    if (rc < 0)
      throw std::runtime_error(std::string(what) + ": " + pxg_last_error());
  }

  template <typename T>
  std::vector<T> toHost(pxg_stream* s, const T* src, size_t n) {
    std::vector<T> v(n);
    if (n)
      checkRc(pxg_copy_to_host(s, v.data(), src, n * sizeof(T)), "pxg_copy_to_host");
    return v;
  }
}  // namespace

// ---- transfers ---------------------------------------------------------------------------

// SiPixelDigisSoAFromCUDA.cc: the digi arrays of the event on the host
class SiPixelDigisSoAFromCUDA : public edm::EDProducer {
public:
  explicit SiPixelDigisSoAFromCUDA(edm::ProductRegistry& reg)
      : in_(reg.consumes<SiPixelDigisSoA>()), out_(reg.produces<SiPixelDigisSoAHost>()) {}
  void produce(edm::Event& ev, edm::EventSetup const&) override {
    auto const& d = ev.get(in_);
    SiPixelDigisSoAHost h;
    h.nModules = d.nModules();
    h.nDigis = d.nDigis();
    h.adc = toHost(d.stream, d.view.adc, h.nDigis);
    h.clus = toHost(d.stream, d.view.clus, h.nDigis);
    h.pdigi = toHost(d.stream, d.view.pdigi, h.nDigis);
    h.rawIdArr = toHost(d.stream, d.view.rawIdArr, h.nDigis);
    ev.emplace(out_, std::move(h));
  }

private:
  edm::EDGetTokenT<SiPixelDigisSoA> in_;
  edm::EDPutTokenT<SiPixelDigisSoAHost> out_;
};
DEFINE_FWK_MODULE(SiPixelDigisSoAFromCUDA);

// PixelTrackSoAFromCUDA.cc: the stored tuples and their fit results on the host
class PixelTrackSoAFromCUDA : public edm::EDProducer {
public:
  explicit PixelTrackSoAFromCUDA(edm::ProductRegistry& reg)
      : in_(reg.consumes<PixelTrackHeterogeneous>()), out_(reg.produces<PixelTrackSoAHost>()) {}
  void produce(edm::Event& ev, edm::EventSetup const&) override {
    auto const& t = ev.get(in_);
    auto const& v = t.view;
    PixelTrackSoAHost h;
    h.nTracks = v.nTuples;
    h.hitOff = toHost(t.stream, v.hitOff, size_t(v.nTuples) + 1);
    h.hitIndices = toHost(t.stream, v.hitIndices, h.hitOff.back());
    h.quality = toHost(t.stream, v.quality, v.nTuples);
    h.chi2 = toHost(t.stream, v.chi2, v.nTuples);
    h.eta = toHost(t.stream, v.eta, v.nTuples);
    h.pt = toHost(t.stream, v.pt, v.nTuples);
    h.state.resize(5 * size_t(v.nTuples));
    for (int k = 0; k < 5 && v.nTuples; ++k)
      checkRc(pxg_copy_to_host(t.stream, h.state.data() + size_t(k) * v.nTuples, v.state + size_t(k) * v.stride,
                               sizeof(float) * v.nTuples),
              "pxg_copy_to_host");
    ev.emplace(out_, std::move(h));
  }

private:
  edm::EDGetTokenT<PixelTrackHeterogeneous> in_;
  edm::EDPutTokenT<PixelTrackSoAHost> out_;
};
DEFINE_FWK_MODULE(PixelTrackSoAFromCUDA);

// PixelVertexSoAFromCUDA.cc: the final vertices on the host
class PixelVertexSoAFromCUDA : public edm::EDProducer {
public:
  explicit PixelVertexSoAFromCUDA(edm::ProductRegistry& reg)
      : in_(reg.consumes<ZVertexHeterogeneous>()), out_(reg.produces<ZVertexSoAHost>()) {}
  void produce(edm::Event& ev, edm::EventSetup const&) override {
    auto const& z = ev.get(in_);
    auto const& v = z.view;
    ZVertexSoAHost h;
    h.nvFinal = v.nvFinal;
    h.zv = toHost(z.stream, v.zv, v.nvFinal);
    h.wv = toHost(z.stream, v.wv, v.nvFinal);
    h.chi2 = toHost(z.stream, v.chi2, v.nvFinal);
    h.ptv2 = toHost(z.stream, v.ptv2, v.nvFinal);
    h.ndof = toHost(z.stream, v.ndof, v.nvFinal);
    h.sortInd = toHost(z.stream, v.sortInd, v.nvFinal);
    ev.emplace(out_, std::move(h));
  }

private:
  edm::EDGetTokenT<ZVertexHeterogeneous> in_;
  edm::EDPutTokenT<ZVertexSoAHost> out_;
};
DEFINE_FWK_MODULE(PixelVertexSoAFromCUDA);

// ---- CountValidator ------------------------------------------------------------------------

namespace {
  std::atomic<int> allEvents{0}, goodEvents{0}, sumVertexDifference{0};
  std::mutex trackDifferenceMutex;
  float sumTrackDifference = 0.f;
}  // namespace

// CountValidator.cc:46-127: modules, digis and clusters exact; tracks (tuples
// with hits) within a relative 0.012; final vertices within one
class CountValidator : public edm::EDProducer {
public:
  explicit CountValidator(edm::ProductRegistry& reg)
      : digiClusterCount_(reg.consumes<DigiClusterCount>()),
        trackCount_(reg.consumes<TrackCount>()),
        vertexCount_(reg.consumes<VertexCount>()),
        digis_(reg.consumes<SiPixelDigisSoA>()),
        clusters_(reg.consumes<SiPixelClustersSoA>()),
        tracks_(reg.consumes<PixelTrackSoAHost>()),
        vertices_(reg.consumes<ZVertexSoAHost>()) {}

  void produce(edm::Event& ev, edm::EventSetup const&) override {
    constexpr float trackTolerance = 0.012f;
    constexpr int vertexTolerance = 1;
    std::stringstream ss;
    bool ok = true;
    ss << "Event " << ev.eventID() << " ";
    {
      auto const& count = ev.get(digiClusterCount_);
      auto const& digis = ev.get(digis_);
      auto const& clusters = ev.get(clusters_);
      if (digis.nModules() != count.nModules()) {
        ss << "\n N(modules) is " << digis.nModules() << " expected " << count.nModules();
        ok = false;
      }
      if (digis.nDigis() != count.nDigis()) {
        ss << "\n N(digis) is " << digis.nDigis() << " expected " << count.nDigis();
        ok = false;
      }
      if (clusters.nClusters() != count.nClusters()) {
        ss << "\n N(clusters) is " << clusters.nClusters() << " expected " << count.nClusters();
        ok = false;
      }
    }
    {
      auto const& count = ev.get(trackCount_);
      auto const& tracks = ev.get(tracks_);
      int nTracks = 0;
      for (uint32_t i = 0; i < tracks.nTracks; ++i)
        nTracks += tracks.nHits(i) > 0;
      const float rel = count.nTracks() ? std::abs(float(nTracks - int(count.nTracks())) / count.nTracks())
                                        : float(nTracks != 0);
      if (uint32_t(nTracks) != count.nTracks()) {
        std::lock_guard<std::mutex> g(trackDifferenceMutex);
        sumTrackDifference += rel;
      }
      if (rel >= trackTolerance) {
        ss << "\n N(tracks) is " << nTracks << " expected " << count.nTracks() << ", relative difference " << rel
           << " is outside tolerance " << trackTolerance;
        ok = false;
      }
    }
    {
      auto const& count = ev.get(vertexCount_);
      auto const& vertices = ev.get(vertices_);
      const int diff = std::abs(int(vertices.nvFinal) - int(count.nVertices()));
      sumVertexDifference += diff;
      if (diff > vertexTolerance) {
        ss << "\n N(vertices) is " << vertices.nvFinal << " expected " << count.nVertices() << ", difference "
           << diff << " is outside tolerance " << vertexTolerance;
        ok = false;
      }
    }
    ++allEvents;
    if (ok)
      ++goodEvents;
    else
      std::cout << ss.str() << std::endl;
  }

  void endJob() override {
    if (reported_.exchange(true))
      return;  // one report for all stream instances
    if (allEvents == goodEvents) {
      std::cout << "CountValidator: all " << allEvents << " events passed validation\n";
      if (sumTrackDifference != 0.f)
        std::cout << " Average relative track difference " << sumTrackDifference / allEvents.load()
                  << " (all within tolerance)\n";
      if (sumVertexDifference != 0)
        std::cout << " Average absolute vertex difference " << float(sumVertexDifference.load()) / allEvents.load()
                  << " (all within tolerance)\n";
    } else {
      std::cout << "CountValidator: " << (allEvents - goodEvents) << " events failed validation (see details above)\n";
      throw std::runtime_error("CountValidator failed");
    }
  }

private:
  static inline std::atomic<bool> reported_{false};
  edm::EDGetTokenT<DigiClusterCount> digiClusterCount_;
  edm::EDGetTokenT<TrackCount> trackCount_;
  edm::EDGetTokenT<VertexCount> vertexCount_;
  edm::EDGetTokenT<SiPixelDigisSoA> digis_;
  edm::EDGetTokenT<SiPixelClustersSoA> clusters_;
  edm::EDGetTokenT<PixelTrackSoAHost> tracks_;
  edm::EDGetTokenT<ZVertexSoAHost> vertices_;
};
DEFINE_FWK_MODULE(CountValidator);

// ---- HistoValidator ------------------------------------------------------------------------

// Fixed-range histogram with under- and overflow bins, filled from many
// threads (SimpleAtomicHisto.h): bin 0 below min, the last bin at or above
// max, and a value that rounds onto the upper edge goes into the last regular bin.
class AtomicHisto {
public:
  AtomicHisto(int nbins, float min, float max) : data_(size_t(nbins) + 2), min_(min), max_(max) {}
  AtomicHisto(AtomicHisto const& o) : data_(o.data_.size()), min_(o.min_), max_(o.max_) {}

  void fill(float value) {
    const int n = int(data_.size()) - 2;
    int i;
    if (value < min_) {
      i = 0;
    } else if (value >= max_) {
      i = n + 1;
    } else {
      i = int((value - min_) / (max_ - min_) * n);
      if (i == n)
        i = n - 1;
      if (i < 0 || i >= n)
        throw std::runtime_error("histogram fill(" + std::to_string(value) + ") out of range");
      ++i;
    }
    data_[size_t(i)].fetch_add(1, std::memory_order_relaxed);
  }

  // "nbins+2 min max count...", the format the histogram comparison reads
  friend std::ostream& operator<<(std::ostream& os, AtomicHisto const& h) {
    os << h.data_.size() << " " << h.min_ << " " << h.max_;
    for (auto const& c : h.data_)
      os << " " << c.load();
    return os;
  }

private:
  std::vector<std::atomic<int>> data_;
  float min_, max_;
};

// HistoValidator.cc:21-151: the 33 histograms of digis, clusters, hits, loose
// tracks and vertices; written to histograms_mi355x.txt at the end of the job
class HistoValidator : public edm::EDProducer {
public:
  explicit HistoValidator(edm::ProductRegistry& reg)
      : digis_(reg.consumes<SiPixelDigisSoA>()),
        clusters_(reg.consumes<SiPixelClustersSoA>()),
        hits_(reg.consumes<TrackingRecHit2DCPU>()),
        tracks_(reg.consumes<PixelTrackSoAHost>()),
        vertices_(reg.consumes<ZVertexSoAHost>()) {}

  void produce(edm::Event& ev, edm::EventSetup const&) override {
    auto const& d = ev.get(digis_);
    auto const& c = ev.get(clusters_);
    auto const& hv = ev.get(hits_);
    auto& H = histos();
    const uint32_t nDigis = d.nDigis(), nModules = d.nModules();
    H.at("digi_n").fill(nDigis);
    for (uint16_t a : toHost(d.stream, d.view.adc, nDigis))
      H.at("digi_adc").fill(a);
    H.at("module_n").fill(nModules);
    H.at("cluster_n").fill(c.nClusters());
    for (uint32_t n : toHost(c.stream, c.view.clusInModule, nModules))
      H.at("cluster_per_module_n").fill(n);

    const uint32_t nHits = hv.nHits();
    auto const& h = hv.view;
    const auto xl = toHost(hv.stream, h.xl, nHits), yl = toHost(hv.stream, h.yl, nHits);
    const auto xe = toHost(hv.stream, h.xerr, nHits), ye = toHost(hv.stream, h.yerr, nHits);
    const auto xg = toHost(hv.stream, h.xg, nHits), yg = toHost(hv.stream, h.yg, nHits);
    const auto zg = toHost(hv.stream, h.zg, nHits), rg = toHost(hv.stream, h.rg, nHits);
    const auto q = toHost(hv.stream, h.charge, nHits);
    const auto sx = toHost(hv.stream, h.xsize, nHits), sy = toHost(hv.stream, h.ysize, nHits);
    H.at("hit_n").fill(nHits);
    for (uint32_t i = 0; i < nHits; ++i) {
      H.at("hit_lx").fill(xl[i]);
      H.at("hit_ly").fill(yl[i]);
      H.at("hit_lex").fill(xe[i]);
      H.at("hit_ley").fill(ye[i]);
      H.at("hit_gx").fill(xg[i]);
      H.at("hit_gy").fill(yg[i]);
      H.at("hit_gz").fill(zg[i]);
      H.at("hit_gr").fill(rg[i]);
      H.at("hit_charge").fill(q[i]);
      H.at("hit_sizex").fill(sx[i]);
      H.at("hit_sizey").fill(sy[i]);
    }

    auto const& t = ev.get(tracks_);
    int nTracks = 0;
    for (uint32_t i = 0; i < t.nTracks; ++i) {
      if (t.nHits(i) > 0 && t.quality[i] >= pxf::kQualityLoose) {
        ++nTracks;
        H.at("track_nhits").fill(t.nHits(i));
        H.at("track_chi2").fill(t.chi2[i]);
        H.at("track_pt").fill(t.pt[i]);
        H.at("track_eta").fill(t.eta[i]);
        H.at("track_phi").fill(t.phi(i));
        H.at("track_tip").fill(t.tip(i));
        H.at("track_tip_zoom").fill(t.tip(i));
        H.at("track_zip").fill(t.zip(i));
        H.at("track_zip_zoom").fill(t.zip(i));
        H.at("track_quality").fill(t.quality[i]);
      }
    }
    H.at("track_n").fill(nTracks);

    auto const& v = ev.get(vertices_);
    H.at("vertex_n").fill(v.nvFinal);
    for (uint32_t i = 0; i < v.nvFinal; ++i) {
      H.at("vertex_z").fill(v.zv[i]);
      H.at("vertex_chi2").fill(v.chi2[i]);
      H.at("vertex_ndof").fill(v.ndof[i]);
      H.at("vertex_pt2").fill(v.ptv2[i]);
    }
  }

  void endJob() override {
    if (written_.exchange(true))
      return;
    std::ofstream out("histograms_mi355x.txt");
    for (auto const& [name, h] : histos())
      out << name << " " << h << "\n";
  }

  static std::map<std::string, AtomicHisto>& histos() {
    // the reference's binning (HistoValidator.cc:29-66)
    static std::map<std::string, AtomicHisto> h = {
        {"digi_n", AtomicHisto(100, 0, 1e5)},
        {"digi_adc", AtomicHisto(250, 0, 5e4)},
        {"module_n", AtomicHisto(100, 1500, 2000)},
        {"cluster_n", AtomicHisto(200, 5000, 25000)},
        {"cluster_per_module_n", AtomicHisto(110, 0, 110)},
        {"hit_n", AtomicHisto(200, 5000, 25000)},
        {"hit_lx", AtomicHisto(200, -1, 1)},
        {"hit_ly", AtomicHisto(800, -4, 4)},
        {"hit_lex", AtomicHisto(100, 0, 5e-5)},
        {"hit_ley", AtomicHisto(100, 0, 1e-4)},
        {"hit_gx", AtomicHisto(200, -20, 20)},
        {"hit_gy", AtomicHisto(200, -20, 20)},
        {"hit_gz", AtomicHisto(600, -60, 60)},
        {"hit_gr", AtomicHisto(200, 0, 20)},
        {"hit_charge", AtomicHisto(400, 0, 4e6)},
        {"hit_sizex", AtomicHisto(800, 0, 800)},
        {"hit_sizey", AtomicHisto(800, 0, 800)},
        {"track_n", AtomicHisto(150, 0, 15000)},
        {"track_nhits", AtomicHisto(3, 3, 6)},
        {"track_chi2", AtomicHisto(100, 0, 40)},
        {"track_pt", AtomicHisto(400, 0, 400)},
        {"track_eta", AtomicHisto(100, -3, 3)},
        {"track_phi", AtomicHisto(100, -3.15, 3.15)},
        {"track_tip", AtomicHisto(100, -1, 1)},
        {"track_tip_zoom", AtomicHisto(100, -0.05, 0.05)},
        {"track_zip", AtomicHisto(100, -15, 15)},
        {"track_zip_zoom", AtomicHisto(100, -0.1, 0.1)},
        {"track_quality", AtomicHisto(6, 0, 6)},
        {"vertex_n", AtomicHisto(60, 0, 60)},
        {"vertex_z", AtomicHisto(100, -15, 15)},
        {"vertex_chi2", AtomicHisto(100, 0, 40)},
        {"vertex_ndof", AtomicHisto(170, 0, 170)},
        {"vertex_pt2", AtomicHisto(100, 0, 4000)}};
    return h;
  }

private:
  static inline std::atomic<bool> written_{false};
  edm::EDGetTokenT<SiPixelDigisSoA> digis_;
  edm::EDGetTokenT<SiPixelClustersSoA> clusters_;
  edm::EDGetTokenT<TrackingRecHit2DCPU> hits_;
  edm::EDGetTokenT<PixelTrackSoAHost> tracks_;
  edm::EDGetTokenT<ZVertexSoAHost> vertices_;
};
DEFINE_FWK_MODULE(HistoValidator);

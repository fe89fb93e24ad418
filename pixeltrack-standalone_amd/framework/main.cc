// Job driver with the reference programs' command line and report
// (serial bin/main.cc:14-205; --transfer as in the CUDA program's
// bin/main.cc:76-146): reads raw.bin and the conditions from the data
// directory (and, with --validation, the expected counts), runs the pixel
// chain modules on N concurrent EDM streams (one thread each, one event at a
// time per stream), and prints the throughput line of the reference.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <mutex>
#include <thread>

#include "DataFormats.h"
#include "Framework.h"

namespace edm {
  PluginFactory& PluginFactory::instance() {
    static PluginFactory f;
    return f;
  }

  std::unique_ptr<EDProducer> PluginFactory::createED(std::string const& name, ProductRegistry& reg) const {
    auto it = ed_.find(name);
    if (it == ed_.end())
      throw std::runtime_error("no module named " + name);
    return it->second(reg);
  }

  std::unique_ptr<ESProducer> PluginFactory::createES(std::string const& name, std::string const& dir) const {
    auto it = es_.find(name);
    if (it == es_.end())
      throw std::runtime_error("no conditions module named " + name);
    return it->second(dir);
  }

  std::vector<std::string> countLines();  // plugins.cc

  // raw.bin (bin/Source.cc:9-21,50-73): per event u32 nfeds, then {u32 fedId, u32 size, bytes}
  std::vector<FEDRawDataCollection> readRawFile(std::string const& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in)
      throw std::runtime_error("cannot open " + path);
    std::vector<FEDRawDataCollection> events;
    uint32_t nfeds;
    while (in.read(reinterpret_cast<char*>(&nfeds), 4)) {
      FEDRawDataCollection c;
      c.record.resize(4);
      std::memcpy(c.record.data(), &nfeds, 4);
      for (uint32_t f = 0; f < nfeds; ++f) {
        uint32_t hdr[2];
        if (!in.read(reinterpret_cast<char*>(hdr), 8))
          throw std::runtime_error("truncated " + path);
        const size_t at = c.record.size();
        c.record.resize(at + 8 + hdr[1]);
        std::memcpy(c.record.data() + at, hdr, 8);
        if (!in.read(reinterpret_cast<char*>(c.record.data() + at + 8), hdr[1]))
          throw std::runtime_error("truncated " + path);
      }
      events.push_back(std::move(c));
    }
    return events;
  }

  // one u32 count per event and per file (bin/Source.cc:58-68): digicluster.bin holds
  // modules, digis and clusters, tracks.bin the tracks, vertices.bin the vertices
  std::vector<uint32_t> readCounts(std::string const& path, size_t perEvent, size_t events) {
    std::ifstream in(path, std::ios::binary);
    if (!in)
      throw std::runtime_error("cannot open " + path);
    std::vector<uint32_t> v(perEvent * events);
    if (!in.read(reinterpret_cast<char*>(v.data()), std::streamsize(v.size() * 4)))
      throw std::runtime_error("truncated " + path);
    return v;
  }

  class EventProcessor {
  public:
    EventProcessor(int maxEvents, int streams, std::vector<std::string> const& ed, std::vector<std::string> const& es,
                   std::string const& dataDir, bool validation)
        : raw_(readRawFile(dataDir + "/raw.bin")), streams_(streams), validation_(validation) {
      if (raw_.empty())
        throw std::runtime_error("no events in " + dataDir + "/raw.bin");
      maxEvents_ = maxEvents < 0 ? int(raw_.size()) : maxEvents;
      if (validation_) {
        digiClusters_ = readCounts(dataDir + "/digicluster.bin", 3, raw_.size());
        tracks_ = readCounts(dataDir + "/tracks.bin", 1, raw_.size());
        vertices_ = readCounts(dataDir + "/vertices.bin", 1, raw_.size());
      }
      for (auto const& name : es)
        PluginFactory::instance().createES(name, dataDir)->produce(setup_);
      // one instance of every module per stream, each stream with the same registry layout
      modules_.resize(streams);
      for (int s = 0; s < streams; ++s) {
        ProductRegistry reg;
        auto raw = reg.produces<FEDRawDataCollection>();
        EDPutTokenT<DigiClusterCount> dc;
        EDPutTokenT<TrackCount> tc;
        EDPutTokenT<VertexCount> vc;
        if (validation_) {
          dc = reg.produces<DigiClusterCount>();
          tc = reg.produces<TrackCount>();
          vc = reg.produces<VertexCount>();
        }
        for (auto const& name : ed)
          modules_[s].push_back(PluginFactory::instance().createED(name, reg));
        if (s == 0) {
          registry_ = reg;
          rawToken_ = raw;
          digiClusterToken_ = dc;
          trackToken_ = tc;
          vertexToken_ = vc;
        }
      }
    }

    int maxEvents() const { return maxEvents_; }
    int processedEvents() const { return processed_; }

    // runs events [next, limit) over the streams, or until `minutes` have passed
    // if it is not negative; events cycle through raw.bin
    void run(int limit, int minutes = -1) {
      const auto start = std::chrono::steady_clock::now();
      std::atomic<bool> timeUp{false};
      std::vector<std::thread> threads;
      std::atomic<bool> failed{false};
      std::string error;
      std::mutex errMutex;
      for (int s = 0; s < streams_; ++s) {
        threads.emplace_back([&, s] {
          try {
            for (;;) {
              // take the next event id, leaving ids at or past the limit for the next run
              int id = next_.load();
              do {
                if (id >= limit || timeUp)
                  return;
              } while (!next_.compare_exchange_weak(id, id + 1));
              if (failed)
                break;
              // the clock is read once per pass over raw.bin (bin/Source.cc:115-128)
              if (minutes >= 0 && id % int(raw_.size()) == 0 &&
                  std::chrono::steady_clock::now() - start >= std::chrono::minutes(minutes)) {
                timeUp = true;
                return;
              }
              Event ev(s, id, registry_);
              const size_t index = size_t(id) % raw_.size();
              ev.emplace(rawToken_, raw_[index]);
              if (validation_) {
                ev.emplace(digiClusterToken_, digiClusters_[3 * index], digiClusters_[3 * index + 1],
                           digiClusters_[3 * index + 2]);
                ev.emplace(trackToken_, tracks_[index]);
                ev.emplace(vertexToken_, vertices_[index]);
              }
              for (auto& m : modules_[s])
                m->produce(ev, setup_);
              ++processed_;
            }
          } catch (std::exception const& e) {
            std::lock_guard<std::mutex> g(errMutex);
            error = e.what();
            failed = true;
          }
        });
      }
      for (auto& t : threads)
        t.join();
      if (failed)
        throw std::runtime_error(error);
    }

    void endJob() {
      for (auto& ms : modules_)
        for (auto& m : ms)
          m->endJob();
      modules_.clear();
    }

  private:
    std::vector<FEDRawDataCollection> raw_;
    int streams_;
    bool validation_;
    std::vector<uint32_t> digiClusters_, tracks_, vertices_;
    EDPutTokenT<DigiClusterCount> digiClusterToken_;
    EDPutTokenT<TrackCount> trackToken_;
    EDPutTokenT<VertexCount> vertexToken_;
    int maxEvents_ = 0;
    EventSetup setup_;
    ProductRegistry registry_;
    EDPutTokenT<FEDRawDataCollection> rawToken_;
    std::vector<std::vector<std::unique_ptr<EDProducer>>> modules_;
    std::atomic<int> next_{0};
    std::atomic<int> processed_{0};

  public:
    void resetCount() { processed_ = 0; }
  };
}  // namespace edm

namespace {
  void printHelp(char const* name) {
    std::cout << "Usage: " << name
              << " [--numberOfThreads NT] [--numberOfStreams NS] [--warmupEvents WE] [--maxEvents ME]"
                 " [--runForMinutes RM] [--data PATH] [--transfer] [--validation] [--histogram] [--counts FILE]"
                 " [--empty]\n\n"
                 "  --numberOfThreads   threads (default 1; 0 = all CPU cores)\n"
                 "  --numberOfStreams   concurrent events (default 0 = numberOfThreads); one thread per stream\n"
                 "  --warmupEvents      events processed before the timed part (default 0)\n"
                 "  --maxEvents         events to process (default -1: all events of raw.bin; more cycle through it)\n"
                 "  --runForMinutes     cycle through raw.bin until this many minutes have passed\n"
                 "  --data              data directory with raw.bin and the conditions files\n"
                 "  --transfer          copy the tracks and vertices to the host\n"
                 "  --validation        check the per-event counts against digicluster.bin, tracks.bin and\n"
                 "                      vertices.bin of the data directory (implies --transfer)\n"
                 "  --histogram         write histograms of the products to histograms_mi355x.txt (implies --transfer)\n"
                 "  --counts FILE       write per-event counts (event modules digis clusters tuples loose vertices)\n"
                 "  --empty             run no producers\n";
  }

  double cpuSeconds() {
    timespec ts;
    clock_gettime(CLOCK_PROCESS_CPUTIME_ID, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
  }
}  // namespace

int main(int argc, char** argv) {
  int threads = 1, streams = 0, warmup = 0, maxEvents = -1, runForMinutes = -1;
  std::string data = "data", counts;
  bool empty = false, transfer = false, validation = false, histogram = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) {
        std::cerr << a << " needs a value\n";
        std::exit(EXIT_FAILURE);
      }
      return argv[++i];
    };
    if (a == "-h" || a == "--help") {
      printHelp(argv[0]);
      return EXIT_SUCCESS;
    } else if (a == "--numberOfThreads") {
      threads = std::stoi(val());
    } else if (a == "--numberOfStreams") {
      streams = std::stoi(val());
    } else if (a == "--warmupEvents") {
      warmup = std::stoi(val());
    } else if (a == "--maxEvents") {
      maxEvents = std::stoi(val());
    } else if (a == "--runForMinutes") {
      runForMinutes = std::stoi(val());
    } else if (a == "--data") {
      data = val();
    } else if (a == "--transfer") {
      transfer = true;
    } else if (a == "--validation") {
      transfer = validation = true;
    } else if (a == "--histogram") {
      transfer = histogram = true;
    } else if (a == "--counts") {
      counts = val();
    } else if (a == "--empty") {
      empty = true;
    } else {
      std::cout << "Invalid parameter " << a << "\n\n";
      printHelp(argv[0]);
      return EXIT_FAILURE;
    }
  }
  if (maxEvents >= 0 && runForMinutes >= 0) {
    std::cout << "Got both --maxEvents and --runForMinutes, please give only one of them" << std::endl;
    return EXIT_FAILURE;
  }
  if (threads == 0)
    threads = int(std::thread::hardware_concurrency());
  if (streams == 0)
    streams = threads;
  std::vector<std::string> ed, es;
  if (!empty) {
    ed = {"BeamSpotToPOD", "SiPixelRawToClusterCUDA", "SiPixelRecHitCUDA", "CAHitNtupletCUDA",
          "PixelVertexProducerCUDA"};
    es = {"BeamSpotESProducer", "SiPixelFedCablingMapGPUWrapperESProducer",
          "SiPixelGainCalibrationForHLTGPUESProducer", "PixelCPEFastESProducer", "PxgDeviceESProducer"};
    if (transfer) {  // after the producer of what they copy
      ed.insert(std::find(ed.begin(), ed.end(), "CAHitNtupletCUDA") + 1, "PixelTrackSoAFromCUDA");
      ed.insert(std::find(ed.begin(), ed.end(), "PixelVertexProducerCUDA") + 1, "PixelVertexSoAFromCUDA");
    }
    if (validation)
      ed.push_back("CountValidator");
    if (histogram)
      ed.push_back("HistoValidator");
    if (!counts.empty())
      ed.push_back("CountLogger");
  }
  try {
    edm::EventProcessor proc(maxEvents, streams, ed, es, data, validation && !empty);
    if (runForMinutes < 0)
      std::cout << "Processing " << proc.maxEvents() << " events,";
    else
      std::cout << "Processing for about " << runForMinutes << " minutes,";
    if (warmup > 0)
      std::cout << " after " << warmup << " events of warm up,";
    std::cout << " with " << streams << " concurrent events and " << threads << " threads." << std::endl;
    proc.run(warmup);
    proc.resetCount();
    const double cpu0 = cpuSeconds();
    const auto t0 = std::chrono::steady_clock::now();
    if (runForMinutes < 0)
      proc.run(warmup + proc.maxEvents());
    else
      proc.run(std::numeric_limits<int>::max(), runForMinutes);
    const auto t1 = std::chrono::steady_clock::now();
    const double cpu = cpuSeconds() - cpu0;
    proc.endJob();
    const double time = std::chrono::duration<double>(t1 - t0).count();
    const int n = proc.processedEvents();
    std::cout << "Processed " << n << " events in " << std::scientific << time << " seconds, throughput "
              << std::defaultfloat << (n / time) << " events/s, CPU usage per thread: " << std::fixed
              << std::setprecision(1) << (cpu / time / threads * 100) << "%" << std::endl;
    if (!counts.empty()) {
      std::ofstream out(counts);
      for (auto const& l : edm::countLines())
        out << l;
    }
  } catch (std::exception const& e) {
    std::cout << "\n----------\nCaught std::exception\n" << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}

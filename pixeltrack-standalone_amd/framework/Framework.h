// Minimal event-processing framework with the reference's module interface
// (Framework/EDProducer.h, Event.h, EventSetup.h, ProductRegistry.h,
// PluginFactory.h of the serial program): modules declare what they produce
// and consume in their constructor, get products from the Event by token and
// put theirs by token, read conditions from the EventSetup, and are found by
// name.  One instance of every module per EDM stream; streams run on their
// own threads.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <typeindex>
#include <unordered_map>
#include <utility>
#include <vector>

namespace edm {

  using StreamID = int;

  template <typename T>
  class EDGetTokenT {
  public:
    EDGetTokenT() = default;
    explicit EDGetTokenT(unsigned i) : index_(i) {}
    unsigned index() const { return index_; }

  private:
    unsigned index_ = ~0u;
  };

  template <typename T>
  class EDPutTokenT {
  public:
    EDPutTokenT() = default;
    explicit EDPutTokenT(unsigned i) : index_(i) {}
    unsigned index() const { return index_; }

  private:
    unsigned index_ = ~0u;
  };

  // Product types are registered in module order; a consumer must come after its producer.
  class ProductRegistry {
  public:
    template <typename T>
    EDPutTokenT<T> produces() {
      const auto [it, fresh] = index_.try_emplace(std::type_index(typeid(T)), unsigned(index_.size()));
      if (!fresh)
        throw std::runtime_error(std::string("product ") + typeid(T).name() + " is produced twice");
      return EDPutTokenT<T>{it->second};
    }

    template <typename T>
    EDGetTokenT<T> consumes() const {
      auto it = index_.find(std::type_index(typeid(T)));
      if (it == index_.end())
        throw std::runtime_error(std::string("product ") + typeid(T).name() + " is consumed but not produced");
      return EDGetTokenT<T>{it->second};
    }

    size_t size() const { return index_.size(); }

  private:
    std::unordered_map<std::type_index, unsigned> index_;
  };

  class WrapperBase {
  public:
    virtual ~WrapperBase() = default;
  };

  template <typename T>
  class Wrapper : public WrapperBase {
  public:
    template <typename... Args>
    explicit Wrapper(Args&&... args) : obj_(std::forward<Args>(args)...) {}
    T const& product() const { return obj_; }

  private:
    T obj_;
  };

  // Owns the products of one event; products are moved in and read as const.
  class Event {
  public:
    Event(StreamID stream, int id, ProductRegistry const& reg) : stream_(stream), id_(id), products_(reg.size()) {}

    StreamID streamID() const { return stream_; }
    int eventID() const { return id_; }

    template <typename T>
    T const& get(EDGetTokenT<T> const& token) const {
      auto const& p = products_.at(token.index());
      if (!p)
        throw std::runtime_error(std::string("product ") + typeid(T).name() + " was not produced in this event");
      return static_cast<Wrapper<T> const&>(*p).product();
    }

    template <typename T, typename... Args>
    void emplace(EDPutTokenT<T> const& token, Args&&... args) {
      products_.at(token.index()) = std::make_unique<Wrapper<T>>(std::forward<Args>(args)...);
    }

  private:
    StreamID stream_;
    int id_;
    std::vector<std::unique_ptr<WrapperBase>> products_;
  };

  // Conditions, filled once by the ES producers before the first event, then read-only.
  class EventSetup {
  public:
    template <typename T>
    void put(std::unique_ptr<T> prod) {
      products_[std::type_index(typeid(T))] = std::shared_ptr<void>(std::move(prod));
    }

    template <typename T>
    T const& get() const {
      auto it = products_.find(std::type_index(typeid(T)));
      if (it == products_.end())
        throw std::runtime_error(std::string("conditions product ") + typeid(T).name() + " is not available");
      return *static_cast<T const*>(it->second.get());
    }

  private:
    std::map<std::type_index, std::shared_ptr<void>> products_;
  };

  class EDProducer {
  public:
    virtual ~EDProducer() = default;
    virtual void produce(Event& event, EventSetup const& setup) = 0;
    virtual void endJob() {}
  };

  class ESProducer {
  public:
    virtual ~ESProducer() = default;
    virtual void produce(EventSetup& setup) = 0;
  };

  // Modules by name (the reference resolves them through plugins.txt and dlopen).
  class PluginFactory {
  public:
    using EDMaker = std::function<std::unique_ptr<EDProducer>(ProductRegistry&)>;
    using ESMaker = std::function<std::unique_ptr<ESProducer>(std::string const& dataDir)>;

    static PluginFactory& instance();
    void addED(std::string const& name, EDMaker m) { ed_[name] = std::move(m); }
    void addES(std::string const& name, ESMaker m) { es_[name] = std::move(m); }
    std::unique_ptr<EDProducer> createED(std::string const& name, ProductRegistry& reg) const;
    std::unique_ptr<ESProducer> createES(std::string const& name, std::string const& dataDir) const;

  private:
    std::map<std::string, EDMaker> ed_;
    std::map<std::string, ESMaker> es_;
  };

  template <typename T>
  struct EDRegistrar {
    explicit EDRegistrar(char const* name) {
      PluginFactory::instance().addED(name, [](ProductRegistry& reg) { return std::make_unique<T>(reg); });
    }
  };

  template <typename T>
  struct ESRegistrar {
    explicit ESRegistrar(char const* name) {
      PluginFactory::instance().addES(name, [](std::string const& dir) { return std::make_unique<T>(dir); });
    }
  };

}  // namespace edm

#define PXF_CAT2(a, b) a##b
#define PXF_CAT(a, b) PXF_CAT2(a, b)
#define DEFINE_FWK_MODULE(type) static edm::EDRegistrar<type> PXF_CAT(edRegistrar_, __LINE__)(#type);
#define DEFINE_FWK_EVENTSETUP_MODULE(type) static edm::ESRegistrar<type> PXF_CAT(esRegistrar_, __LINE__)(#type);

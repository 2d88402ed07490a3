// Minimal plugin registry with the reference's semantics: a module exports
// `extern "C" CPluginBase *get_plugin_interface()` returning a chain of
// plugins (append_interface / next_interface, mia/core/plugin_base.hh:70-100);
// the handler keeps, per (type, name), the plugin with the higher priority and
// discards the other (mia/core/handler.cxx:70-120).  Loading a module with a
// higher-priority plugin of the same name is how a drop-in replaces a loaded
// plugin.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "nonrigid.hh"

namespace mia_amd {

class CPluginBase {
public:
    CPluginBase(const char *name, const char *type) : m_name(name), m_type(type) {}
    virtual ~CPluginBase() = default;
    const std::string &get_name() const { return m_name; }
    const std::string &get_type() const { return m_type; }
    void set_priority(unsigned p) { m_priority = p; }
    unsigned get_priority() const { return m_priority; }
    // link another plugin at the end of this chain
    void append_interface(CPluginBase *p);
    CPluginBase *next_interface() const { return m_next; }
    void detach() { m_next = nullptr; }
    virtual std::string get_descr() const = 0;

private:
    std::string m_name, m_type;
    unsigned m_priority = 0;
    CPluginBase *m_next = nullptr;
};

// type "2dtransform" / "3dtransform", name "spline"
class CSplineTransformCreatorPlugin : public CPluginBase {
public:
    explicit CSplineTransformCreatorPlugin(int dim);
    std::shared_ptr<CSplineTransformCreator> create(const std::string &descr) const;
    std::string get_descr() const override;

private:
    int m_dim;
};

// type "2dimage-cost" / "3dimage-cost", name "ssd"
class CSSDCostPlugin : public CPluginBase {
public:
    explicit CSSDCostPlugin(int dim);
    CSSDCost create(const std::string &descr) const;
    std::string get_descr() const override;
};

// type "2dfullcost" / "3dfullcost", name "image"
class CImageFullCostPlugin : public CPluginBase {
public:
    explicit CImageFullCostPlugin(int dim);
    PImageFullCost create(const std::string &descr, int device = 0) const;
    std::string get_descr() const override;

private:
    int m_dim;
};

// type "3dsplinepenalty", name "divcurl"
class CDivcurlPenaltyPlugin : public CPluginBase {
public:
    CDivcurlPenaltyPlugin();
    PSplinePenalty create(const std::string &descr, int device = 0) const;
    std::string get_descr() const override;
};

class CPluginHandler {
public:
    // the handler with this library's plugins registered
    static CPluginHandler &instance();
    ~CPluginHandler();
    // add a chain of plugins; `module` names where it came from
    void add_chain(CPluginBase *head, const std::string &module);
    // dlopen a module and add what its get_plugin_interface() returns;
    // throws std::runtime_error if it cannot be loaded
    void load_module(const std::string &path);
    // the plugin of (type, name) or std::invalid_argument
    CPluginBase *get(const std::string &type, const std::string &name) const;
    // module the kept plugin came from
    const std::string &get_module(const std::string &type, const std::string &name) const;
    std::vector<std::string> get_names(const std::string &type) const;

private:
    struct Entry {
        std::unique_ptr<CPluginBase> plugin;
        std::string module;
    };
    std::map<std::pair<std::string, std::string>, Entry> m_plugins;
    std::vector<void *> m_libs;
};

// descriptor -> product through the handler, dimension 2 or 3
std::shared_ptr<CSplineTransformCreator> produce_transform_creator(int dim, const std::string &descr);
PImageFullCost produce_full_cost(int dim, const std::string &descr, int device = 0);

} // namespace mia_amd

extern "C" mia_amd::CPluginBase *get_plugin_interface();

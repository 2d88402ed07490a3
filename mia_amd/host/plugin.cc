// Plugin registry and this library's plugins (declarations in plugin.hh).
#include "plugin.hh"

#include <dlfcn.h>

#include <stdexcept>

#include "descr.hh"

namespace mia_amd {

PSplinePenalty produce_spline_penalty(int dim, const std::string &descr, int device);

void CPluginBase::append_interface(CPluginBase *p)
{
    CPluginBase *t = this;
    while (t->m_next)
        t = t->m_next;
    t->m_next = p;
}

namespace {
const char *dim_type(int dim, const char *what)
{
    static const std::string names[] = {"2dtransform", "3dtransform", "2dimage-cost", "3dimage-cost",
                                        "2dfullcost",  "3dfullcost"};
    const std::string w(what);
    const int base = w == "transform" ? 0 : w == "image-cost" ? 2 : 4;
    return names[base + (dim == 3 ? 1 : 0)].c_str();
}
// plugins of this library take priority over the reference's (priority 0)
constexpr unsigned kPriority = 1;
} // namespace

CSplineTransformCreatorPlugin::CSplineTransformCreatorPlugin(int dim)
    : CPluginBase("spline", dim_type(dim, "transform")), m_dim(dim)
{
    set_priority(kPriority);
}

std::shared_ptr<CSplineTransformCreator> CSplineTransformCreatorPlugin::create(const std::string &descr) const
{
    return std::make_shared<CSplineTransformCreator>(m_dim, descr);
}

std::string CSplineTransformCreatorPlugin::get_descr() const
{
    return "Free-form transformation that can be described by a set of B-spline coefficients and an underlying "
           "B-spline kernel (cubic kernel; GPU cost, gradient and warp).";
}

CSSDCostPlugin::CSSDCostPlugin(int dim) : CPluginBase("ssd", dim_type(dim, "image-cost"))
{
    set_priority(kPriority);
}

CSSDCost CSSDCostPlugin::create(const std::string &descr) const
{
    return CSSDCost::produce(descr);
}

std::string CSSDCostPlugin::get_descr() const
{
    return "Sum of squared differences image cost (norm, autothresh).";
}

CImageFullCostPlugin::CImageFullCostPlugin(int dim) : CPluginBase("image", dim_type(dim, "fullcost")), m_dim(dim)
{
    set_priority(kPriority);
}

PImageFullCost CImageFullCostPlugin::create(const std::string &descr, int device) const
{
    return CImageFullCost::produce(m_dim, descr, device);
}

std::string CImageFullCostPlugin::get_descr() const
{
    return "Image similarity full cost with multiresolution resampling; the cost kernel is ssd.";
}

CDivcurlPenaltyPlugin::CDivcurlPenaltyPlugin() : CPluginBase("divcurl", "3dsplinepenalty")
{
    set_priority(kPriority);
}

PSplinePenalty CDivcurlPenaltyPlugin::create(const std::string &descr, int device) const
{
    return produce_spline_penalty(3, descr, device);
}

std::string CDivcurlPenaltyPlugin::get_descr() const
{
    return "divcurl penalty on the transformation";
}

CPluginHandler &CPluginHandler::instance()
{
    static CPluginHandler h;
    static bool init = false;
    if (!init) {
        init = true;
        h.add_chain(get_plugin_interface(), "libmia_host");
    }
    return h;
}

CPluginHandler::~CPluginHandler()
{
    m_plugins.clear();
    // modules stay loaded until the process ends: their plugins' code may
    // still be referenced by products handed out
}

void CPluginHandler::add_chain(CPluginBase *head, const std::string &module)
{
    CPluginBase *p = head;
    while (p) {
        CPluginBase *next = p->next_interface();
        p->detach();
        const auto key = std::make_pair(p->get_type(), p->get_name());
        auto it = m_plugins.find(key);
        if (it == m_plugins.end()) {
            m_plugins[key] = Entry{std::unique_ptr<CPluginBase>(p), module};
        } else if (it->second.plugin->get_priority() < p->get_priority()) {
            it->second = Entry{std::unique_ptr<CPluginBase>(p), module};
        } else {
            delete p;
        }
        p = next;
    }
}

void CPluginHandler::load_module(const std::string &path)
{
    void *h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h)
        throw std::runtime_error("plugin module '" + path + "': " + dlerror());
    using Fn = CPluginBase *(*)();
    auto fn = (Fn)dlsym(h, "get_plugin_interface");
    if (!fn) {
        dlclose(h);
        throw std::runtime_error("plugin module '" + path + "' has no get_plugin_interface");
    }
    m_libs.push_back(h);
    add_chain(fn(), path);
}

CPluginBase *CPluginHandler::get(const std::string &type, const std::string &name) const
{
    auto it = m_plugins.find(std::make_pair(type, name));
    if (it == m_plugins.end())
        throw std::invalid_argument("no plugin '" + name + "' of type '" + type + "'");
    return it->second.plugin.get();
}

const std::string &CPluginHandler::get_module(const std::string &type, const std::string &name) const
{
    auto it = m_plugins.find(std::make_pair(type, name));
    if (it == m_plugins.end())
        throw std::invalid_argument("no plugin '" + name + "' of type '" + type + "'");
    return it->second.module;
}

std::vector<std::string> CPluginHandler::get_names(const std::string &type) const
{
    std::vector<std::string> r;
    for (const auto &kv : m_plugins)
        if (kv.first.first == type)
            r.push_back(kv.first.second);
    return r;
}

std::shared_ptr<CSplineTransformCreator> produce_transform_creator(int dim, const std::string &descr)
{
    const CParsedDescr d(descr);
    auto *p = dynamic_cast<CSplineTransformCreatorPlugin *>(
        CPluginHandler::instance().get(dim == 3 ? "3dtransform" : "2dtransform", d.name));
    if (!p)
        throw std::invalid_argument("plugin '" + d.name + "' is not a transformation creator");
    return p->create(descr);
}

PImageFullCost produce_full_cost(int dim, const std::string &descr, int device)
{
    const CParsedDescr d(descr);
    auto *p = dynamic_cast<CImageFullCostPlugin *>(
        CPluginHandler::instance().get(dim == 3 ? "3dfullcost" : "2dfullcost", d.name));
    if (!p)
        throw std::invalid_argument("plugin '" + d.name + "' is not a full cost");
    return p->create(descr, device);
}

} // namespace mia_amd

extern "C" mia_amd::CPluginBase *get_plugin_interface()
{
    using namespace mia_amd;
    auto *p = new CSplineTransformCreatorPlugin(2);
    p->append_interface(new CSplineTransformCreatorPlugin(3));
    p->append_interface(new CSSDCostPlugin(2));
    p->append_interface(new CSSDCostPlugin(3));
    p->append_interface(new CImageFullCostPlugin(2));
    p->append_interface(new CImageFullCostPlugin(3));
    p->append_interface(new CDivcurlPenaltyPlugin());
    return p;
}

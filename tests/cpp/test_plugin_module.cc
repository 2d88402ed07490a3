// A plugin module for the registry test: an `ssd` image cost of higher
// priority than the library's (it replaces it) and an `image` full cost of
// lower priority (the handler keeps the library's).
#include "../../mia_amd/host/plugin.hh"

namespace {
class TestSSD : public mia_amd::CPluginBase {
public:
    TestSSD() : CPluginBase("ssd", "3dimage-cost") { set_priority(5); }
    std::string get_descr() const override { return "test module ssd"; }
};
class TestImage : public mia_amd::CPluginBase {
public:
    TestImage() : CPluginBase("image", "3dfullcost") { set_priority(0); }
    std::string get_descr() const override { return "test module image"; }
};
} // namespace

extern "C" mia_amd::CPluginBase *get_plugin_interface()
{
    auto *p = new TestSSD;
    p->append_interface(new TestImage);
    return p;
}

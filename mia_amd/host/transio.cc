// Transformation files in the layout of MIA's "xml" transformation IO (see
// nonrigid.hh): the element sequence a Boost XML archive of
// serialization.hh's save() holds.
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "nonrigid.hh"
#include "plugin.hh"

namespace mia_amd {

namespace {
// the five predefined XML entities (what a Boost XML archive escapes)
std::string xml_escape(const std::string &s)
{
    std::string r;
    for (char c : s) {
        if (c == '&')
            r += "&amp;";
        else if (c == '<')
            r += "&lt;";
        else if (c == '>')
            r += "&gt;";
        else if (c == '"')
            r += "&quot;";
        else if (c == '\'')
            r += "&apos;";
        else
            r += c;
    }
    return r;
}

// inverse of xml_escape; an '&' without its ';' or an unknown entity is a
// malformed file
std::string xml_unescape(const std::string &s)
{
    static const char *const names[5][2] = {
        {"&amp;", "&"}, {"&lt;", "<"}, {"&gt;", ">"}, {"&quot;", "\""}, {"&apos;", "'"}};
    std::string r;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] != '&') {
            r += s[i];
            continue;
        }
        const size_t e = s.find(';', i);
        if (e == std::string::npos)
            throw std::runtime_error("load_transformation: malformed file, '&' without ';'");
        const std::string ent = s.substr(i, e - i + 1);
        bool known = false;
        for (const auto &n : names)
            if (ent == n[0]) {
                r += n[1];
                known = true;
                break;
            }
        if (!known)
            throw std::runtime_error("load_transformation: malformed file, unknown entity " + ent);
        i = e;
    }
    return r;
}

// text of the first <tag ...>...</tag> at or after pos
bool element(const std::string &x, const std::string &tag, size_t &pos, std::string &out)
{
    // the tag itself, not a longer name starting with it (item vs item_version)
    size_t a = pos;
    for (;;) {
        a = x.find("<" + tag, a);
        if (a == std::string::npos)
            return false;
        const char next = a + tag.size() + 1 < x.size() ? x[a + tag.size() + 1] : 0;
        if (next == '>' || next == ' ')
            break;
        ++a;
    }
    const size_t b = x.find('>', a);
    const size_t c = x.find("</" + tag + ">", b);
    if (b == std::string::npos || c == std::string::npos)
        return false;
    out = x.substr(b + 1, c - b - 1);
    pos = c + tag.size() + 3;
    return true;
}
} // namespace

void save_transformation(const std::string &path, const CSplineTransformation &t)
{
    std::ofstream f(path);
    if (!f)
        throw std::runtime_error("save_transformation: cannot write '" + path + "'");
    const auto p = t.get_parameters();
    std::string creator = t.get_creator_string();
    if (creator.empty()) {
        // a transformation built without a creator: describe it by its rates
        const Vec3 &r = t.get_target_rate();
        std::ostringstream c;
        c << std::setprecision(9) << "spline:anisorate=[" << r[0] << "," << r[1];
        if (t.get_dim() == 3)
            c << "," << r[2];
        c << "]";
        if (t.get_image_degree() == 1)
            c << ",imgkernel=[bspline:d=1]";
        creator = c.str();
    }
    f << "<?xml version=\"1.0\" encoding=\"UTF-8\" standalone=\"yes\" ?>\n"
      << "<!DOCTYPE boost_serialization>\n"
      << "<boost_serialization signature=\"serialization::archive\" version=\"17\">\n"
      << "<creator>" << xml_escape(creator) << "</creator>\n"
      << "<size_x>" << t.get_size().x << "</size_x>\n"
      << "<size_y>" << t.get_size().y << "</size_y>\n";
    if (t.get_dim() == 3)
        f << "<size_z>" << t.get_size().z << "</size_z>\n";
    // attributes as strings, the std::map<string, string> of
    // serialization.hh:43-49: count, item_version, then (first, second) items
    const auto &attr = t.attributes();
    f << "<attributes class_id=\"0\" tracking_level=\"0\" version=\"0\">\n"
      << "\t<count>" << attr.size() << "</count>\n\t<item_version>0</item_version>\n";
    bool first_item = true;
    for (const auto &kv : attr) {
        f << (first_item ? "\t<item class_id=\"1\" tracking_level=\"0\" version=\"0\">\n" : "\t<item>\n")
          << "\t\t<first>" << xml_escape(kv.first) << "</first>\n"
          << "\t\t<second>" << xml_escape(kv.second) << "</second>\n\t</item>\n";
        first_item = false;
    }
    f << "</attributes>\n"
      << "<params>\n\t<count>" << p.size() << "</count>\n\t<item_version>0</item_version>\n";
    f << std::setprecision(17);
    for (double v : p)
        f << "\t<item>" << v << "</item>\n";
    f << "</params>\n</boost_serialization>\n";
    if (!f.good())
        throw std::runtime_error("save_transformation: write error on '" + path + "'");
}

PSplineTransformation load_transformation(const std::string &path, int dim)
{
    std::ifstream f(path);
    if (!f)
        throw std::runtime_error("load_transformation: cannot read '" + path + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string x = ss.str();
    size_t pos = 0;
    std::string creator, sx, sy, sz = "1", params, count;
    if (!element(x, "creator", pos, creator) || !element(x, "size_x", pos, sx) || !element(x, "size_y", pos, sy))
        throw std::runtime_error("load_transformation: '" + path + "' is not a transformation file");
    if (dim == 3 && !element(x, "size_z", pos, sz))
        throw std::runtime_error("load_transformation: '" + path + "' has no size_z (a 2D transformation?)");
    // attributes (serialization.hh:66-77: read as strings, restored after
    // the parameters)
    std::vector<std::pair<std::string, std::string>> attr;
    {
        std::string block, acount, item;
        size_t ap = pos;
        if (!element(x, "attributes", ap, block))
            throw std::runtime_error("load_transformation: '" + path + "' has no attributes");
        pos = ap;
        size_t bp = 0;
        if (!element(block, "count", bp, acount))
            throw std::runtime_error("load_transformation: attributes without count in '" + path + "'");
        while (element(block, "item", bp, item)) {
            size_t ip = 0;
            std::string k, v;
            if (!element(item, "first", ip, k) || !element(item, "second", ip, v))
                throw std::runtime_error("load_transformation: malformed attribute in '" + path + "'");
            attr.emplace_back(xml_unescape(k), xml_unescape(v));
        }
        if (attr.size() != std::stoul(acount))
            throw std::runtime_error("load_transformation: attribute count and items differ in '" + path + "'");
    }
    if (!element(x, "params", pos, params))
        throw std::runtime_error("load_transformation: '" + path + "' has no params");
    size_t pp = 0;
    if (!element(params, "count", pp, count))
        throw std::runtime_error("load_transformation: params without count in '" + path + "'");
    std::vector<double> v;
    std::string item;
    while (element(params, "item", pp, item))
        v.push_back(std::stod(item));
    if (v.size() != std::stoul(count))
        throw std::runtime_error("load_transformation: params count and items differ in '" + path + "'");
    const Size3 size{(unsigned)std::stoul(sx), (unsigned)std::stoul(sy), (unsigned)std::stoul(sz)};
    auto t = produce_transform_creator(dim, xml_unescape(creator))->create(size);
    if (t->degrees_of_freedom() != v.size())
        throw std::runtime_error("load_transformation: bogus input, expected " +
                                 std::to_string(t->degrees_of_freedom()) + " parameters, got " +
                                 std::to_string(v.size()));
    t->set_parameters(v);
    for (const auto &kv : attr)
        t->set_attribute(kv.first, kv.second);
    return t;
}

} // namespace mia_amd

// Plugin descriptors "name:key=value,key=[nested:descr],..." and typed
// parameter tables, with the reference's error behaviour: an unknown key, a
// malformed value or a value out of range throws std::invalid_argument
// (mia/core/factory.hh, mia/core/parameter.cc).
#pragma once
#include <functional>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mia_amd {

struct CParsedDescr {
    std::string name;
    std::vector<std::pair<std::string, std::string>> params; // in the order given

    // Split at the first ':' and at the commas outside brackets; a value in
    // brackets is a nested descriptor and is kept without them.
    explicit CParsedDescr(const std::string &descr)
    {
        const size_t c = descr.find(':');
        name = descr.substr(0, c);
        if (name.empty())
            throw std::invalid_argument("descriptor '" + descr + "' has no plugin name");
        if (c == std::string::npos)
            return;
        const std::string rest = descr.substr(c + 1);
        int depth = 0;
        std::string item;
        auto flush = [&]() {
            if (item.empty())
                return;
            const size_t e = item.find('=');
            if (e == std::string::npos || e == 0)
                throw std::invalid_argument("descriptor '" + descr + "': expected key=value, got '" + item + "'");
            std::string v = item.substr(e + 1);
            if (v.size() >= 2 && v.front() == '[' && v.back() == ']')
                v = v.substr(1, v.size() - 2);
            params.emplace_back(item.substr(0, e), v);
            item.clear();
        };
        for (char ch : rest) {
            if (ch == '[')
                ++depth;
            else if (ch == ']' && --depth < 0)
                throw std::invalid_argument("descriptor '" + descr + "': unbalanced ']'");
            if (ch == ',' && depth == 0)
                flush();
            else
                item += ch;
        }
        if (depth != 0)
            throw std::invalid_argument("descriptor '" + descr + "': unbalanced '['");
        flush();
    }
};

// A table of named, typed parameters with defaults.  set() parses a value
// string into the bound variable and checks its range.
class CParamTable {
public:
    void add_float(const std::string &key, float &v, float lo, float hi)
    {
        m_setters[key] = [&v, lo, hi, key](const std::string &s) {
            v = (float)to_double(key, s);
            if (!(v >= lo && v <= hi))
                throw std::invalid_argument("parameter '" + key + "'=" + s + " is out of range");
        };
    }
    void add_double(const std::string &key, double &v, double lo, double hi)
    {
        m_setters[key] = [&v, lo, hi, key](const std::string &s) {
            v = to_double(key, s);
            if (!(v >= lo && v <= hi))
                throw std::invalid_argument("parameter '" + key + "'=" + s + " is out of range");
        };
    }
    void add_int(const std::string &key, int &v, int lo, int hi)
    {
        m_setters[key] = [&v, lo, hi, key](const std::string &s) {
            size_t pos = 0;
            long r = 0;
            try {
                r = std::stol(s, &pos);
            } catch (const std::exception &) {
                pos = 0;
            }
            if (pos != s.size() || s.empty())
                throw std::invalid_argument("parameter '" + key + "': '" + s + "' is not an integer");
            if (r < lo || r > hi)
                throw std::invalid_argument("parameter '" + key + "'=" + s + " is out of range");
            v = (int)r;
        };
    }
    void add_bool(const std::string &key, bool &v)
    {
        m_setters[key] = [&v, key](const std::string &s) {
            if (s == "1" || s == "true" || s == "yes")
                v = true;
            else if (s == "0" || s == "false" || s == "no")
                v = false;
            else
                throw std::invalid_argument("parameter '" + key + "': '" + s + "' is not a boolean");
        };
    }
    void add_string(const std::string &key, std::string &v)
    {
        m_setters[key] = [&v](const std::string &s) { v = s; };
    }
    // x,y[,z] float vector
    void add_fvector(const std::string &key, std::vector<float> &v, size_t n)
    {
        m_setters[key] = [&v, n, key](const std::string &s) {
            std::vector<float> r;
            std::stringstream ss(s);
            std::string item;
            while (std::getline(ss, item, ','))
                r.push_back((float)to_double(key, item));
            if (r.size() != n)
                throw std::invalid_argument("parameter '" + key + "' expects " + std::to_string(n) + " values");
            v = r;
        };
    }

    void set(const std::string &plugin, const CParsedDescr &d)
    {
        for (const auto &kv : d.params) {
            auto it = m_setters.find(kv.first);
            if (it == m_setters.end())
                throw std::invalid_argument("plugin '" + plugin + "' has no parameter '" + kv.first + "'");
            it->second(kv.second);
        }
    }

private:
    static double to_double(const std::string &key, const std::string &s)
    {
        size_t pos = 0;
        double r = 0.0;
        try {
            r = std::stod(s, &pos);
        } catch (const std::exception &) {
            pos = 0;
        }
        if (pos != s.size() || s.empty())
            throw std::invalid_argument("parameter '" + key + "': '" + s + "' is not a number");
        return r;
    }
    std::map<std::string, std::function<void(const std::string &)>> m_setters;
};

} // namespace mia_amd

// C ABI of the multiresolution driver (include/mia_host.h): the C++ mirror
// behind a plain-pointer entry point for callers outside C++ (bench, Python).
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/mia_host.h"
#include "nonrigid.hh"
#include "plugin.hh"

namespace mia_amd {
namespace detail {
int host_fail(int code, const std::string &msg);
}
} // namespace mia_amd

using namespace mia_amd;

extern "C" int mia_host_nonrigid_register(int device, int dim, const uint32_t *size, const float *src,
                                          const float *ref, const char *transform, const char *cost,
                                          const char *minimizer, int levels, double *params_out, size_t params_cap,
                                          size_t *n_params, uint32_t *coeff_size, mia_host_level_stats *stats,
                                          int max_levels, double *final_cost)
{
    if ((dim != 2 && dim != 3) || !size || !src || !ref || !transform || !cost || !minimizer)
        return detail::host_fail(MIA_HIP_ERR_ARG, "nonrigid_register: bad arguments");
    try {
        const Size3 s{size[0], size[1], dim == 3 ? size[2] : 1u};
        auto a = std::make_shared<CImage>(dim, s, src);
        auto b = std::make_shared<CImage>(dim, s, ref);
        CFullCostList costs;
        std::string rest(cost);
        size_t pos;
        do {
            pos = rest.find('+');
            costs.push(produce_full_cost(dim, rest.substr(0, pos), device));
            if (pos != std::string::npos)
                rest = rest.substr(pos + 1);
        } while (pos != std::string::npos);
        auto creator = produce_transform_creator(dim, transform);
        CNonrigidRegister reg(costs, minimizer, *creator, levels);
        auto t = reg.run(std::move(a), std::move(b));
        const auto p = t->get_parameters();
        if (n_params)
            *n_params = p.size();
        if (params_out) {
            if (params_cap < p.size())
                return detail::host_fail(MIA_HIP_ERR_SIZE, "nonrigid_register: params_out too small");
            std::memcpy(params_out, p.data(), sizeof(double) * p.size());
        }
        if (coeff_size) {
            coeff_size[0] = t->get_coeff_size().x;
            coeff_size[1] = t->get_coeff_size().y;
            coeff_size[2] = t->get_coeff_size().z;
        }
        if (stats) {
            const auto &st = reg.get_level_stats();
            for (int l = 0; l < (int)st.size() && l < max_levels; ++l) {
                stats[l].size[0] = st[l].size.x;
                stats[l].size[1] = st[l].size.y;
                stats[l].size[2] = st[l].size.z;
                stats[l].coeff_size[0] = st[l].coeff_size.x;
                stats[l].coeff_size[1] = st[l].coeff_size.y;
                stats[l].coeff_size[2] = st[l].coeff_size.z;
                stats[l].evaluations = st[l].evaluations;
                stats[l].cost = st[l].cost;
                stats[l].seconds = st[l].seconds;
                stats[l].minimize_seconds = st[l].minimize_seconds;
                stats[l].loop_seconds = st[l].loop_seconds;
            }
        }
        if (final_cost) {
            costs.set_size(s);
            *final_cost = costs.cost_value(*t) + t->get_energy_penalty();
        }
        return MIA_HIP_OK;
    } catch (const std::invalid_argument &e) {
        return detail::host_fail(MIA_HIP_ERR_ARG, e.what());
    } catch (const std::exception &e) {
        return detail::host_fail(MIA_HIP_ERR_DEVICE, e.what());
    }
}

extern "C" int mia_host_transform_save(const char *path, int dim, const uint32_t *size, const char *creator,
                                       const double *params, size_t n, const char *attributes)
{
    if (!path || (dim != 2 && dim != 3) || !size || !creator || (!params && n))
        return detail::host_fail(MIA_HIP_ERR_ARG, "transform_save: bad arguments");
    try {
        auto t = produce_transform_creator(dim, creator)->create(Size3{size[0], size[1], dim == 3 ? size[2] : 1u});
        t->set_parameters(params, n);
        // "key\0value\0...\0": pairs up to the empty key
        for (const char *a = attributes; a && *a;) {
            const std::string key(a);
            a += key.size() + 1;
            const std::string value(a);
            a += value.size() + 1;
            t->set_attribute(key, value);
        }
        save_transformation(path, *t);
        return MIA_HIP_OK;
    } catch (const std::invalid_argument &e) {
        return detail::host_fail(MIA_HIP_ERR_ARG, e.what());
    } catch (const std::exception &e) {
        return detail::host_fail(MIA_HIP_ERR_SIZE, e.what());
    }
}

extern "C" int mia_host_transform_load(const char *path, int dim, uint32_t *size, uint32_t *coeff_size,
                                       double *params, size_t params_cap, size_t *n_params, char *creator,
                                       int creator_len, char *attributes, int attributes_len)
{
    if (!path || (dim != 2 && dim != 3))
        return detail::host_fail(MIA_HIP_ERR_ARG, "transform_load: bad arguments");
    try {
        auto t = load_transformation(path, dim);
        const auto p = t->get_parameters();
        if (n_params)
            *n_params = p.size();
        if (size) {
            size[0] = t->get_size().x;
            size[1] = t->get_size().y;
            size[2] = t->get_size().z;
        }
        if (coeff_size) {
            coeff_size[0] = t->get_coeff_size().x;
            coeff_size[1] = t->get_coeff_size().y;
            coeff_size[2] = t->get_coeff_size().z;
        }
        if (params) {
            if (params_cap < p.size())
                return detail::host_fail(MIA_HIP_ERR_SIZE, "transform_load: params buffer too small");
            std::memcpy(params, p.data(), sizeof(double) * p.size());
        }
        if (creator && creator_len > 0) {
            std::strncpy(creator, t->get_creator_string().c_str(), (size_t)creator_len - 1);
            creator[creator_len - 1] = 0;
        }
        if (attributes) {
            std::string packed;
            for (const auto &kv : t->attributes()) {
                packed.append(kv.first).push_back('\0');
                packed.append(kv.second).push_back('\0');
            }
            packed.push_back('\0');
            if (attributes_len < 0 || packed.size() > (size_t)attributes_len)
                return detail::host_fail(MIA_HIP_ERR_SIZE, "transform_load: attributes buffer too small (" +
                                                               std::to_string(packed.size()) + " bytes needed)");
            std::memcpy(attributes, packed.data(), packed.size());
        }
        return MIA_HIP_OK;
    } catch (const std::invalid_argument &e) {
        return detail::host_fail(MIA_HIP_ERR_ARG, e.what());
    } catch (const std::exception &e) {
        return detail::host_fail(MIA_HIP_ERR_SIZE, e.what());
    }
}

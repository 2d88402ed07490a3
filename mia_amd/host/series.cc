// Batched series registration over libmia_hip (include/mia_host.h).
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mia_hip.h"
#include "../../include/mia_host.h"
#include "minimizer.hh"

namespace mia_amd {

namespace detail {
thread_local std::string g_err;

int host_fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}
} // namespace detail

namespace {
using detail::g_err;

int fail(int code, const std::string &msg)
{
    return detail::host_fail(code, msg);
}

int hip_fail(int code)
{
    g_err = mia_hip_last_error();
    return code;
}

// name:key=value,key=value
bool parse_descr(const std::string &d, std::string &name, std::vector<std::pair<std::string, std::string>> &kv)
{
    const size_t c = d.find(':');
    name = d.substr(0, c);
    kv.clear();
    if (c == std::string::npos)
        return true;
    std::stringstream ss(d.substr(c + 1));
    std::string item;
    while (std::getline(ss, item, ',')) {
        const size_t e = item.find('=');
        if (e == std::string::npos)
            return false;
        kv.emplace_back(item.substr(0, e), item.substr(e + 1));
    }
    return true;
}
} // namespace

ReverseMinimizer *produce_minimizer(const std::string &descr, size_t n, std::string &err)
{
    std::string name;
    std::vector<std::pair<std::string, std::string>> kv;
    if (!parse_descr(descr, name, kv)) {
        err = "malformed minimizer descriptor '" + descr + "'";
        return nullptr;
    }
    try {
        if (name == "lbfgs") {
            int maxeval = 100, memory = 10;
            double gtol = 0.0, ftolr = 0.0, step = 1.0;
            for (auto &p : kv) {
                if (p.first == "maxiter")
                    maxeval = std::stoi(p.second);
                else if (p.first == "memory")
                    memory = std::stoi(p.second);
                else if (p.first == "gtol")
                    gtol = std::stod(p.second);
                else if (p.first == "ftolr")
                    ftolr = std::stod(p.second);
                else if (p.first == "step")
                    step = std::stod(p.second);
                else if (p.first == "host")
                    (void)std::stoi(p.second); // where it runs, not how
                else {
                    err = "lbfgs: unknown parameter '" + p.first + "'";
                    return nullptr;
                }
            }
            if (maxeval < 1 || memory < 1 || step <= 0.0) {
                err = "lbfgs: maxiter, memory and step must be positive";
                return nullptr;
            }
            return new LBFGSMinimizer(n, maxeval, memory, gtol, ftolr, step);
        }
        if (name == "gdsq") {
            unsigned maxiter = 100;
            double step = 0.1, scale = 2.0, xtol = 0.0, gtol = 0.0, ftolr = 0.0;
            for (auto &p : kv) {
                if (p.first == "maxiter")
                    maxiter = (unsigned)std::stoul(p.second);
                else if (p.first == "step")
                    step = std::stod(p.second);
                else if (p.first == "scale")
                    scale = std::stod(p.second);
                else if (p.first == "xtola")
                    xtol = std::stod(p.second);
                else if (p.first == "gtola")
                    gtol = std::stod(p.second);
                else if (p.first == "ftolr")
                    ftolr = std::stod(p.second);
                else {
                    err = "gdsq: unknown parameter '" + p.first + "'";
                    return nullptr;
                }
            }
            // parameter bounds of the reference plugin (gdsq.cc:135-146)
            if (maxiter < 1 || !(step > 0.0) || !(scale > 1.0) || xtol < 0.0 || gtol < 0.0 || ftolr < 0.0) {
                err = "gdsq: parameter out of range";
                return nullptr;
            }
            return new GDSQMinimizer(n, maxiter, step, scale, xtol, gtol, ftolr);
        }
    } catch (const std::exception &e) {
        err = std::string("bad minimizer parameter value: ") + e.what();
        return nullptr;
    }
    err = "unknown minimizer '" + name + "'";
    return nullptr;
}

// Joint linear map of src and ref to mean 0, standard deviation 1
// (FScaleFilterCreator, mia/template/nonrigidregister.cxx:147-184); identity
// when both images are constant.
static void normalize_pair(const float *a, const float *b, size_t n, std::vector<float> &oa, std::vector<float> &ob)
{
    double sum = 0.0, sum2 = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double fa = a[i], fb = b[i];
        sum += fa + fb;
        sum2 += fa * fa + fb * fb;
    }
    const double m = 2.0 * n;
    const double mean = sum / m;
    const double sigma = std::sqrt((sum2 - sum * sum / m) / (m - 1));
    oa.assign(a, a + n);
    ob.assign(b, b + n);
    if (sigma == 0.0)
        return;
    const double sa = 1.0 / sigma, sb = -mean / sigma;
    for (size_t i = 0; i < n; ++i) {
        oa[i] = (float)(a[i] * sa + sb);
        ob[i] = (float)(b[i] * sa + sb);
    }
}

} // namespace mia_amd

using namespace mia_amd;

// batches of MIA_HOST_DEVICE_SPLIT (one per stream of mia_hip_optimize_many)
constexpr int kSeriesParts = 3;

struct mia_host_series {
    int device = 0, dim = 2, n = 0, degree = 3;
    std::vector<uint32_t> sizes, csizes;
    std::vector<size_t> n_par, off;
    size_t n_total = 0;
    mia_hip_ctx *batch = nullptr;
    std::vector<mia_hip_ctx *> single;
    std::vector<char> frame_set;
    // MIA_HOST_DEVICE_SPLIT: the frames in kSeriesParts batches of about
    // equal pixel counts (none when there is one frame); frame i is instance
    // part_idx[i] of batch part_of[i]
    std::vector<mia_hip_ctx *> part;
    std::vector<std::vector<int>> part_frames;
    std::vector<int> part_of, part_idx;
};

bool mia_amd::device_lbfgs_params(const std::string &descr, int &maxeval, int &memory, double &step,
                                   std::string &err)
{
    std::string name;
    std::vector<std::pair<std::string, std::string>> kv;
    err.clear();
    if (!parse_descr(descr, name, kv)) {
        err = "malformed minimizer descriptor '" + descr + "'";
        return false;
    }
    if (name != "lbfgs")
        return false;
    maxeval = 100;
    memory = 10;
    step = 1.0;
    bool device = true;
    try {
        for (auto &p : kv) {
            if (p.first == "maxiter")
                maxeval = std::stoi(p.second);
            else if (p.first == "memory")
                memory = std::stoi(p.second);
            else if (p.first == "step")
                step = std::stod(p.second);
            else if (p.first == "host")
                device = device && std::stoi(p.second) == 0;
            else if (p.first == "gtol" || p.first == "ftolr")
                device = device && std::stod(p.second) == 0.0;
            else
                device = false;
        }
    } catch (const std::exception &e) {
        err = std::string("bad minimizer parameter value: ") + e.what();
        return false;
    }
    return device && maxeval >= 1 && memory >= 1 && memory <= 16 && step > 0.0;
}

// On-device L-BFGS through mia_hip_optimize: all frames in the batched context
// (one graph covers every frame), or each frame in its own context, one frame
// after the other; or, through mia_hip_optimize_many, the frames in
// kSeriesParts batches whose iterations overlap.  Only `lbfgs` with gtol = ftolr = 0 runs
// on the device.
static int run_device(mia_host_series *s, const char *minimizer, int mode, double *params_out, double *costs_out,
                      int *evals_out, double *seconds)
{
    int maxeval = 100, memory = 10;
    double step = 1.0;
    std::string err;
    if (!device_lbfgs_params(minimizer ? minimizer : "", maxeval, memory, step, err))
        return fail(MIA_HIP_ERR_ARG, err.empty() ? "series: the device modes run lbfgs (maxiter, memory, step; "
                                                   "gtol = ftolr = 0, memory <= 16)"
                                                 : "series: " + err);
    const int chunk = 10;
    int st;
    if (mode == MIA_HOST_DEVICE_SPLIT && !s->part.empty()) {
        const int np = (int)s->part.size();
        std::vector<std::vector<double>> x0(np), xo(np), co(np);
        std::vector<std::vector<int>> eo(np);
        std::vector<const double *> px0(np);
        std::vector<double *> pxo(np), pco(np);
        std::vector<int *> peo(np);
        for (int p = 0; p < np; ++p) {
            size_t n = 0, o = 0;
            const int last = (int)s->part_frames[p].size() - 1;
            if ((st = mia_hip_n_params(s->part[p], last, &n, &o)))
                return hip_fail(st);
            x0[p].assign(o + n, 0.0);
            xo[p].resize(o + n);
            co[p].resize(s->part_frames[p].size());
            eo[p].resize(s->part_frames[p].size());
            px0[p] = x0[p].data();
            pxo[p] = xo[p].data();
            pco[p] = co[p].data();
            peo[p] = eo[p].data();
            if ((st = mia_hip_set_cost(s->part[p], 0, 0.0f, 1.0, MIA_HIP_GRAD_REFERENCE)))
                return hip_fail(st);
        }
        if ((st = mia_hip_optimize_many(s->part.data(), np, px0.data(), maxeval, memory, step, chunk, pxo.data(),
                                        pco.data(), peo.data(), seconds)))
            return hip_fail(st);
        for (int i = 0; i < s->n; ++i) {
            const int p = s->part_of[i], j = s->part_idx[i];
            size_t n = 0, o = 0;
            if ((st = mia_hip_n_params(s->part[p], j, &n, &o)))
                return hip_fail(st);
            if (params_out)
                std::copy(xo[p].begin() + (long)o, xo[p].begin() + (long)(o + n), params_out + s->off[i]);
            if (costs_out)
                costs_out[i] = co[p][j];
            if (evals_out)
                evals_out[i] = eo[p][j];
        }
        return MIA_HIP_OK;
    }
    if (mode == MIA_HOST_DEVICE_BATCHED || mode == MIA_HOST_DEVICE_SPLIT) {
        std::vector<double> x0(s->n_total, 0.0);
        if ((st = mia_hip_set_cost(s->batch, 0, 0.0f, 1.0, MIA_HIP_GRAD_REFERENCE)))
            return hip_fail(st);
        if ((st = mia_hip_optimize(s->batch, x0.data(), maxeval, memory, step, chunk, params_out, costs_out, evals_out,
                                   seconds)))
            return hip_fail(st);
        return MIA_HIP_OK;
    }
    double total = 0.0;
    for (int i = 0; i < s->n; ++i) {
        mia_hip_ctx *c = s->single[i];
        std::vector<double> x0(s->n_par[i], 0.0);
        double sec = 0.0, cost = 0.0;
        int ev = 0;
        if ((st = mia_hip_set_cost(c, 0, 0.0f, 1.0, MIA_HIP_GRAD_REFERENCE)))
            return hip_fail(st);
        if ((st = mia_hip_optimize(c, x0.data(), maxeval, memory, step, chunk,
                                   params_out ? params_out + s->off[i] : nullptr, &cost, &ev, &sec)))
            return hip_fail(st);
        total += sec;
        if (costs_out)
            costs_out[i] = cost;
        if (evals_out)
            evals_out[i] = ev;
    }
    if (seconds)
        *seconds = total;
    return MIA_HIP_OK;
}

extern "C" {

const char *mia_host_last_error(void)
{
    return g_err.c_str();
}

int mia_host_series_destroy(mia_host_series *s)
{
    if (!s)
        return MIA_HIP_OK;
    if (s->batch)
        mia_hip_ctx_destroy(s->batch);
    for (auto *c : s->part)
        if (c)
            mia_hip_ctx_destroy(c);
    for (auto *c : s->single)
        if (c)
            mia_hip_ctx_destroy(c);
    delete s;
    return MIA_HIP_OK;
}

int mia_host_series_create(int device, int dim, int n_frames, const uint32_t *sizes, float rate, int degree,
                           mia_host_series **out)
{
    if (!out || !sizes || n_frames < 1 || (dim != 2 && dim != 3))
        return fail(MIA_HIP_ERR_ARG, "series: need dim 2 or 3, at least one frame and its sizes");
    // on a failure below, the contexts created so far go with the series
    auto s = std::unique_ptr<mia_host_series, int (*)(mia_host_series *)>(new mia_host_series(),
                                                                            mia_host_series_destroy);
    s->device = device;
    s->dim = dim;
    s->n = n_frames;
    s->degree = degree;
    s->sizes.assign(sizes, sizes + 3 * n_frames);
    s->csizes.assign(3 * n_frames, 1);
    for (int i = 0; i < n_frames; ++i) {
        if (dim == 2)
            s->sizes[3 * i + 2] = 1;
        for (int a = 0; a < dim; ++a) {
            const int cs = mia_hip_spline_csize(s->sizes[3 * i + a], rate, dim);
            if (cs < 0)
                return fail(MIA_HIP_ERR_ARG, "series: spline rate too large for frame " + std::to_string(i));
            s->csizes[3 * i + a] = (uint32_t)cs;
        }
    }
    int st = mia_hip_ctx_create(device, &s->batch);
    if (st)
        return hip_fail(st);
    if ((st = mia_hip_setup(s->batch, dim, n_frames, s->sizes.data(), s->csizes.data(), degree)))
        return hip_fail(st);
    s->n_par.resize(n_frames);
    s->off.resize(n_frames);
    for (int i = 0; i < n_frames; ++i)
        if ((st = mia_hip_n_params(s->batch, i, &s->n_par[i], &s->off[i])))
            return hip_fail(st);
    s->n_total = s->off.back() + s->n_par.back();
    s->single.assign(n_frames, nullptr);
    s->frame_set.assign(n_frames, 0);
    const int n_parts = std::min(n_frames, kSeriesParts);
    if (n_parts >= 2) {
        // batches of about equal pixel counts: frames by decreasing size,
        // each to the batch with the fewest pixels so far (ties to the
        // first); a batch keeps its frames in ascending order
        std::vector<int> order(n_frames);
        for (int i = 0; i < n_frames; ++i)
            order[i] = i;
        auto pixels = [&](int i) { return (double)s->sizes[3 * i] * s->sizes[3 * i + 1] * s->sizes[3 * i + 2]; };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pixels(a) > pixels(b); });
        s->part_frames.assign(n_parts, {});
        std::vector<double> load(n_parts, 0.0);
        for (int i : order) {
            const int p = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            s->part_frames[p].push_back(i);
            load[p] += pixels(i);
        }
        s->part_of.assign(n_frames, 0);
        s->part_idx.assign(n_frames, 0);
        s->part.assign(n_parts, nullptr);
        for (int p = 0; p < n_parts; ++p) {
            auto &fr = s->part_frames[p];
            std::sort(fr.begin(), fr.end());
            std::vector<uint32_t> sz, cs;
            for (size_t j = 0; j < fr.size(); ++j) {
                s->part_of[fr[j]] = p;
                s->part_idx[fr[j]] = (int)j;
                sz.insert(sz.end(), &s->sizes[3 * fr[j]], &s->sizes[3 * fr[j]] + 3);
                cs.insert(cs.end(), &s->csizes[3 * fr[j]], &s->csizes[3 * fr[j]] + 3);
            }
            if ((st = mia_hip_ctx_create(device, &s->part[p])) ||
                (st = mia_hip_setup(s->part[p], dim, (int)fr.size(), sz.data(), cs.data(), degree)))
                return hip_fail(st);
        }
    }
    *out = s.release();
    return MIA_HIP_OK;
}

int mia_host_series_set_frame(mia_host_series *s, int frame, const float *src, const float *ref)
{
    if (!s || frame < 0 || frame >= s->n || !src || !ref)
        return fail(MIA_HIP_ERR_ARG, "series: bad frame index or null image");
    const size_t nv = (size_t)s->sizes[3 * frame] * s->sizes[3 * frame + 1] * s->sizes[3 * frame + 2];
    std::vector<float> a, b;
    normalize_pair(src, ref, nv, a, b);
    int st = mia_hip_set_images(s->batch, frame, a.data(), b.data());
    if (st)
        return hip_fail(st);
    if (!s->part.empty() && (st = mia_hip_set_images(s->part[s->part_of[frame]], s->part_idx[frame], a.data(), b.data())))
        return hip_fail(st);
    if (!s->single[frame]) {
        if ((st = mia_hip_ctx_create(s->device, &s->single[frame])))
            return hip_fail(st);
        if ((st = mia_hip_setup(s->single[frame], s->dim, 1, &s->sizes[3 * frame], &s->csizes[3 * frame],
                                s->degree)))
            return hip_fail(st);
    }
    if ((st = mia_hip_set_images(s->single[frame], 0, a.data(), b.data())))
        return hip_fail(st);
    s->frame_set[frame] = 1;
    return MIA_HIP_OK;
}

int mia_host_series_n_params(mia_host_series *s, int frame, size_t *n, size_t *offset)
{
    if (!s || frame < 0 || frame >= s->n)
        return fail(MIA_HIP_ERR_ARG, "series: bad frame index");
    if (n)
        *n = s->n_par[frame];
    if (offset)
        *offset = s->off[frame];
    return MIA_HIP_OK;
}

int mia_host_series_run(mia_host_series *s, const char *minimizer, int mode, double *params_out, double *costs_out,
                        int *evals_out, double *seconds)
{
    if (!s || !minimizer)
        return fail(MIA_HIP_ERR_ARG, "series: null argument");
    for (int i = 0; i < s->n; ++i)
        if (!s->frame_set[i])
            return fail(MIA_HIP_ERR_STATE, "series: frame " + std::to_string(i) + " has no images");
    if (mode == MIA_HOST_DEVICE_BATCHED || mode == MIA_HOST_DEVICE_PER_FRAME || mode == MIA_HOST_DEVICE_SPLIT)
        return run_device(s, minimizer, mode, params_out, costs_out, evals_out, seconds);
    std::vector<std::unique_ptr<ReverseMinimizer>> opt(s->n);
    for (int i = 0; i < s->n; ++i) {
        std::string err;
        opt[i].reset(produce_minimizer(minimizer, s->n_par[i], err));
        if (!opt[i])
            return fail(MIA_HIP_ERR_ARG, err);
        std::vector<double> x0(s->n_par[i], 0.0);
        opt[i]->start(x0.data());
    }
    int st;
    auto t0 = std::chrono::steady_clock::now();
    if (mode == MIA_HOST_BATCHED_GRAPH || mode == MIA_HOST_BATCHED) {
        if ((st = mia_hip_set_cost(s->batch, 0, 0.0f, 1.0, MIA_HIP_GRAD_REFERENCE)))
            return hip_fail(st);
        double *P = nullptr, *G = nullptr, *C = nullptr;
        std::vector<double> params(s->n_total), grad(s->n_total), costs(s->n);
        if (mode == MIA_HOST_BATCHED_GRAPH) {
            if ((st = mia_hip_graph_capture(s->batch)) || (st = mia_hip_graph_params(s->batch, &P)) ||
                (st = mia_hip_graph_outputs(s->batch, &G, &C)))
                return hip_fail(st);
        } else {
            P = params.data();
            G = grad.data();
            C = costs.data();
        }
        for (;;) {
            bool any = false;
            for (int i = 0; i < s->n; ++i) {
                const double *x = opt[i]->done() ? opt[i]->best_x() : opt[i]->point();
                if (!opt[i]->done())
                    any = true;
                if (x)
                    std::memcpy(P + s->off[i], x, sizeof(double) * s->n_par[i]);
                else
                    std::memset(P + s->off[i], 0, sizeof(double) * s->n_par[i]);
            }
            if (!any)
                break;
            st = mode == MIA_HOST_BATCHED_GRAPH ? mia_hip_graph_run(s->batch) : mia_hip_cost_grad(s->batch, P, G, C);
            if (st)
                return hip_fail(st);
            for (int i = 0; i < s->n; ++i)
                if (!opt[i]->done())
                    opt[i]->tell(C[i], G + s->off[i]);
        }
    } else if (mode == MIA_HOST_PER_FRAME) {
        for (int i = 0; i < s->n; ++i) {
            mia_hip_ctx *c = s->single[i];
            if ((st = mia_hip_set_cost(c, 0, 0.0f, 1.0, MIA_HIP_GRAD_REFERENCE)))
                return hip_fail(st);
            std::vector<double> g(s->n_par[i]);
            double cost = 0.0;
            while (!opt[i]->done()) {
                if ((st = mia_hip_cost_grad(c, opt[i]->point(), g.data(), &cost)))
                    return hip_fail(st);
                opt[i]->tell(cost, g.data());
            }
        }
    } else {
        return fail(MIA_HIP_ERR_ARG, "series: unknown run mode");
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds)
        *seconds = std::chrono::duration<double>(t1 - t0).count();
    for (int i = 0; i < s->n; ++i) {
        if (params_out && opt[i]->best_x())
            std::memcpy(params_out + s->off[i], opt[i]->best_x(), sizeof(double) * s->n_par[i]);
        if (costs_out)
            costs_out[i] = opt[i]->best_f();
        if (evals_out)
            evals_out[i] = opt[i]->evaluations();
    }
    return MIA_HIP_OK;
}

} // extern "C"

// Host mirror of the spline transformation, the image full cost with the ssd
// kernel, the cost list, the minimizer problem and the multiresolution
// driver (declarations and reference citations in nonrigid.hh).  The cost,
// gradient, warp, image resampling and penalty run on the GPU through
// libmia_hip; what stays on the host is per-level bookkeeping on coefficient
// grids (upscale, refine) and single-point queries.
#include "nonrigid.hh"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <sstream>
#include <thread>

#include "descr.hh"
#include "minimizer.hh"

namespace mia_amd {

ReverseMinimizer *produce_minimizer(const std::string &descr, size_t n, std::string &err);

void hip_check(int status, const char *what)
{
    if (status == MIA_HIP_OK)
        return;
    const std::string msg = std::string(what) + ": " + mia_hip_last_error();
    if (status == MIA_HIP_ERR_ARG || status == MIA_HIP_ERR_SIZE)
        throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

CDatapool &CDatapool::instance()
{
    static CDatapool pool;
    return pool;
}

PImage CDatapool::get(const std::string &key) const
{
    auto it = m_map.find(key);
    if (it == m_map.end() || !it->second)
        throw std::invalid_argument("CDatapool: no data for key '" + key + "'");
    return it->second;
}

namespace {

// floor after rounding to a multiple of 2^-16 (mia/core/splinekernel.cc:164-192)
double fastfloor16(double v)
{
    return std::floor(std::nearbyint(v * 65536.0) / 65536.0);
}

// cubic B-spline weights at fraction f (mia/core/splinekernel/bspline.cc:258-285)
void bspline3_weights(double f, double w[4])
{
    const double fm1 = 1.0 - f;
    w[3] = f * f * f / 6.0;
    w[0] = fm1 * fm1 * fm1 / 6.0;
    w[2] = f - 2.0 * w[3] + w[0];
    w[1] = fm1 - 2.0 * w[0] + w[3];
}

// start index and weights of the cubic kernel at x (splinekernel.cc:250-255)
int kernel_at(double x, double w[4])
{
    const double fl = fastfloor16(x);
    bspline3_weights(x - fl, w);
    return (int)fl - 1;
}

// mirror boundary index map (mia/core/splinebc/bc.cc:52-64)
int mirror_index(int i, int n)
{
    if (n == 1)
        return 0;
    const int n2 = 2 * n - 2;
    int idx = i < 0 ? -i : i;
    if (idx >= n2)
        idx %= n2;
    if (idx >= n)
        idx = n2 - idx;
    return idx;
}

// cubic B-spline prefilter of one line with mirror boundaries
// (mia/core/boundary_conditions.cc:62-103, mia/core/splinebc/bc.cc:67-88)
void prefilter_line(double *c, int n)
{
    if (n < 2)
        return;
    const double p = std::sqrt(3.0) - 2.0;
    const double lambda = 2.0 - p - 1.0 / p;
    for (int k = 0; k < n; ++k)
        c[k] *= lambda;
    double zn = p, z2n = std::pow(p, (double)(n - 1));
    const double iz = 1.0 / p;
    double sum = c[0] + z2n * c[n - 1];
    z2n *= z2n * iz;
    for (int k = 1; k < n - 1; ++k) {
        sum += (zn + z2n) * c[k];
        zn *= p;
        z2n *= iz;
    }
    c[0] = sum / (1.0 - zn * zn);
    for (int k = 1; k < n; ++k)
        c[k] += p * c[k - 1];
    c[n - 1] = (p / (p * p - 1.0)) * (p * c[n - 2] + c[n - 1]);
    for (int k = n - 2; k >= 0; --k)
        c[k] = p * (c[k + 1] - c[k]);
}

// coefficient grid for a rate; an axis whose grid would have at most one
// interior coefficient comes back as 0 (the caller throws)
Size3 spline_csize(int dim, Size3 range, const Vec3 &rate)
{
    auto one = [&](unsigned n, float r) {
        const int c = mia_hip_spline_csize(n, r, dim);
        return c < 0 ? 0u : (unsigned)c;
    };
    Size3 cs;
    cs.x = one(range.x, rate[0]);
    cs.y = one(range.y, rate[1]);
    cs.z = dim == 3 ? one(range.z, rate[2]) : 1u;
    return cs;
}

std::string dim_prefix(int dim)
{
    return dim == 2 ? "C2D" : "C3D";
}

} // namespace

// ---------------------------------------------------------------------------
// penalty
// ---------------------------------------------------------------------------

void CSplinePenalty::initialize(int dim, Size3 coeff_size, Vec3 range)
{
    m_dim = dim;
    m_size = coeff_size;
    m_range = range;
    do_initialize();
}

double CSplinePenalty::effective_weight() const
{
    const double vol = (double)m_range[0] * m_range[1] * (m_dim == 3 ? m_range[2] : 1.0f);
    return m_normalize ? m_weight / vol : m_weight;
}

double CSplinePenalty::value(const std::vector<float> &coefficients) const
{
    return effective_weight() * do_value(coefficients);
}

double CSplinePenalty::value_and_gradient(const std::vector<float> &coefficients, std::vector<double> &gradient) const
{
    const double w = effective_weight();
    gradient.assign(coefficients.size(), 0.0);
    const double v = w * do_value_and_gradient(coefficients, gradient);
    for (double &g : gradient)
        g = -w * g;
    return v;
}

struct CDivcurlSplinePenalty::Impl {
    mia_hip_divcurl *p = nullptr;
    ~Impl()
    {
        if (p)
            mia_hip_divcurl_destroy(p);
    }
};

CDivcurlSplinePenalty::CDivcurlSplinePenalty(double weight, bool normalize, double div, double curl, int device)
    : CSplinePenalty(weight, normalize), m_div(div), m_curl(curl), m_device(device), m_impl(new Impl)
{
    if (!(weight > 0.0))
        throw std::invalid_argument("divcurl: weight must be positive");
    if (!(div >= 0.0) || !(curl >= 0.0))
        throw std::invalid_argument("divcurl: div and curl must be non-negative");
}

CDivcurlSplinePenalty::~CDivcurlSplinePenalty() = default;

CSplinePenalty *CDivcurlSplinePenalty::clone() const
{
    auto *r = new CDivcurlSplinePenalty(get_weight(), get_normalize(), m_div, m_curl, m_device);
    if (m_impl->p)
        r->initialize(m_dim, m_size, m_range);
    return r;
}

mia_hip_divcurl *CDivcurlSplinePenalty::device_handle() const
{
    return m_impl->p;
}

void CDivcurlSplinePenalty::do_initialize()
{
    if (m_dim != 3)
        throw std::invalid_argument("divcurl: this build provides the 3D penalty");
    if (m_impl->p) {
        mia_hip_divcurl_destroy(m_impl->p);
        m_impl->p = nullptr;
    }
    const uint32_t cs[3] = {m_size.x, m_size.y, m_size.z};
    hip_check(mia_hip_divcurl_create(m_device, cs, m_range.data(), m_div, m_curl, &m_impl->p), "divcurl");
}

double CDivcurlSplinePenalty::do_value(const std::vector<float> &coefficients) const
{
    if (!m_impl->p)
        throw std::runtime_error("divcurl: penalty not initialized");
    double v = 0.0;
    hip_check(mia_hip_divcurl_evaluate(m_impl->p, coefficients.data(), &v, nullptr), "divcurl");
    return v;
}

double CDivcurlSplinePenalty::do_value_and_gradient(const std::vector<float> &coefficients,
                                                    std::vector<double> &gradient) const
{
    if (!m_impl->p)
        throw std::runtime_error("divcurl: penalty not initialized");
    gradient.resize(coefficients.size());
    double v = 0.0;
    hip_check(mia_hip_divcurl_evaluate(m_impl->p, coefficients.data(), &v, gradient.data()), "divcurl");
    return v;
}

CImage scale_image(const CImage &in, Size3 out_size, int device)
{
    if (in.dim == 2)
        out_size.z = 1;
    CImage out(in.dim, out_size);
    const uint32_t a[3] = {in.size.x, in.size.y, in.size.z};
    const uint32_t b[3] = {out_size.x, out_size.y, out_size.z};
    hip_check(mia_hip_scale_image(device, in.dim, a, in.data.data(), b, out.data.data()), "scale");
    return out;
}

// ---------------------------------------------------------------------------
// spline transformation
// ---------------------------------------------------------------------------

CSplineTransformation::CSplineTransformation(int dim, Size3 range, Vec3 rate, int img_degree, PSplinePenalty penalty)
    : m_dim(dim), m_range(range), m_rate(rate), m_img_degree(img_degree)
{
    if (dim != 2 && dim != 3)
        throw std::invalid_argument("CSplineTransformation: dimension must be 2 or 3");
    if (dim == 2)
        m_range.z = 1;
    if (img_degree != 1 && img_degree != 3)
        throw std::invalid_argument("CSplineTransformation: image interpolation supports bspline:d=1 and d=3");
    if (range.x == 0 || range.y == 0 || range.z == 0)
        throw std::invalid_argument("CSplineTransformation: empty image range");
    for (int a = 0; a < dim; ++a)
        if (!(rate[a] >= 1.0f))
            throw std::invalid_argument("CSplineTransformation: the coefficient rate must be >= 1");
    m_csize = spline_csize(dim, m_range, rate);
    for (int a = 0; a < dim; ++a)
        if (m_csize[a] <= 3) {
            std::ostringstream s;
            s << dim_prefix(dim) << "SplineTransformation: your coefficient rate [" << rate[0] << "," << rate[1];
            if (dim == 3)
                s << "," << rate[2];
            s << "] is too large of the current image dimensions, reduce the crate parameter of the "
                 "transformation or the number of multi-resolution levels if you run a registration algorithm.";
            throw std::invalid_argument(s.str());
        }
    m_coeffs.assign(m_csize.product() * dim, 0.0f);
    if (penalty)
        m_penalty.reset(penalty->clone());
}

CSplineTransformation::CSplineTransformation(const CSplineTransformation &org)
    : m_dim(org.m_dim), m_range(org.m_range), m_rate(org.m_rate), m_img_degree(org.m_img_degree),
      m_csize(org.m_csize), m_coeffs(org.m_coeffs), m_creator(org.m_creator), m_attributes(org.m_attributes)
{
    if (org.m_penalty)
        m_penalty.reset(org.m_penalty->clone());
}

std::vector<double> CSplineTransformation::get_parameters() const
{
    return std::vector<double>(m_coeffs.begin(), m_coeffs.end());
}

void CSplineTransformation::set_parameters(const double *params, size_t n)
{
    if (n != m_coeffs.size())
        throw std::invalid_argument(dim_prefix(m_dim) + "SplineTransformation::set_parameters: expected " +
                                    std::to_string(m_coeffs.size()) + " parameters, got " + std::to_string(n));
    for (size_t k = 0; k < n; ++k)
        m_coeffs[k] = (float)params[k];
}

void CSplineTransformation::set_parameters(const std::vector<double> &params)
{
    set_parameters(params.data(), params.size());
}

void CSplineTransformation::set_identity()
{
    std::fill(m_coeffs.begin(), m_coeffs.end(), 0.0f);
}

void CSplineTransformation::set_coefficients(const std::vector<float> &field, Size3 fsize)
{
    if (m_dim == 2)
        fsize.z = 1;
    for (int a = 0; a < m_dim; ++a)
        if (fsize[a] <= 3)
            throw std::invalid_argument(dim_prefix(m_dim) +
                                        "SplineTransformation::set_coefficients: your coefficient field is too "
                                        "small, dimensions must be larger than 3");
    if (field.size() != fsize.product() * m_dim)
        throw std::invalid_argument("set_coefficients: field size does not match its dimensions");
    m_csize = fsize;
    m_coeffs = field;
    m_penalty_valid = false;
}

void CSplineTransformation::set_coefficients_and_prefilter(const std::vector<float> &field, Size3 fsize)
{
    if (m_dim == 2)
        fsize.z = 1;
    if (field.size() != fsize.product() * m_dim)
        throw std::invalid_argument("set_coefficients_and_prefilter: field size does not match its dimensions");
    // separable prefilter of every component along x, y (and z)
    std::vector<double> c(field.begin(), field.end());
    const int n[3] = {(int)fsize.x, (int)fsize.y, (int)fsize.z};
    std::vector<double> line;
    for (int axis = 0; axis < m_dim; ++axis) {
        line.resize(n[axis]);
        const size_t stride = axis == 0 ? 1 : axis == 1 ? (size_t)n[0] : (size_t)n[0] * n[1];
        const int o1 = axis == 0 ? n[1] : n[0], o2 = axis == 2 ? n[1] : n[2];
        for (int b = 0; b < o2; ++b)
            for (int a = 0; a < o1; ++a) {
                size_t base;
                if (axis == 0)
                    base = ((size_t)b * n[1] + a) * n[0];
                else if (axis == 1)
                    base = (size_t)b * n[0] * n[1] + a;
                else
                    base = (size_t)b * n[0] + a;
                for (int comp = 0; comp < m_dim; ++comp) {
                    for (int k = 0; k < n[axis]; ++k)
                        line[k] = c[(base + k * stride) * m_dim + comp];
                    prefilter_line(line.data(), n[axis]);
                    for (int k = 0; k < n[axis]; ++k)
                        c[(base + k * stride) * m_dim + comp] = line[k];
                }
            }
    }
    set_coefficients(std::vector<float>(c.begin(), c.end()), fsize);
}

Vec3 CSplineTransformation::scale(const Vec3 &x) const
{
    Vec3 r{0.0f, 0.0f, 0.0f};
    for (int a = 0; a < m_dim; ++a) {
        const float s = (float)(m_csize[a] - 3) / (float)(m_range[a] - 1);
        r[a] = x[a] * s + 1.0f;
    }
    return r;
}

Vec3 CSplineTransformation::get_displacement_at(const Vec3 &x) const
{
    const Vec3 p = scale(x);
    double w[3][4];
    int st[3] = {0, 0, 0};
    for (int a = 0; a < m_dim; ++a)
        st[a] = kernel_at(p[a], w[a]);
    // sum over the 4^dim coefficients; reads beyond the grid are skipped
    // (spline.cc:433-466)
    Vec3 r{0.0f, 0.0f, 0.0f};
    const int nzc = m_dim == 3 ? 4 : 1;
    for (int c = 0; c < nzc; ++c) {
        const int kz = m_dim == 3 ? st[2] + c : 0;
        if (kz < 0 || kz >= (int)m_csize.z)
            continue;
        float hy[3] = {0.0f, 0.0f, 0.0f};
        for (int b = 0; b < 4; ++b) {
            const int ky = st[1] + b;
            if (ky < 0 || ky >= (int)m_csize.y)
                continue;
            float hx[3] = {0.0f, 0.0f, 0.0f};
            for (int a = 0; a < 4; ++a) {
                const int kx = st[0] + a;
                if (kx < 0 || kx >= (int)m_csize.x)
                    continue;
                const float *cf = &m_coeffs[(((size_t)kz * m_csize.y + ky) * m_csize.x + kx) * m_dim];
                for (int k = 0; k < m_dim; ++k)
                    hx[k] += (float)(w[0][a] * cf[k]);
            }
            for (int k = 0; k < m_dim; ++k)
                hy[k] += (float)(hx[k] * w[1][b]);
        }
        for (int k = 0; k < m_dim; ++k)
            r[k] += m_dim == 3 ? (float)(hy[k] * w[2][c]) : hy[k];
    }
    return r;
}

Vec3 CSplineTransformation::operator()(const Vec3 &x) const
{
    const Vec3 u = get_displacement_at(x);
    return Vec3{x[0] - u[0], x[1] - u[1], m_dim == 3 ? x[2] - u[2] : 0.0f};
}

CImage CSplineTransformation::operator()(const CImage &image, int device) const
{
    if (image.dim != m_dim || image.size != m_range)
        throw std::invalid_argument("CSplineTransformation: image and transformation sizes differ");
    mia_hip_ctx *ctx = nullptr;
    hip_check(mia_hip_ctx_create(device, &ctx), "warp: context");
    std::unique_ptr<mia_hip_ctx, int (*)(mia_hip_ctx *)> guard(ctx, mia_hip_ctx_destroy);
    const uint32_t sz[3] = {m_range.x, m_range.y, m_range.z};
    const uint32_t cs[3] = {m_csize.x, m_csize.y, m_csize.z};
    hip_check(mia_hip_setup(ctx, m_dim, 1, sz, cs, m_img_degree), "warp: setup");
    std::vector<float> zero(image.data.size(), 0.0f);
    hip_check(mia_hip_set_images(ctx, 0, image.data.data(), zero.data()), "warp: images");
    const std::vector<double> p = get_parameters();
    CImage out(m_dim, m_range);
    hip_check(mia_hip_warp(ctx, 0, p.data(), out.data.data()), "warp");
    return out;
}

std::shared_ptr<CSplineTransformation> CSplineTransformation::upscale(Size3 size) const
{
    if (m_dim == 2)
        size.z = 1;
    // the new transformation keeps this coefficient grid, with the
    // displacements scaled to the new image size: size / range in 3D
    // (spline.cc:399-412), (size - 1) / (range - 1) in 2D
    // (mia/2d/transform/spline.cc:356-370)
    float mx[3];
    for (int a = 0; a < 3; ++a)
        mx[a] = m_dim == 3 ? (float)size[a] / (float)m_range[a]
                           : (a < 2 ? ((float)size[a] - 1.0f) / ((float)m_range[a] - 1.0f) : 1.0f);
    auto help = std::make_shared<CSplineTransformation>(*this);
    help->m_range = size;
    std::vector<float> c(m_coeffs.size());
    for (size_t k = 0; k < c.size(); ++k)
        c[k] = mx[k % m_dim] * m_coeffs[k];
    help->set_coefficients(c, m_csize);
    help->m_rate = m_rate;
    return help;
}

bool CSplineTransformation::refine()
{
    // coefficient grid for the target rate at the current size
    // (spline.cc:363-365; 2D mia/2d/transform/spline.cc:321-322)
    Size3 cs;
    for (int a = 0; a < 3; ++a) {
        unsigned v = 1;
        if (a < m_dim)
            v = (unsigned)(((float)m_range[a] + m_rate[a] - 1.0f) / m_rate[a]) + 2u;
        (a == 0 ? cs.x : a == 1 ? cs.y : cs.z) = v;
    }
    bool grow = false;
    for (int a = 0; a < m_dim; ++a)
        grow |= cs[a] > m_csize[a];
    if (!grow)
        return false;
    // interpolate the old coefficient field, as an image of vectors, at the
    // new knots: prefilter, then cubic B-spline sums with mirror indices,
    // clamped per component to [min, max] of the field where min is set to
    // zero when all of its components are positive (mia/3d/interpolator.cxx:
    // 67-130,170-180; vector comparison mia/3d/vector.hh:541-544)
    const int d = m_dim;
    std::vector<double> c(m_coeffs.begin(), m_coeffs.end());
    double vmin[3] = {0, 0, 0}, vmax[3] = {0, 0, 0};
    for (int k = 0; k < d; ++k) {
        vmin[k] = std::numeric_limits<double>::infinity();
        vmax[k] = -std::numeric_limits<double>::infinity();
    }
    for (size_t i = 0; i < m_csize.product(); ++i)
        for (int k = 0; k < d; ++k) {
            vmin[k] = std::min(vmin[k], (double)m_coeffs[i * d + k]);
            vmax[k] = std::max(vmax[k], (double)m_coeffs[i * d + k]);
        }
    bool all_pos = true;
    for (int k = 0; k < d; ++k)
        all_pos &= 0.0 < vmin[k];
    if (all_pos)
        for (int k = 0; k < d; ++k)
            vmin[k] = 0.0;
    {
        const int n[3] = {(int)m_csize.x, (int)m_csize.y, (int)m_csize.z};
        std::vector<double> line;
        for (int axis = 0; axis < d; ++axis) {
            line.resize(n[axis]);
            const size_t stride = axis == 0 ? 1 : axis == 1 ? (size_t)n[0] : (size_t)n[0] * n[1];
            const int o1 = axis == 0 ? n[1] : n[0], o2 = axis == 2 ? n[1] : n[2];
            for (int b = 0; b < o2; ++b)
                for (int a = 0; a < o1; ++a) {
                    const size_t base = axis == 0   ? ((size_t)b * n[1] + a) * n[0]
                                        : axis == 1 ? (size_t)b * n[0] * n[1] + a
                                                    : (size_t)b * n[0] + a;
                    for (int comp = 0; comp < d; ++comp) {
                        for (int k = 0; k < n[axis]; ++k)
                            line[k] = c[(base + k * stride) * d + comp];
                        prefilter_line(line.data(), n[axis]);
                        for (int k = 0; k < n[axis]; ++k)
                            c[(base + k * stride) * d + comp] = line[k];
                    }
                }
        }
    }
    float dx[3] = {1.0f, 1.0f, 1.0f};
    for (int a = 0; a < d; ++a)
        dx[a] = (float)(m_csize[a] - 3) / (float)(cs[a] - 3);
    std::vector<float> out(cs.product() * d);
    size_t o = 0;
    for (unsigned z = 0; z < cs.z; ++z)
        for (unsigned y = 0; y < cs.y; ++y)
            for (unsigned x = 0; x < cs.x; ++x, ++o) {
                const unsigned X[3] = {x, y, z};
                double w[3][4];
                int idx[3][4];
                for (int a = 0; a < 3; ++a) {
                    if (a >= d) {
                        w[a][0] = 1.0;
                        idx[a][0] = 0;
                        continue;
                    }
                    const float p = dx[a] * ((float)X[a] - 1.0f) + 1.0f;
                    const int st = kernel_at(p, w[a]);
                    for (int k = 0; k < 4; ++k)
                        idx[a][k] = mirror_index(st + k, (int)m_csize[a]);
                }
                const int nz = d == 3 ? 4 : 1;
                for (int comp = 0; comp < d; ++comp) {
                    double v = 0.0;
                    for (int cz = 0; cz < nz; ++cz) {
                        double vy = 0.0;
                        for (int b = 0; b < 4; ++b) {
                            double vx = 0.0;
                            for (int a = 0; a < 4; ++a)
                                vx += w[0][a] *
                                      c[(((size_t)idx[2][cz] * m_csize.y + idx[1][b]) * m_csize.x + idx[0][a]) * d +
                                        comp];
                            vy += w[1][b] * vx;
                        }
                        v += w[2][cz] * vy;
                    }
                    v = std::min(std::max(v, vmin[comp]), vmax[comp]);
                    out[o * d + comp] = (float)v;
                }
            }
    set_coefficients(out, cs);
    return true;
}

void CSplineTransformation::reinit_penalty() const
{
    if (m_penalty && !m_penalty_valid) {
        m_penalty->initialize(m_dim, m_csize, Vec3{(float)m_range.x, (float)m_range.y, (float)m_range.z});
        m_penalty_valid = true;
    }
}

double CSplineTransformation::get_energy_penalty() const
{
    if (!m_penalty)
        return 0.0;
    reinit_penalty();
    return m_penalty->value(m_coeffs);
}

bool CSplineTransformation::device_penalty(mia_hip_divcurl **p, double *weight) const
{
    *p = nullptr;
    *weight = 0.0;
    if (!m_penalty)
        return true;
    reinit_penalty();
    *p = m_penalty->device_handle();
    *weight = m_penalty->effective_weight();
    return *p != nullptr;
}

double CSplineTransformation::get_energy_penalty_and_gradient(std::vector<double> &gradient) const
{
    if (!m_penalty) {
        gradient.assign(m_coeffs.size(), 0.0);
        return 0.0;
    }
    reinit_penalty();
    return m_penalty->value_and_gradient(m_coeffs, gradient);
}

// ---------------------------------------------------------------------------
// plugin products: spline creator, ssd, image full cost
// ---------------------------------------------------------------------------

PSplinePenalty produce_spline_penalty(int dim, const std::string &descr, int device)
{
    if (descr.empty())
        return {};
    CParsedDescr d(descr);
    if (d.name != "divcurl")
        throw std::invalid_argument("unknown spline penalty '" + d.name + "'");
    if (dim != 3)
        throw std::invalid_argument("divcurl: this build provides the 3D penalty");
    float weight = 1.0f, div = 1.0f, curl = 1.0f;
    bool norm = false;
    CParamTable t;
    t.add_float("weight", weight, std::numeric_limits<float>::min(), std::numeric_limits<float>::max());
    t.add_bool("norm", norm);
    t.add_float("div", div, 0.0f, std::numeric_limits<float>::max());
    t.add_float("curl", curl, 0.0f, std::numeric_limits<float>::max());
    t.set("divcurl", d);
    return std::make_shared<CDivcurlSplinePenalty>(weight, norm, div, curl, device);
}

CSplineTransformCreator::CSplineTransformCreator(int dim, const std::string &descr) : m_dim(dim), m_descr(descr)
{
    CParsedDescr d(descr);
    if (d.name != "spline")
        throw std::invalid_argument("this build provides the 'spline' transformation, not '" + d.name + "'");
    float rate = 10.0f;
    std::vector<float> aniso(dim, 0.0f);
    std::string kernel = "bspline:d=3", penalty, imgkernel = "bspline:d=3", imgboundary = "mirror";
    bool debug = false;
    CParamTable t;
    t.add_float("rate", rate, 1.0f, std::numeric_limits<float>::max());
    t.add_fvector("anisorate", aniso, dim);
    t.add_string("kernel", kernel);
    t.add_string("penalty", penalty);
    t.add_string("imgkernel", imgkernel);
    t.add_string("imgboundary", imgboundary);
    if (dim == 3)
        t.add_bool("debug", debug);
    t.set("spline", d);
    if (kernel != "bspline:d=3" && kernel != "bspline")
        throw std::invalid_argument("spline: the transformation kernel of this build is bspline:d=3");
    if (imgkernel == "bspline:d=3" || imgkernel == "bspline")
        m_img_degree = 3;
    else if (imgkernel == "bspline:d=1")
        m_img_degree = 1;
    else
        throw std::invalid_argument("spline: imgkernel must be bspline:d=3 or bspline:d=1");
    if (imgboundary != "mirror")
        throw std::invalid_argument("spline: imgboundary must be mirror");
    m_rate = Vec3{rate, rate, dim == 3 ? rate : 1.0f};
    for (int a = 0; a < dim; ++a)
        if (aniso[a] > 0.0f)
            m_rate[a] = aniso[a];
    m_penalty = produce_spline_penalty(dim, penalty, 0);
}

PSplineTransformation CSplineTransformCreator::create(Size3 size) const
{
    auto t = std::make_shared<CSplineTransformation>(m_dim, size, m_rate, m_img_degree, m_penalty);
    t->set_creator_string(m_descr);
    return t;
}

CSSDCost CSSDCost::produce(const std::string &descr)
{
    CParsedDescr d(descr);
    if (d.name != "ssd")
        throw std::invalid_argument("this build provides the 'ssd' image cost, not '" + d.name + "'");
    CSSDCost c;
    CParamTable t;
    t.add_bool("norm", c.norm);
    t.add_float("autothresh", c.autothresh, 0.0f, 1000.0f);
    t.set("ssd", d);
    return c;
}

// ---------------------------------------------------------------------------
// image full cost
// ---------------------------------------------------------------------------

CImageFullCost::CImageFullCost(int dim, const std::string &src_key, const std::string &ref_key, CSSDCost cost,
                               double weight, int grad_mode, int device)
    : m_dim(dim), m_src_key(src_key), m_ref_key(ref_key), m_cost(cost), m_weight(weight), m_grad_mode(grad_mode),
      m_device(device)
{
}

CImageFullCost::~CImageFullCost()
{
    if (m_ctx)
        mia_hip_ctx_destroy(m_ctx);
}

std::shared_ptr<CImageFullCost> CImageFullCost::produce(int dim, const std::string &descr, int device)
{
    CParsedDescr d(descr);
    if (d.name != "image")
        throw std::invalid_argument("this build provides the 'image' full cost, not '" + d.name + "'");
    std::string src = "src.@", ref = "ref.@", cost = "ssd";
    float weight = 1.0f;
    bool debug = false;
    CParamTable t;
    t.add_string("src", src);
    t.add_string("ref", ref);
    t.add_string("cost", cost);
    t.add_bool("debug", debug);
    t.add_float("weight", weight, -std::numeric_limits<float>::max(), std::numeric_limits<float>::max());
    t.set("image", d);
    return std::make_shared<CImageFullCost>(dim, src, ref, CSSDCost::produce(cost), weight, MIA_HIP_GRAD_REFERENCE,
                                            device);
}

void CImageFullCost::reinit()
{
    m_src = CDatapool::instance().get(m_src_key);
    m_ref = CDatapool::instance().get(m_ref_key);
    if (m_src->dim != m_dim || m_ref->dim != m_dim)
        throw std::invalid_argument(dim_prefix(m_dim) + "ImageFullCost: image dimension mismatch");
    if (m_src->size != m_ref->size)
        throw std::runtime_error(dim_prefix(m_dim) + "ImageFullCost only works with images of equal size");
    m_dsrc.reset();
    m_dref.reset();
    m_dsrc_level.reset();
    m_dref_level.reset();
    m_current = m_src->size;
    m_ctx_size = Size3{0, 0, 0};
}

bool CImageFullCost::get_full_size(Size3 &size) const
{
    if (!m_src)
        throw std::runtime_error("CImageFullCost: call reinit() before get_full_size()");
    if (size == Size3{0, 0, 0}) {
        size = m_src->size;
        return true;
    }
    return size == m_src->size;
}

void CImageFullCost::set_size(Size3 size)
{
    if (!m_src)
        throw std::runtime_error("CImageFullCost: call reinit() before set_size()");
    if (m_dim == 2)
        size.z = 1;
    if (size == m_current)
        return;
    m_dsrc_level.reset();
    m_dref_level.reset();
    m_current = size;
    m_ctx_size = Size3{0, 0, 0};
}

void CImageFullCost::level_images() const
{
    if (m_dsrc_level && m_dref_level)
        return;
    auto upload = [&](const CImage &img) {
        const uint32_t sz[3] = {img.size.x, img.size.y, img.size.z};
        mia_hip_image *p = nullptr;
        hip_check(mia_hip_image_create(m_device, img.dim, sz, img.data.data(), &p), "image cost: upload");
        return PDeviceImage(p, mia_hip_image_destroy);
    };
    if (!m_dsrc || !m_dref) {
        m_dsrc = upload(*m_src);
        m_dref = upload(*m_ref);
    }
    if (m_current == m_src->size) {
        m_dsrc_level = m_dsrc;
        m_dref_level = m_dref;
        return;
    }
    auto scale = [&](const PDeviceImage &full) {
        const uint32_t sz[3] = {m_current.x, m_current.y, m_dim == 3 ? m_current.z : 1u};
        mia_hip_image *p = nullptr;
        hip_check(mia_hip_image_scale(full.get(), sz, &p), "image cost: scale");
        return PDeviceImage(p, mia_hip_image_destroy);
    };
    m_dsrc_level = scale(m_dsrc);
    m_dref_level = scale(m_dref);
}

void CImageFullCost::prepare(const CSplineTransformation &t) const
{
    if (t.get_dim() != m_dim)
        throw std::invalid_argument("CImageFullCost: transformation dimension mismatch");
    if (t.get_size() != m_current)
        throw std::invalid_argument("CImageFullCost: transformation and cost sizes differ");
    if (m_ctx && m_ctx_size == m_current && m_ctx_csize == t.get_coeff_size() &&
        m_ctx_degree == t.get_image_degree())
        return;
    if (!m_ctx)
        hip_check(mia_hip_ctx_create(m_device, &m_ctx), "image cost: context");
    const uint32_t sz[3] = {m_current.x, m_current.y, m_current.z};
    const Size3 cs = t.get_coeff_size();
    const uint32_t c3[3] = {cs.x, cs.y, cs.z};
    if (m_ctx_size == m_current && m_ctx_degree == t.get_image_degree()) {
        // a refined grid at the same level: the images on the device stay
        hip_check(mia_hip_setup_grid(m_ctx, c3), "image cost: grid");
    } else {
        level_images();
        hip_check(mia_hip_setup(m_ctx, m_dim, 1, sz, c3, t.get_image_degree()), "image cost: setup");
        hip_check(mia_hip_set_images_image(m_ctx, 0, m_dsrc_level.get(), m_dref_level.get()), "image cost: images");
    }
    hip_check(mia_hip_set_cost(m_ctx, m_cost.norm ? 1 : 0, m_cost.autothresh, m_weight, m_grad_mode),
              "image cost: cost");
    m_ctx_size = m_current;
    m_ctx_csize = cs;
    m_ctx_degree = t.get_image_degree();
}

mia_hip_ctx *CImageFullCost::device_context(const CSplineTransformation &t) const
{
    prepare(t);
    return m_ctx;
}

// the weight is applied on the device (the cost and the gradient's scale)
double CImageFullCost::evaluate(const CSplineTransformation &t, double *gradient) const
{
    prepare(t);
    const std::vector<double> p = t.get_parameters();
    double cost = 0.0;
    hip_check(mia_hip_cost_grad(m_ctx, p.data(), gradient, &cost), "image cost: evaluate");
    return cost;
}

double CImageFullCost::evaluate(const CSplineTransformation &t, std::vector<double> &gradient) const
{
    gradient.resize(t.degrees_of_freedom());
    return evaluate(t, gradient.data());
}

double CImageFullCost::cost_value(const CSplineTransformation &t) const
{
    m_grad_buf.resize(t.degrees_of_freedom());
    return evaluate(t, m_grad_buf.data());
}

// ---------------------------------------------------------------------------
// cost list, problem, driver
// ---------------------------------------------------------------------------

void CFullCostList::reinit()
{
    for (auto &c : m_costs)
        c->reinit();
}

bool CFullCostList::get_full_size(Size3 &size) const
{
    bool r = true;
    for (auto it = m_costs.begin(); it != m_costs.end() && r; ++it)
        r = (*it)->get_full_size(size);
    return r;
}

void CFullCostList::set_size(Size3 size)
{
    for (auto &c : m_costs)
        c->set_size(size);
}

double CFullCostList::evaluate(const CSplineTransformation &t, std::vector<double> &gradient) const
{
    double result = 0.0;
    std::vector<double> tmp(gradient.size());
    for (const auto &c : m_costs) {
        std::fill(tmp.begin(), tmp.end(), 0.0);
        result += c->evaluate(t, tmp.data());
        for (size_t k = 0; k < gradient.size(); ++k)
            gradient[k] += tmp[k];
    }
    return result;
}

double CFullCostList::cost_value(const CSplineTransformation &t) const
{
    double result = 0.0;
    for (const auto &c : m_costs)
        result += c->cost_value(t);
    return result;
}

double CNonrigRegGradientProblem::fdf(size_t n, const double *x, double *g)
{
    m_t.set_parameters(x, n);
    std::vector<double> gv(n, 0.0);
    double result = m_costs.evaluate(m_t, gv);
    if (m_t.has_energy_penalty()) {
        std::vector<double> help;
        result += m_t.get_energy_penalty_and_gradient(help);
        for (size_t k = 0; k < n; ++k)
            gv[k] += help[k];
    }
    std::copy(gv.begin(), gv.end(), g);
    ++m_grad_evals;
    ++m_func_evals;
    return result;
}

double CNonrigRegGradientProblem::f(size_t n, const double *x)
{
    m_t.set_parameters(x, n);
    double result = m_costs.cost_value(m_t);
    if (m_t.has_energy_penalty())
        result += m_t.get_energy_penalty();
    ++m_func_evals;
    return result;
}

bool normalize_images(CImage &src, CImage &ref)
{
    double sum = 0.0, sum2 = 0.0;
    const size_t n1 = src.data.size();
    const double n = 2.0 * (double)n1;
    for (size_t i = 0; i < n1; ++i) {
        const double a = src.data[i], b = ref.data[i];
        sum += a + b;
        sum2 += a * a + b * b;
    }
    const double mean = sum / n;
    const double sigma = std::sqrt((sum2 - sum * sum / n) / (n - 1.0));
    if (sigma == 0.0)
        return false;
    // "convert:repn=float,map=linear,b=-mean/sigma,a=1/sigma"; elementwise,
    // so host threads over contiguous chunks give the same values as one loop
    const double a = 1.0 / sigma, b = -mean / sigma;
    auto map = [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            src.data[i] = (float)(a * src.data[i] + b);
            ref.data[i] = (float)(a * ref.data[i] + b);
        }
    };
    const size_t chunk = 1 << 20;
    const size_t nt = std::min<size_t>({(n1 + chunk - 1) / chunk, 16, std::max(1u, std::thread::hardware_concurrency())});
    if (nt <= 1) {
        map(0, n1);
        return true;
    }
    std::vector<std::thread> pool;
    const size_t per = (n1 + nt - 1) / nt;
    for (size_t t = 1; t < nt; ++t)
        pool.emplace_back(map, std::min(n1, t * per), std::min(n1, (t + 1) * per));
    map(0, std::min(n1, per));
    for (auto &th : pool)
        th.join();
    return true;
}

CNonrigidRegister::CNonrigidRegister(CFullCostList &costs, const std::string &minimizer,
                                     const CSplineTransformCreator &creator, int mg_levels, int idx)
    : m_costs(costs), m_minimizer(minimizer), m_creator(creator), m_levels(mg_levels), m_idx(idx)
{
    if (mg_levels < 1)
        throw std::invalid_argument("CNonrigidRegister: at least one multiresolution level is needed");
}

double CNonrigidRegister::minimize(CNonrigRegGradientProblem &p, std::vector<double> &x, LevelStats &st) const
{
    std::string err;
    std::unique_ptr<ReverseMinimizer> opt(produce_minimizer(m_minimizer, x.size(), err));
    if (!opt)
        throw std::invalid_argument(err);
    std::vector<double> g(x.size());
    opt->start(x.data());
    while (!opt->done()) {
        const double f = p.fdf(x.size(), opt->point(), g.data());
        opt->tell(f, g.data());
    }
    if (opt->evaluations() > 0 && std::isfinite(opt->best_f()))
        x.assign(opt->best_x(), opt->best_x() + x.size());
    st.evaluations += opt->evaluations();
    return opt->best_f();
}

bool CNonrigidRegister::minimize_on_device(CSplineTransformation &t, std::vector<double> &x, LevelStats &st,
                                           double &best) const
{
    int maxeval = 0, memory = 0;
    double step = 0.0;
    std::string err;
    if (!device_lbfgs_params(m_minimizer, maxeval, memory, step, err)) {
        if (!err.empty())
            throw std::invalid_argument(err);
        return false;
    }
    mia_hip_divcurl *pen = nullptr;
    double pen_w = 0.0;
    if (m_costs.size() != 1 || !t.device_penalty(&pen, &pen_w))
        return false;
    mia_hip_ctx *ctx = m_costs.at(0)->device_context(t);
    // the penalty is part of this minimisation's evaluations only
    struct Detach {
        mia_hip_ctx *c;
        ~Detach() { (void)mia_hip_set_penalty(c, 0, nullptr, 0.0); }
    } detach{ctx};
    hip_check(mia_hip_set_penalty(ctx, 0, pen, pen_w), "nonrigid register: penalty");
    std::vector<double> xo(x.size());
    double cost = 0.0, loop = 0.0;
    int evals = 0;
    hip_check(mia_hip_optimize(ctx, x.data(), maxeval, memory, step, 10, xo.data(), &cost, &evals, &loop),
              "nonrigid register: device L-BFGS");
    st.loop_seconds += loop;
    if (evals > 0 && std::isfinite(cost))
        x = xo;
    st.evaluations += evals;
    best = cost;
    return true;
}

PSplineTransformation CNonrigidRegister::run(const CImage &src_in, const CImage &ref_in)
{
    return run(std::make_shared<CImage>(src_in), std::make_shared<CImage>(ref_in));
}

PSplineTransformation CNonrigidRegister::run(PImage src, PImage ref)
{
    if (!src || !ref || src->size != ref->size || src->dim != ref->dim)
        throw std::invalid_argument("CNonrigidRegister: src and ref differ in size");
    m_stats.clear();
    if (!normalize_images(*src, *ref))
        return m_creator.create(src->size);
    std::string src_name = "src.@", ref_name = "ref.@";
    if (m_idx >= 0) {
        src_name = "src" + std::to_string(m_idx) + ".@";
        ref_name = "ref" + std::to_string(m_idx) + ".@";
    }
    CDatapool::instance().add(src_name, src);
    CDatapool::instance().add(ref_name, ref);
    m_costs.reinit();
    Size3 global{0, 0, 0};
    if (!m_costs.get_full_size(global))
        throw std::invalid_argument("Nonrigidregister: the given combination of cost functions doesn't agree on "
                                    "the size of the registration problem");
    PSplineTransformation t;
    int shift = m_levels;
    do {
        --shift;
        const unsigned f = 1u << shift;
        Size3 local{global.x / f, global.y / f, src->dim == 3 ? global.z / f : 1u};
        const auto t0 = std::chrono::steady_clock::now();
        t = t ? t->upscale(local) : m_creator.create(local);
        m_costs.set_size(local);
        LevelStats st;
        st.size = local;
        CNonrigRegGradientProblem gp(m_costs, *t);
        std::vector<double> x = t->get_parameters();
        auto m0 = std::chrono::steady_clock::now();
        if (!minimize_on_device(*t, x, st, st.cost))
            st.cost = minimize(gp, x, st);
        st.minimize_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - m0).count();
        t->set_parameters(x);
        if (t->refine()) {
            gp.reset_counters();
            x = t->get_parameters();
            m0 = std::chrono::steady_clock::now();
            if (!minimize_on_device(*t, x, st, st.cost))
                st.cost = minimize(gp, x, st);
            st.minimize_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - m0).count();
            t->set_parameters(x);
        }
        st.coeff_size = t->get_coeff_size();
        st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        m_stats.push_back(st);
    } while (shift);
    return t;
}

} // namespace mia_amd

// C++ host mirror of the reference's classes on the spline + ssd nonrigid
// registration path, running on libmia_hip.so:
//
//   C2DSplineTransformation / C3DSplineTransformation
//       mia/2d/transform/spline.{hh,cc}, mia/3d/transform/spline.{hh,cc}
//   CSSDCost (the `ssd` image cost parameters)     mia/template/ssd.hh:201-240
//   CImageFullCost (the `image` full cost)         mia/3d/fullcost/image.cc, mia/2d/fullcost/image.cc,
//                                                  mia/template/fullcost.cxx
//   CFullCostList                                  mia/template/multicost.cxx
//   CNonrigRegGradientProblem (Problem::fdf)       mia/template/nonrigidregister.cxx:380-440
//   CNonrigidRegister (multiresolution driver)     mia/template/nonrigidregister.cxx:192-296
//   CSplinePenalty / CDivcurlSplinePenalty         mia/3d/splinetransformpenalty.cc, mia/3d/splinepenalty/divcurl.cc
//
// Error behaviour follows the reference: std::invalid_argument for bad sizes
// and parameters, std::runtime_error for inconsistent images; a failing GPU
// call throws std::runtime_error with the library's message.  Names follow the
// reference; one runtime class serves 2D and 3D (dim 2 or 3), vectors carry
// three components with z unused in 2D.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mia_hip.h"

namespace mia_amd {

struct Size3 {
    unsigned x = 1, y = 1, z = 1;
    size_t product() const { return (size_t)x * y * z; }
    bool operator==(const Size3 &o) const { return x == o.x && y == o.y && z == o.z; }
    bool operator!=(const Size3 &o) const { return !(*this == o); }
    unsigned operator[](int i) const { return i == 0 ? x : i == 1 ? y : z; }
};

using Vec3 = std::array<float, 3>;

// Throw std::runtime_error (or std::invalid_argument for argument errors)
// carrying the library's message when a libmia_hip call fails.
void hip_check(int status, const char *what);

// A float image of dimension 2 (size.z == 1) or 3, x fastest.
struct CImage {
    int dim = 3;
    Size3 size;
    std::vector<float> data;
    CImage() = default;
    CImage(int d, Size3 s) : dim(d), size(s), data(s.product(), 0.0f) {}
    // a copy of s.product() values at `values` (one pass over fresh memory)
    CImage(int d, Size3 s, const float *values) : dim(d), size(s), data(values, values + s.product()) {}
};
using PImage = std::shared_ptr<CImage>;

// The data pool the `image` full cost reads its images from ("src.@",
// "ref.@" keys; mia/core/datapool.hh).
class CDatapool {
public:
    static CDatapool &instance();
    void add(const std::string &key, PImage image) { m_map[key] = std::move(image); }
    PImage get(const std::string &key) const;
    bool has_key(const std::string &key) const { return m_map.count(key) != 0; }
    void clear() { m_map.clear(); }

private:
    std::map<std::string, PImage> m_map;
};

// Image resampling to a new size with MIA's `scale` filter semantics
// (mia/3d/filter/scale.cc:38-104, mia/core/scaler1d.cc), on the GPU.
CImage scale_image(const CImage &in, Size3 out_size, int device = 0);

// Spline transformation penalty (mia/3d/splinetransformpenalty.cc:40-120).
class CSplinePenalty {
public:
    CSplinePenalty(double weight, bool normalize) : m_weight(weight), m_normalize(normalize) {}
    virtual ~CSplinePenalty() = default;
    void initialize(int dim, Size3 coeff_size, Vec3 range);
    // weight (divided by the image volume with norm=1) times the energy
    double value(const std::vector<float> &coefficients) const;
    // as value(); writes gradient = -weight * d(energy), the reference's sign
    // (splinetransformpenalty.cc:60-71, test_divcurl.cc:51-58)
    double value_and_gradient(const std::vector<float> &coefficients, std::vector<double> &gradient) const;
    virtual CSplinePenalty *clone() const = 0;
    double get_weight() const { return m_weight; }
    bool get_normalize() const { return m_normalize; }
    // the weight value() applies (divided by the image volume with norm=1)
    double effective_weight() const;
    // the penalty's device object once initialized, for evaluation inside a
    // context (mia_hip_set_penalty); nullptr where there is none
    virtual mia_hip_divcurl *device_handle() const { return nullptr; }

protected:
    virtual void do_initialize() = 0;
    virtual double do_value(const std::vector<float> &coefficients) const = 0;
    virtual double do_value_and_gradient(const std::vector<float> &coefficients,
                                         std::vector<double> &gradient) const = 0;
    int m_dim = 3;
    Size3 m_size;
    Vec3 m_range{};

private:
    double m_weight;
    bool m_normalize;
};
using PSplinePenalty = std::shared_ptr<CSplinePenalty>;

// The `divcurl` penalty, 3D (mia/3d/splinepenalty/divcurl.cc, mia/3d/ppmatrix.cc),
// evaluated on the GPU.
class CDivcurlSplinePenalty : public CSplinePenalty {
public:
    CDivcurlSplinePenalty(double weight, bool normalize, double div, double curl, int device = 0);
    ~CDivcurlSplinePenalty() override;
    CSplinePenalty *clone() const override;
    mia_hip_divcurl *device_handle() const override;

private:
    void do_initialize() override;
    double do_value(const std::vector<float> &coefficients) const override;
    double do_value_and_gradient(const std::vector<float> &coefficients,
                                 std::vector<double> &gradient) const override;
    double m_div, m_curl;
    int m_device;
    struct Impl;
    std::unique_ptr<Impl> m_impl;
};

// B-spline (cubic kernel) free-form transformation T(x) = x - u(x).
class CSplineTransformation {
public:
    // range: image size; rate: coefficient rate per axis (>= 1); img_degree:
    // image interpolation kernel bspline:d=img_degree (1 or 3).  Throws
    // std::invalid_argument when the rate is too large for the image
    // (mia/3d/transform/spline.cc:130-143).
    CSplineTransformation(int dim, Size3 range, Vec3 rate, int img_degree = 3, PSplinePenalty penalty = {});
    CSplineTransformation(const CSplineTransformation &org);
    virtual ~CSplineTransformation() = default;

    int get_dim() const { return m_dim; }
    const Size3 &get_size() const { return m_range; }
    const Size3 &get_coeff_size() const { return m_csize; }
    int get_image_degree() const { return m_img_degree; }
    const Vec3 &get_target_rate() const { return m_rate; }
    size_t degrees_of_freedom() const { return m_coeffs.size(); }

    std::vector<double> get_parameters() const;
    // throws std::invalid_argument on a size mismatch
    void set_parameters(const std::vector<double> &params);
    void set_parameters(const double *params, size_t n);
    void set_identity();
    // coefficient field (dim components per coefficient, x fastest); throws
    // std::invalid_argument if a dimension is <= 3 (spline.cc:158-177)
    void set_coefficients(const std::vector<float> &field, Size3 field_size);
    void set_coefficients_and_prefilter(const std::vector<float> &field, Size3 field_size);
    const std::vector<float> &get_coefficients() const { return m_coeffs; }

    // scaled grid position x * scale + 1 (spline.cc:321-326)
    Vec3 scale(const Vec3 &x) const;
    // u(x) (spline.cc:299-304), T(x) = x - u(x) (:314-318)
    Vec3 get_displacement_at(const Vec3 &x) const;
    Vec3 operator()(const Vec3 &x) const;
    // the warped image M(T(x)) (mia/3d/transform.cc:280-353), on the GPU
    CImage operator()(const CImage &image, int device = 0) const;

    // the same displacement on an image grid of `size` (spline.cc:399-412)
    std::shared_ptr<CSplineTransformation> upscale(Size3 size) const;
    // grow the coefficient grid to the target rate (spline.cc:360-397)
    bool refine();

    bool has_energy_penalty() const { return (bool)m_penalty; }
    // The penalty as a device object for the current coefficient grid and
    // the weight it is applied with: true with *p = nullptr without a
    // penalty, false if the penalty has no device form.
    bool device_penalty(mia_hip_divcurl **p, double *weight) const;
    double get_energy_penalty() const;
    double get_energy_penalty_and_gradient(std::vector<double> &gradient) const;

    // "spline:rate=...", what the reference's create() stores
    const std::string &get_creator_string() const { return m_creator; }
    void set_creator_string(const std::string &s) { m_creator = s; }

    // Attributes in their string form (CAttributedData,
    // mia/core/attributes.hh:259-349: set_attribute(key, string),
    // get_attribute_as_string, the attribute map); the transformation file
    // stores them as strings (mia/3d/transio/serialization.hh:43-49).
    void set_attribute(const std::string &key, const std::string &value) { m_attributes[key] = value; }
    bool has_attribute(const std::string &key) const { return m_attributes.count(key) != 0; }
    // "" for a key that is not set (get_attribute_as_string of a missing key)
    std::string get_attribute_as_string(const std::string &key) const
    {
        const auto i = m_attributes.find(key);
        return i == m_attributes.end() ? std::string() : i->second;
    }
    const std::map<std::string, std::string> &attributes() const { return m_attributes; }

private:
    void reinit_penalty() const;
    int m_dim;
    Size3 m_range;
    Vec3 m_rate;
    int m_img_degree;
    Size3 m_csize;
    std::vector<float> m_coeffs;
    PSplinePenalty m_penalty;
    mutable bool m_penalty_valid = false;
    std::string m_creator;
    std::map<std::string, std::string> m_attributes;
};
using PSplineTransformation = std::shared_ptr<CSplineTransformation>;

class C2DSplineTransformation : public CSplineTransformation {
public:
    C2DSplineTransformation(Size3 range, Vec3 rate, int img_degree = 3, PSplinePenalty penalty = {})
        : CSplineTransformation(2, range, rate, img_degree, std::move(penalty))
    {
    }
};

class C3DSplineTransformation : public CSplineTransformation {
public:
    C3DSplineTransformation(Size3 range, Vec3 rate, int img_degree = 3, PSplinePenalty penalty = {})
        : CSplineTransformation(3, range, rate, img_degree, std::move(penalty))
    {
    }
};

// The `spline` transform creator: "spline:rate=10,anisorate=[0,0,0],
// kernel=[bspline:d=3],penalty=[...],imgkernel=[bspline:d=3],imgboundary=mirror"
// (spline.cc:995-1027, mia/template/transformfactory.cxx:58-59).
class CSplineTransformCreator {
public:
    CSplineTransformCreator(int dim, const std::string &descr);
    PSplineTransformation create(Size3 size) const;
    const std::string &get_init_string() const { return m_descr; }

private:
    int m_dim;
    std::string m_descr;
    Vec3 m_rate;
    int m_img_degree;
    PSplinePenalty m_penalty;
};

// `ssd` cost parameters (mia/template/ssd.hh:217-231).
struct CSSDCost {
    bool norm = false;
    float autothresh = 0.0f;
    static CSSDCost produce(const std::string &descr);
};

// The `image` full cost with the ssd kernel:
// "image:src=src.@,ref=ref.@,cost=[ssd:norm=0,autothresh=0],weight=1".
class CImageFullCost {
public:
    CImageFullCost(int dim, const std::string &src_key, const std::string &ref_key, CSSDCost cost, double weight,
                   int grad_mode = MIA_HIP_GRAD_REFERENCE, int device = 0);
    ~CImageFullCost();
    static std::shared_ptr<CImageFullCost> produce(int dim, const std::string &descr, int device = 0);

    // read src / ref from the data pool (image.cc:111-127)
    void reinit();
    // true and the full image size if size is 0, else whether it matches
    bool get_full_size(Size3 &size) const;
    // resample the images to `size` when it differs (image.cc:80-97)
    void set_size(Size3 size);
    const Size3 &get_current_size() const { return m_current; }
    // weight * cost; gradient (overwritten) = weight * d cost / d params
    // (fullcost.cxx:36-45, image.cc:67-78)
    double evaluate(const CSplineTransformation &t, std::vector<double> &gradient) const;
    double evaluate(const CSplineTransformation &t, double *gradient) const;
    double cost_value(const CSplineTransformation &t) const;
    double get_weight() const { return m_weight; }
    int get_dim() const { return m_dim; }
    // the device context that evaluates this cost for t's size and grid
    // (images at the current size, weight applied on the device)
    mia_hip_ctx *device_context(const CSplineTransformation &t) const;

private:
    void prepare(const CSplineTransformation &t) const;
    int m_dim;
    std::string m_src_key, m_ref_key;
    CSSDCost m_cost;
    double m_weight;
    int m_grad_mode;
    int m_device;
    PImage m_src, m_ref;
    // the images on the device: full size (uploaded once per reinit) and at
    // the current size (resampled there from the full-size ones, or shared
    // with them), made when a context first needs them
    using PDeviceImage = std::shared_ptr<mia_hip_image>;
    void level_images() const;
    mutable PDeviceImage m_dsrc, m_dref, m_dsrc_level, m_dref_level;
    Size3 m_current;
    mutable mia_hip_ctx *m_ctx = nullptr;
    mutable Size3 m_ctx_size, m_ctx_csize;
    mutable int m_ctx_degree = -1;
    mutable std::vector<double> m_grad_buf;
};
using PImageFullCost = std::shared_ptr<CImageFullCost>;

// Sum of full costs (multicost.cxx:58-110): each cost writes into a zeroed
// temporary that is added to the caller's gradient.
class CFullCostList {
public:
    void push(PImageFullCost c) { m_costs.push_back(std::move(c)); }
    bool empty() const { return m_costs.empty(); }
    size_t size() const { return m_costs.size(); }
    const PImageFullCost &at(size_t i) const { return m_costs.at(i); }
    void reinit();
    bool get_full_size(Size3 &size) const;
    void set_size(Size3 size);
    // adds each cost's gradient into `gradient`
    double evaluate(const CSplineTransformation &t, std::vector<double> &gradient) const;
    double cost_value(const CSplineTransformation &t) const;

private:
    std::vector<PImageFullCost> m_costs;
};

// The minimizer's problem (nonrigidregister.cxx:380-440): fdf sets the
// parameters, zero-fills g, adds the costs' gradients and the penalty.
class CNonrigRegGradientProblem {
public:
    CNonrigRegGradientProblem(const CFullCostList &costs, CSplineTransformation &t) : m_costs(costs), m_t(t) {}
    size_t size() const { return m_t.degrees_of_freedom(); }
    double fdf(size_t n, const double *x, double *g);
    double f(size_t n, const double *x);
    int get_func_evals() const { return m_func_evals; }
    int get_grad_evals() const { return m_grad_evals; }
    void reset_counters() { m_func_evals = m_grad_evals = 0; }

private:
    const CFullCostList &m_costs;
    CSplineTransformation &m_t;
    std::vector<double> m_g;
    int m_func_evals = 0, m_grad_evals = 0;
};

// Multiresolution nonrigid registration (nonrigidregister.cxx:192-296):
// images mapped jointly to mean 0 / sigma 1, then per level (1/2^k of the
// full size) create or upscale the transformation, resample the images,
// minimise, refine the coefficient grid and minimise again if it grew.
class CNonrigidRegister {
public:
    struct LevelStats {
        Size3 size, coeff_size;
        int evaluations = 0;
        double cost = 0.0, seconds = 0.0;
        // of `seconds`: in the minimisations (with the device optimizer's
        // set-up and graph capture), and in its iteration loop alone
        double minimize_seconds = 0.0, loop_seconds = 0.0;
    };
    CNonrigidRegister(CFullCostList &costs, const std::string &minimizer, const CSplineTransformCreator &creator,
                      int mg_levels, int idx = -1);
    PSplineTransformation run(const CImage &src, const CImage &ref);
    // the same on images the registration may keep and normalise in place
    PSplineTransformation run(PImage src, PImage ref);
    const std::vector<LevelStats> &get_level_stats() const { return m_stats; }

private:
    double minimize(CNonrigRegGradientProblem &p, std::vector<double> &x, LevelStats &st) const;
    // the same minimisation with the on-device L-BFGS (one image cost, an
    // optional divcurl penalty, lbfgs without tolerances); false if the
    // problem does not qualify
    bool minimize_on_device(CSplineTransformation &t, std::vector<double> &x, LevelStats &st, double &best) const;
    CFullCostList &m_costs;
    std::string m_minimizer;
    const CSplineTransformCreator &m_creator;
    int m_levels;
    int m_idx;
    std::vector<LevelStats> m_stats;
};

// Transformation files in the layout of MIA's "xml" transformation IO
// (mia/3d/transio/xml.cc, mia/3d/transio/serialization.hh:36-79; 2D
// without size_z): creator string, size, attributes (the string map,
// restored on load) and the parameter vector.  load_transformation re-creates
// the transformation from its creator string and size and sets the parameters;
// it throws std::runtime_error on a malformed file or a parameter count that
// does not match the creator (as the reference's loaders do).
void save_transformation(const std::string &path, const CSplineTransformation &t);
PSplineTransformation load_transformation(const std::string &path, int dim);

// Joint mean-0 / sigma-1 map of two images (nonrigidregister.cxx:147-184);
// false if both are of one value (nothing to register).
bool normalize_images(CImage &src, CImage &ref);

} // namespace mia_amd

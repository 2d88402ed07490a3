// Host-side tests of the C++ mirror of the reference's classes (no GPU
// calls): descriptor and parameter handling, the plugin registry's priority
// rule, the spline transformation's parameter layout, size checks,
// single-point displacement (against the FP64 oracle), upscale and refine.
// Cases follow the reference's own tests where they exist
// (mia/2d/transform/test_spline.cc, mia/3d/transform/test_spline.cc,
// mia/core/test_handler.cc).
#include <cmath>
#include <cstdlib>
#include <random>
#include <stdexcept>

#include "../../mia_amd/host/descr.hh"
#include "../../mia_amd/host/nonrigid.hh"
#include "../../mia_amd/host/plugin.hh"
#include "check.hh"

extern "C" void mo_displacement_at(int dim, const int size[3], const int csize[3], const double *params,
                                   const double *pos, double *out);

using namespace mia_amd;

static std::string g_module; // path of the test plugin module (argv / env)

TEST_CASE(descriptor_nested_and_errors)
{
    CParsedDescr d("spline:rate=8,penalty=[divcurl:weight=0.01,div=2],imgkernel=[bspline:d=1]");
    CHECK(d.name == "spline");
    CHECK(d.params.size() == 3);
    CHECK(d.params[1].first == "penalty" && d.params[1].second == "divcurl:weight=0.01,div=2");
    CHECK(d.params[2].second == "bspline:d=1");
    CHECK_THROW(CParsedDescr(":rate=8"), std::invalid_argument);
    CHECK_THROW(CParsedDescr("spline:rate"), std::invalid_argument);
    CHECK_THROW(CParsedDescr("spline:penalty=[divcurl:weight=1"), std::invalid_argument);
}

TEST_CASE(spline_creator_parameters)
{
    CSplineTransformCreator c(3, "spline:rate=8");
    auto t = c.create(Size3{256, 256, 256});
    // csize = floor((n - 1 + rate) / rate) + 2 (mia/3d/transform/spline.cc:130-143): 34 at 256 / rate 8
    CHECK(t->get_coeff_size() == (Size3{34, 34, 34}));
    CHECK(t->degrees_of_freedom() == 3u * 34 * 34 * 34);
    CHECK(t->get_creator_string() == "spline:rate=8");
    CSplineTransformCreator a(3, "spline:rate=8,anisorate=[4,0,16]");
    auto ta = a.create(Size3{64, 64, 64});
    CHECK(ta->get_coeff_size() == (Size3{18, 10, 6}));
    CHECK_THROW(CSplineTransformCreator(3, "spline:rate=0.5"), std::invalid_argument);
    CHECK_THROW(CSplineTransformCreator(3, "spline:ratex=8"), std::invalid_argument);
    CHECK_THROW(CSplineTransformCreator(3, "spline:kernel=[bspline:d=4]"), std::invalid_argument);
    CHECK_THROW(CSplineTransformCreator(3, "spline:imgkernel=[omoms:d=3]"), std::invalid_argument);
    CHECK_THROW(CSplineTransformCreator(3, "spline:imgboundary=zero"), std::invalid_argument);
    CHECK_THROW(CSplineTransformCreator(2, "spline:debug=1"), std::invalid_argument); // 2D has no debug
    CHECK_THROW(CSplineTransformCreator(3, "grid:rate=8"), std::invalid_argument);
    CSplineTransformCreator l(2, "spline:rate=16,imgkernel=[bspline:d=1]");
    CHECK(l.create(Size3{128, 128, 1})->get_image_degree() == 1);
    CHECK_THROW(CSplineTransformCreator(3, "spline:penalty=[divcurl:weight=0]"), std::invalid_argument);
    CHECK_THROW(CSplineTransformCreator(3, "spline:penalty=[divcurl:div=-1]"), std::invalid_argument);
    CHECK_THROW(CSplineTransformCreator(2, "spline:penalty=[divcurl]"), std::invalid_argument);
    CHECK_THROW(CSplineTransformCreator(3, "spline:penalty=[laplace]"), std::invalid_argument);
}

TEST_CASE(transformation_small_range_throws)
{
    // mia/2d/transform/test_spline.cc:835-845: 10x10 with rate (10,2) or (2,10)
    CHECK_THROW(C2DSplineTransformation(Size3{10, 10, 1}, Vec3{10.0f, 2.0f, 1.0f}), std::invalid_argument);
    CHECK_THROW(C2DSplineTransformation(Size3{10, 10, 1}, Vec3{2.0f, 10.0f, 1.0f}), std::invalid_argument);
    C2DSplineTransformation ok(Size3{10, 10, 1}, Vec3{2.0f, 2.0f, 1.0f});
    CHECK(ok.get_coeff_size().x == 7u);
}

TEST_CASE(parameters_layout_and_sizes)
{
    C3DSplineTransformation t(Size3{20, 24, 28}, Vec3{4.0f, 4.0f, 4.0f});
    const size_t n = t.degrees_of_freedom();
    std::vector<double> p(n);
    for (size_t k = 0; k < n; ++k)
        p[k] = 0.25 * (double)(k % 17) - 1.0;
    t.set_parameters(p);
    const auto q = t.get_parameters();
    CHECK(q.size() == n);
    bool same = true;
    for (size_t k = 0; k < n; ++k)
        same &= q[k] == (double)(float)p[k];
    CHECK(same);
    // the coefficient field is interleaved (x, y, z) per coefficient, x fastest (spline.cc:329-358)
    CHECK(t.get_coefficients()[3] == (float)p[3]);
    p.pop_back();
    CHECK_THROW(t.set_parameters(p), std::invalid_argument);
    std::vector<float> small(3 * 3 * 8 * 8, 0.0f);
    CHECK_THROW(t.set_coefficients(small, Size3{3, 8, 8}), std::invalid_argument);
    t.set_identity();
    bool zero = true;
    for (float v : t.get_coefficients())
        zero &= v == 0.0f;
    CHECK(zero);
}

static void check_displacement_at(int dim, Size3 size, float rate, unsigned seed)
{
    CSplineTransformation t(dim, size, Vec3{rate, rate, rate});
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.5);
    std::vector<double> p(t.degrees_of_freedom());
    for (double &v : p)
        v = nd(rng);
    t.set_parameters(p);
    const Size3 cs = t.get_coeff_size();
    const int sz[3] = {(int)size.x, (int)size.y, (int)size.z};
    const int c3[3] = {(int)cs.x, (int)cs.y, (int)cs.z};
    std::vector<double> pf(t.get_parameters());
    std::uniform_real_distribution<double> ud(0.0, 1.0);
    double worst = 0.0;
    for (int k = 0; k < 50; ++k) {
        Vec3 x{(float)(ud(rng) * (size.x - 1)), (float)(ud(rng) * (size.y - 1)),
               dim == 3 ? (float)(ud(rng) * (size.z - 1)) : 0.0f};
        const Vec3 u = t.get_displacement_at(x);
        const double pos[3] = {x[0], x[1], x[2]};
        double ref[3] = {0, 0, 0};
        mo_displacement_at(dim, sz, c3, pf.data(), pos, ref);
        for (int a = 0; a < dim; ++a)
            worst = std::max(worst, std::fabs(u[a] - ref[a]));
        const Vec3 T = t(x);
        CHECK(std::fabs(T[0] - (x[0] - u[0])) < 1e-6f);
    }
    CHECK(worst < 2e-5);
}

TEST_CASE(displacement_at_matches_oracle_2d)
{
    check_displacement_at(2, Size3{60, 80, 1}, 5.0f, 11);
}

TEST_CASE(displacement_at_matches_oracle_3d)
{
    check_displacement_at(3, Size3{30, 26, 22}, 4.0f, 12);
}

TEST_CASE(upscale_scales_displacement)
{
    for (int dim = 2; dim <= 3; ++dim) {
        const Size3 small{17, 21, dim == 3 ? 13u : 1u};
        const Size3 big{34, 42, dim == 3 ? 26u : 1u};
        CSplineTransformation t(dim, small, Vec3{4.0f, 4.0f, 4.0f});
        std::vector<double> p(t.degrees_of_freedom());
        for (size_t k = 0; k < p.size(); ++k)
            p[k] = std::sin(0.37 * (double)k);
        t.set_parameters(p);
        auto u = t.upscale(big);
        // same coefficient grid and target rate, size changed
        CHECK(u->get_coeff_size() == t.get_coeff_size());
        CHECK(u->get_size() == big);
        CHECK(u->get_target_rate() == t.get_target_rate());
        // 3D scales the displacement by size / range (spline.cc:399-412), 2D by
        // (size - 1) / (range - 1) (mia/2d/transform/spline.cc:356-370); the
        // displacement at corresponding grid positions scales accordingly
        double worst = 0.0;
        for (int k = 0; k < 20; ++k) {
            const Vec3 xs{(float)(k % 7) * 2.0f + 0.5f, (float)(k % 5) * 3.0f + 0.25f,
                          dim == 3 ? (float)(k % 4) * 2.5f : 0.0f};
            Vec3 xb{};
            for (int a = 0; a < dim; ++a)
                xb[a] = xs[a] * (float)(big[a] - 1) / (float)(small[a] - 1);
            const Vec3 us = t.get_displacement_at(xs), ub = u->get_displacement_at(xb);
            for (int a = 0; a < dim; ++a) {
                const float m = dim == 3 ? (float)big[a] / small[a] : (float)(big[a] - 1) / (small[a] - 1);
                worst = std::max(worst, (double)std::fabs(ub[a] - m * us[a]));
            }
        }
        CHECK(worst < 1e-4);
    }
}

TEST_CASE(refine_constant_field_and_growth)
{
    // a transformation upscaled from a coarse level keeps its grid until
    // refine() grows it to the target rate (spline.cc:360-397)
    C3DSplineTransformation t(Size3{24, 24, 20}, Vec3{8.0f, 8.0f, 8.0f});
    const Size3 cs0 = t.get_coeff_size();
    CHECK(!t.refine()); // already at the rate
    std::vector<float> c(t.get_coefficients().size());
    for (size_t k = 0; k < c.size(); ++k)
        c[k] = k % 3 == 0 ? 1.5f : (k % 3 == 1 ? -0.75f : 0.25f);
    t.set_coefficients(c, cs0);
    auto u = t.upscale(Size3{48, 48, 40});
    CHECK(u->get_coeff_size() == cs0);
    CHECK(u->refine());
    const Size3 cs1 = u->get_coeff_size();
    CSplineTransformation direct(3, Size3{48, 48, 40}, Vec3{8.0f, 8.0f, 8.0f});
    CHECK(cs1 == direct.get_coeff_size());
    // a constant field stays constant (prefilter and interpolation reproduce constants)
    double worst = 0.0;
    const auto &r = u->get_coefficients();
    for (size_t k = 0; k < r.size(); ++k) {
        const float e = (k % 3 == 0 ? 1.5f : (k % 3 == 1 ? -0.75f : 0.25f)) * 2.0f;
        worst = std::max(worst, (double)std::fabs(r[k] - e));
    }
    CHECK(worst < 1e-5);
    CHECK(!u->refine());
}

TEST_CASE(refine_preserves_smooth_displacement)
{
    for (int dim = 2; dim <= 3; ++dim) {
        const Size3 size{64, 56, dim == 3 ? 48u : 1u};
        CSplineTransformation coarse(dim, size, Vec3{16.0f, 16.0f, 16.0f});
        const Size3 cs = coarse.get_coeff_size();
        std::vector<float> c(coarse.get_coefficients().size());
        for (unsigned z = 0; z < cs.z; ++z)
            for (unsigned y = 0; y < cs.y; ++y)
                for (unsigned x = 0; x < cs.x; ++x)
                    for (int a = 0; a < dim; ++a)
                        c[((((size_t)z * cs.y + y) * cs.x) + x) * dim + a] =
                            (float)(2.0 * std::sin(0.5 * x + 0.3 * a) * std::cos(0.4 * y) * std::cos(0.3 * z));
        coarse.set_coefficients(c, cs);
        CSplineTransformation fine(coarse);
        // target rate 8 on the same image: refine interpolates the coarse grid
        CSplineTransformation target(dim, size, Vec3{8.0f, 8.0f, 8.0f});
        auto up = fine.upscale(size);
        (void)up;
        CSplineTransformation t8(dim, size, Vec3{8.0f, 8.0f, 8.0f});
        t8.set_coefficients(c, cs); // coarse grid, fine target rate
        CHECK(t8.refine());
        CHECK(t8.get_coeff_size() == target.get_coeff_size());
        double worst = 0.0, scale = 0.0;
        for (int k = 0; k < 40; ++k) {
            const Vec3 x{(float)(k * 13 % 60) + 1.5f, (float)(k * 7 % 50) + 2.25f,
                         dim == 3 ? (float)(k * 5 % 44) + 1.0f : 0.0f};
            const Vec3 a = coarse.get_displacement_at(x), b = t8.get_displacement_at(x);
            for (int i = 0; i < dim; ++i) {
                worst = std::max(worst, (double)std::fabs(a[i] - b[i]));
                scale = std::max(scale, (double)std::fabs(a[i]));
            }
        }
        CHECK(worst < 0.1 * scale);
    }
}

TEST_CASE(ssd_and_fullcost_parameters)
{
    const CSSDCost c = CSSDCost::produce("ssd:norm=1,autothresh=0.5");
    CHECK(c.norm && c.autothresh == 0.5f);
    CHECK_THROW(CSSDCost::produce("ssd:autothresh=2000"), std::invalid_argument);
    CHECK_THROW(CSSDCost::produce("ssd:norm=maybe"), std::invalid_argument);
    CHECK_THROW(CSSDCost::produce("ncc"), std::invalid_argument);
    auto f = CImageFullCost::produce(3, "image:src=a.@,ref=b.@,cost=[ssd:norm=1],weight=0.5");
    CHECK(f->get_weight() == 0.5);
    CDatapool::instance().clear();
    CHECK_THROW(f->reinit(), std::invalid_argument); // nothing in the pool
    CDatapool::instance().add("a.@", std::make_shared<CImage>(3, Size3{8, 8, 8}));
    CDatapool::instance().add("b.@", std::make_shared<CImage>(3, Size3{8, 8, 9}));
    CHECK_THROW(f->reinit(), std::runtime_error); // sizes differ (image.cc:115-116)
    CDatapool::instance().add("b.@", std::make_shared<CImage>(3, Size3{8, 8, 8}));
    f->reinit();
    Size3 s{0, 0, 0};
    CHECK(f->get_full_size(s) && s == (Size3{8, 8, 8}));
    Size3 o{8, 8, 7};
    CHECK(!f->get_full_size(o));
    CHECK_THROW(CImageFullCost::produce(3, "image:cost=[ssd],wieght=1"), std::invalid_argument);
    CDatapool::instance().clear();
}

TEST_CASE(normalize_images_moments)
{
    CImage a(2, Size3{30, 20, 1}), b(2, Size3{30, 20, 1});
    for (size_t k = 0; k < a.data.size(); ++k) {
        a.data[k] = (float)(k % 37) * 3.0f;
        b.data[k] = 50.0f + (float)(k % 11);
    }
    CHECK(normalize_images(a, b));
    double s = 0, s2 = 0;
    for (size_t k = 0; k < a.data.size(); ++k) {
        s += a.data[k] + b.data[k];
        s2 += (double)a.data[k] * a.data[k] + (double)b.data[k] * b.data[k];
    }
    const double n = 2.0 * a.data.size();
    CHECK(std::fabs(s / n) < 1e-5);
    CHECK(std::fabs(std::sqrt((s2 - s * s / n) / (n - 1)) - 1.0) < 1e-5);
    CImage c(2, Size3{4, 4, 1}), d(2, Size3{4, 4, 1});
    CHECK(!normalize_images(c, d)); // one value: nothing to register
}

TEST_CASE(plugin_registry_priority)
{
    auto &h = CPluginHandler::instance();
    CHECK(h.get_names("3dtransform") == std::vector<std::string>{"spline"});
    CHECK(h.get("3dfullcost", "image") != nullptr);
    CHECK(h.get("3dsplinepenalty", "divcurl")->get_priority() == 1);
    CHECK_THROW(h.get("3dtransform", "affine"), std::invalid_argument);
    auto tc = produce_transform_creator(2, "spline:rate=4");
    CHECK(tc->create(Size3{32, 32, 1})->get_coeff_size().x == 10u);
    CHECK_THROW(produce_full_cost(3, "lsd:weight=1"), std::invalid_argument);
    if (g_module.empty()) {
        std::printf("  (no test module given: priority override not checked)\n");
        return;
    }
    // the module brings 3dimage-cost/ssd at priority 5 (replaces this
    // library's, priority 1) and 3dfullcost/image at priority 0 (ignored)
    h.load_module(g_module);
    CHECK(h.get_module("3dimage-cost", "ssd") == g_module);
    CHECK(h.get("3dimage-cost", "ssd")->get_priority() == 5);
    CHECK(h.get_module("3dfullcost", "image") == "libmia_host");
    CHECK(h.get("3dimage-cost", "ssd")->get_descr() == "test module ssd");
    CHECK_THROW(h.load_module("/nonexistent/module.so"), std::runtime_error);
}

int main_impl(int argc, char **argv);

#define main main_impl
CHECK_MAIN()
#undef main

int main(int argc, char **argv)
{
    if (const char *m = std::getenv("MIA_TEST_PLUGIN_MODULE"))
        g_module = m;
    return main_impl(argc, argv);
}

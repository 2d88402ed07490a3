// GPU tests of the C++ mirror: the image full cost, the cost list, the
// minimizer problem with the divcurl penalty, image resampling and the warp
// against the FP64 oracle, the reference's divcurl plugin test, and
// multiresolution registrations end to end.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <random>
#include <stdexcept>

#include "../../mia_amd/host/nonrigid.hh"
#include "../../mia_amd/host/plugin.hh"
#include "check.hh"

extern "C" {
double mo_cost_grad(int dim, const int size[3], const float *moving, const float *fixed, const int csize[3],
                    const double *params, int degree, int norm, float autothresh, int grad_mode, double weight,
                    double *grad, double *warped, double *force);
void mo_scale_image(const float *img, const int in_size[3], const int out_size[3], float *out);
void mo_pp_cells(const int csize[3], const float range[3], double wdiv, double wrot, double *cells);
double mo_divcurl(const int cs[3], const double *cells, const float *coef, double *grad);
}

using namespace mia_amd;

namespace {

// sum of Gaussian blobs at x - d(x) with a smooth analytic displacement d
// of amplitude amp (no interpolation involved)
CImage blobs(int dim, Size3 s, unsigned seed, double amp)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    struct B {
        double c[3], sig, a;
    };
    std::vector<B> bs(8);
    for (auto &b : bs) {
        b.c[0] = u(rng) * s.x;
        b.c[1] = u(rng) * s.y;
        b.c[2] = u(rng) * s.z;
        b.sig = 0.06 * std::min(s.x, s.y) + 0.06 * std::min(s.x, s.y) * u(rng);
        b.a = 0.3 + 0.7 * u(rng);
    }
    CImage img(dim, s);
    size_t k = 0;
    for (unsigned z = 0; z < s.z; ++z)
        for (unsigned y = 0; y < s.y; ++y)
            for (unsigned x = 0; x < s.x; ++x, ++k) {
                const double p[3] = {x - amp * std::sin(6.28 * y / s.y), y - amp * std::cos(6.28 * x / s.x),
                                     dim == 3 ? z - 0.5 * amp * std::sin(6.28 * x / s.x) : 0.0};
                double v = 0.0;
                for (const auto &b : bs) {
                    double r2 = 0.0;
                    for (int a = 0; a < dim; ++a)
                        r2 += (p[a] - b.c[a]) * (p[a] - b.c[a]);
                    v += b.a * std::exp(-0.5 * r2 / (b.sig * b.sig));
                }
                img.data[k] = (float)v;
            }
    return img;
}

std::vector<double> random_params(size_t n, unsigned seed, double sd)
{
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, sd);
    std::vector<double> p(n);
    for (double &v : p)
        v = std::max(-3.0, std::min(3.0, nd(rng)));
    return p;
}

double rel_inf(const std::vector<double> &a, const std::vector<double> &b)
{
    double d = 0.0, m = 0.0;
    for (size_t k = 0; k < a.size(); ++k) {
        d = std::max(d, std::fabs(a[k] - b[k]));
        m = std::max(m, std::fabs(b[k]));
    }
    return m > 0.0 ? d / m : d;
}

void fullcost_vs_oracle(int dim, Size3 s, float rate, bool norm, double weight)
{
    CDatapool::instance().clear();
    auto src = std::make_shared<CImage>(blobs(dim, s, 1, 0.0));
    auto ref = std::make_shared<CImage>(blobs(dim, s, 1, 2.0));
    CDatapool::instance().add("src.@", src);
    CDatapool::instance().add("ref.@", ref);
    CSSDCost c;
    c.norm = norm;
    CImageFullCost cost(dim, "src.@", "ref.@", c, weight);
    cost.reinit();
    CSplineTransformation t(dim, s, Vec3{rate, rate, rate});
    t.set_parameters(random_params(t.degrees_of_freedom(), 7, 1.0));
    std::vector<double> g;
    const double v = cost.evaluate(t, g);
    const int sz[3] = {(int)s.x, (int)s.y, (int)s.z};
    const Size3 cs = t.get_coeff_size();
    const int c3[3] = {(int)cs.x, (int)cs.y, (int)cs.z};
    std::vector<double> gr(g.size());
    const auto p = t.get_parameters();
    const double vr = mo_cost_grad(dim, sz, src->data.data(), ref->data.data(), c3, p.data(), 3, norm ? 1 : 0, 0.0f,
                                   0, weight, gr.data(), nullptr, nullptr);
    CHECK_CLOSE(v, vr, 1e-3);
    CHECK(rel_inf(g, gr) < 2e-5);
    CHECK_CLOSE(cost.cost_value(t), vr, 1e-3);
}

} // namespace

TEST_CASE(fullcost_matches_oracle_2d)
{
    fullcost_vs_oracle(2, Size3{70, 60, 1}, 8.0f, false, 1.0);
}

TEST_CASE(fullcost_matches_oracle_3d_norm_weight)
{
    fullcost_vs_oracle(3, Size3{36, 30, 28}, 6.0f, true, 0.5);
}

TEST_CASE(costlist_adds_costs)
{
    CDatapool::instance().clear();
    const Size3 s{48, 40, 1};
    CDatapool::instance().add("a.@", std::make_shared<CImage>(blobs(2, s, 3, 0.0)));
    CDatapool::instance().add("b.@", std::make_shared<CImage>(blobs(2, s, 3, 1.5)));
    CDatapool::instance().add("c.@", std::make_shared<CImage>(blobs(2, s, 3, -1.0)));
    auto c1 = produce_full_cost(2, "image:src=a.@,ref=b.@,weight=1");
    auto c2 = produce_full_cost(2, "image:src=a.@,ref=c.@,cost=[ssd:norm=1],weight=0.25");
    CFullCostList list;
    list.push(c1);
    list.push(c2);
    list.reinit();
    Size3 full{0, 0, 0};
    CHECK(list.get_full_size(full) && full == s);
    CSplineTransformation t(2, s, Vec3{8.0f, 8.0f, 1.0f});
    t.set_parameters(random_params(t.degrees_of_freedom(), 5, 0.5));
    std::vector<double> g(t.degrees_of_freedom(), 1.0); // the list adds into g
    const double v = list.evaluate(t, g);
    std::vector<double> g1, g2;
    const double v1 = c1->evaluate(t, g1), v2 = c2->evaluate(t, g2);
    CHECK_CLOSE(v, v1 + v2, 1e-9);
    double worst = 0.0;
    for (size_t k = 0; k < g.size(); ++k)
        worst = std::max(worst, std::fabs(g[k] - (1.0 + g1[k] + g2[k])));
    CHECK(worst < 1e-9);
    CHECK_CLOSE(list.cost_value(t), v1 + v2, 1e-9);
}

TEST_CASE(divcurl_gpu_matches_oracle)
{
    const int cs[3] = {13, 11, 9};
    const float range[3] = {40.0f, 36.0f, 30.0f};
    for (int w = 0; w < 3; ++w) {
        const double wd = w == 2 ? 0.0 : 1.0, wr = w == 1 ? 0.0 : 1.5;
        std::vector<double> cells(9 * 9 * 9 * 6);
        mo_pp_cells(cs, range, wd, wr, cells.data());
        const size_t n = (size_t)cs[0] * cs[1] * cs[2];
        std::vector<float> coef(3 * n);
        std::mt19937_64 rng(17 + w);
        std::normal_distribution<float> nd(0.0f, 1.0f);
        for (float &c : coef)
            c = nd(rng);
        std::vector<double> gr(3 * n), gg(3 * n);
        const double vr = mo_divcurl(cs, cells.data(), coef.data(), gr.data());
        mia_hip_divcurl *p = nullptr;
        const uint32_t csu[3] = {13, 11, 9};
        hip_check(mia_hip_divcurl_create(0, csu, range, wd, wr, &p), "create");
        double v = 0.0;
        hip_check(mia_hip_divcurl_evaluate(p, coef.data(), &v, gg.data()), "evaluate");
        double v2 = 0.0;
        hip_check(mia_hip_divcurl_evaluate(p, coef.data(), &v2, nullptr), "value only");
        mia_hip_divcurl_destroy(p);
        CHECK_CLOSE(v, vr, 1e-9);
        CHECK(v2 == v);
        CHECK(rel_inf(gg, gr) < 1e-11);
        // a quadratic form: c . grad = 2 value
        double cg = 0.0;
        for (size_t k = 0; k < 3 * n; ++k)
            cg += coef[k] * gg[k];
        CHECK_CLOSE(cg, 2.0 * v, 1e-8);
    }
}

TEST_CASE(divcurl_cost_w2_d1_c1)
{
    // mia/3d/splinepenalty/test_divcurl.cc:37-58: size 10^3, range 32^3,
    // weight 2; value 0 for zero coefficients; with (2,3,4) at (5,5,5) the
    // value is 2 * the form and the gradient -2 * its derivative
    const int cs[3] = {10, 10, 10};
    const float range[3] = {32.0f, 32.0f, 32.0f};
    std::vector<double> cells(9 * 9 * 9 * 6);
    mo_pp_cells(cs, range, 1.0, 1.0, cells.data());
    CDivcurlSplinePenalty pen(2.0, false, 1.0, 1.0);
    pen.initialize(3, Size3{10, 10, 10}, Vec3{32.0f, 32.0f, 32.0f});
    std::vector<float> coef(3 * 1000, 0.0f);
    CHECK(pen.value(coef) == 0.0);
    const size_t i = (5 * 10 + 5) * 10 + 5;
    coef[3 * i] = 2.0f;
    coef[3 * i + 1] = 3.0f;
    coef[3 * i + 2] = 4.0f;
    std::vector<double> gref(3 * 1000);
    const double form = mo_divcurl(cs, cells.data(), coef.data(), gref.data());
    CHECK_CLOSE(pen.value(coef), 2.0 * form, 1e-9);
    std::vector<double> g;
    CHECK_CLOSE(pen.value_and_gradient(coef, g), 2.0 * form, 1e-9);
    double worst = 0.0;
    for (size_t k = 0; k < g.size(); ++k)
        worst = std::max(worst, std::fabs(g[k] + 2.0 * gref[k]));
    CHECK(worst < 1e-9 * std::max(1.0, std::fabs(form)));
    auto pn = produce_transform_creator(3, "spline:rate=4,penalty=[divcurl:weight=2,norm=1]");
    (void)pn;
}

TEST_CASE(problem_fdf_adds_penalty)
{
    CDatapool::instance().clear();
    const Size3 s{32, 28, 24};
    CDatapool::instance().add("src.@", std::make_shared<CImage>(blobs(3, s, 9, 0.0)));
    CDatapool::instance().add("ref.@", std::make_shared<CImage>(blobs(3, s, 9, 1.0)));
    CFullCostList list;
    list.push(produce_full_cost(3, "image"));
    list.reinit();
    CSplineTransformCreator cr(3, "spline:rate=6,penalty=[divcurl:weight=0.05]");
    auto t = cr.create(s);
    CHECK(t->has_energy_penalty());
    CNonrigRegGradientProblem prob(list, *t);
    const auto x = random_params(prob.size(), 4, 0.7);
    std::vector<double> g(x.size());
    const double f = prob.fdf(x.size(), x.data(), g.data());
    std::vector<double> gi(x.size(), 0.0), gp;
    t->set_parameters(x);
    const double ci = list.evaluate(*t, gi);
    const double pv = t->get_energy_penalty_and_gradient(gp);
    CHECK(pv > 0.0);
    CHECK_CLOSE(f, ci + pv, 1e-9);
    double worst = 0.0;
    for (size_t k = 0; k < g.size(); ++k)
        worst = std::max(worst, std::fabs(g[k] - (gi[k] + gp[k])));
    CHECK(worst < 1e-9);
    CHECK_CLOSE(prob.f(x.size(), x.data()), f, 1e-9);
    CHECK(prob.get_grad_evals() == 1 && prob.get_func_evals() == 2);
    std::vector<double> bad(3);
    CHECK_THROW(prob.fdf(bad.size(), bad.data(), g.data()), std::invalid_argument);
}

TEST_CASE(scale_image_matches_oracle)
{
    struct C {
        int dim;
        Size3 a, b;
    } cases[] = {{2, {100, 80, 1}, {50, 40, 1}},  {2, {60, 45, 1}, {170, 150, 1}}, {2, {64, 64, 1}, {64, 17, 1}},
                 {3, {40, 36, 32}, {20, 18, 16}}, {3, {24, 20, 16}, {48, 20, 33}}};
    for (const auto &c : cases) {
        const CImage in = blobs(c.dim, c.a, 21, 1.0);
        const CImage out = scale_image(in, c.b);
        std::vector<float> ref(c.b.product());
        const int a3[3] = {(int)c.a.x, (int)c.a.y, (int)c.a.z}, b3[3] = {(int)c.b.x, (int)c.b.y, (int)c.b.z};
        mo_scale_image(in.data.data(), a3, b3, ref.data());
        double d = 0.0, m = 0.0;
        for (size_t k = 0; k < ref.size(); ++k) {
            d = std::max(d, (double)std::fabs(out.data[k] - ref[k]));
            m = std::max(m, (double)std::fabs(ref[k]));
        }
        CHECK(d <= 1e-6 * m);
    }
    const CImage same = blobs(3, Size3{10, 10, 10}, 2, 0.0);
    CHECK(scale_image(same, Size3{10, 10, 10}).data == same.data);
}

TEST_CASE(warp_matches_oracle)
{
    for (int dim = 2; dim <= 3; ++dim) {
        const Size3 s = dim == 2 ? Size3{64, 50, 1} : Size3{30, 28, 26};
        const CImage img = blobs(dim, s, 31, 0.0);
        for (int deg : {1, 3}) {
            CSplineTransformation t(dim, s, Vec3{7.0f, 7.0f, 7.0f}, deg);
            t.set_parameters(random_params(t.degrees_of_freedom(), 8, 1.2));
            const CImage w = t(img);
            std::vector<double> wr(s.product());
            const int sz[3] = {(int)s.x, (int)s.y, (int)s.z};
            const Size3 cs = t.get_coeff_size();
            const int c3[3] = {(int)cs.x, (int)cs.y, (int)cs.z};
            const auto p = t.get_parameters();
            mo_cost_grad(dim, sz, img.data.data(), img.data.data(), c3, p.data(), deg, 0, 0.0f, 0, 1.0, nullptr,
                         wr.data(), nullptr);
            double d = 0.0;
            for (size_t k = 0; k < wr.size(); ++k)
                d = std::max(d, std::fabs(w.data[k] - wr[k]));
            CHECK(d < 2e-5);
        }
    }
}

static void registration_case(int dim, Size3 s, const char *tdescr, int levels, const char *minimizer)
{
    CDatapool::instance().clear();
    const CImage src = blobs(dim, s, 41, 0.0), ref = blobs(dim, s, 41, 2.5);
    CFullCostList list;
    list.push(produce_full_cost(dim, "image:cost=[ssd]"));
    CSplineTransformCreator cr(dim, tdescr);
    CNonrigidRegister reg(list, minimizer, cr, levels);
    const auto t0 = std::chrono::steady_clock::now();
    auto t = reg.run(src, ref);
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const auto &st = reg.get_level_stats();
    CHECK((int)st.size() == levels);
    CHECK(st.back().size == s);
    for (int l = 0; l < levels; ++l) {
        const unsigned f = 1u << (levels - 1 - l);
        CHECK(st[l].size.x == s.x / f && st[l].size.y == s.y / f);
        std::printf("  level %d: %ux%ux%u, grid %ux%ux%u, %d evaluations, cost %.6g, %.3f s\n", l, st[l].size.x,
                    st[l].size.y, st[l].size.z, st[l].coeff_size.x, st[l].coeff_size.y, st[l].coeff_size.z,
                    st[l].evaluations, st[l].cost, st[l].seconds);
    }
    // the final cost against the identity on the normalised full-size images
    CSplineTransformation id(*t);
    id.set_identity();
    list.set_size(s);
    const double c0 = list.cost_value(id) + id.get_energy_penalty();
    const double c1 = list.cost_value(*t) + t->get_energy_penalty();
    std::printf("  identity cost %.6g -> registered %.6g (%.2f s)\n", c0, c1, sec);
    CHECK(c1 < 0.5 * c0);
    CHECK(t->get_coeff_size() == cr.create(s)->get_coeff_size());
}

TEST_CASE(registration_2d_three_levels)
{
    registration_case(2, Size3{128, 112, 1}, "spline:rate=8", 3, "lbfgs:maxiter=40");
}

TEST_CASE(registration_3d_two_levels_divcurl)
{
    registration_case(3, Size3{48, 44, 40}, "spline:rate=8,penalty=[divcurl:weight=0.01]", 2, "lbfgs:maxiter=30");
}

TEST_CASE(registration_2d_gdsq)
{
    registration_case(2, Size3{80, 80, 1}, "spline:rate=8", 2, "gdsq:maxiter=40");
}

CHECK_MAIN()

// CPU test of the reverse-communication minimizers (mia_amd/host/minimizer.hh)
// on problems with known minimizers; prints one line per check, exits non-zero
// on failure.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../mia_amd/host/minimizer.hh"

using namespace mia_amd;

static int failures = 0;
#define CHECK(cond, ...)                                                                                     \
    do {                                                                                                     \
        if (!(cond)) {                                                                                       \
            std::printf("FAIL %s:%d ", __FILE__, __LINE__);                                                  \
            std::printf(__VA_ARGS__);                                                                        \
            std::printf("\n");                                                                               \
            ++failures;                                                                                      \
        }                                                                                                    \
    } while (0)

// f(x) = 0.5 sum_i a_i (x_i - c_i)^2, a_i in [1, 50]
struct Quad {
    std::vector<double> a, c;
    explicit Quad(int n) : a(n), c(n)
    {
        for (int i = 0; i < n; ++i) {
            a[i] = 1.0 + 49.0 * ((i * 37) % 17) / 16.0;
            c[i] = std::sin(0.7 * i);
        }
    }
    double fdf(const double *x, double *g) const
    {
        double f = 0.0;
        for (size_t i = 0; i < a.size(); ++i) {
            const double d = x[i] - c[i];
            f += 0.5 * a[i] * d * d;
            g[i] = a[i] * d;
        }
        return f;
    }
};

// Rosenbrock in 2D
static double rosen(const double *x, double *g)
{
    const double a = 1.0 - x[0], b = x[1] - x[0] * x[0];
    g[0] = -2.0 * a - 400.0 * x[0] * b;
    g[1] = 200.0 * b;
    return a * a + 100.0 * b * b;
}

template <typename F>
static void run(ReverseMinimizer &m, const std::vector<double> &x0, F fdf, size_t n)
{
    std::vector<double> g(n);
    m.start(x0.data());
    while (!m.done()) {
        const double f = fdf(m.point(), g.data());
        m.tell(f, g.data());
    }
}

int main()
{
    {
        const int n = 40;
        Quad q(n);
        LBFGSMinimizer m(n, 200, 10, 1e-10);
        run(m, std::vector<double>(n, 0.0), [&](const double *x, double *g) { return q.fdf(x, g); }, n);
        CHECK(m.best_f() < 1e-12, "lbfgs quadratic f=%g", m.best_f());
        double err = 0.0;
        for (int i = 0; i < n; ++i)
            err = std::max(err, std::fabs(m.best_x()[i] - q.c[i]));
        CHECK(err < 1e-6, "lbfgs quadratic |x - c| = %g", err);
        CHECK(m.evaluations() < 120, "lbfgs quadratic evals %d", m.evaluations());
        std::printf("lbfgs quadratic: f=%.3e evals=%d err=%.2e\n", m.best_f(), m.evaluations(), err);
    }
    {
        LBFGSMinimizer m(2, 500, 10, 1e-12);
        run(m, {-1.2, 1.0}, rosen, 2);
        CHECK(std::fabs(m.best_x()[0] - 1.0) < 1e-4 && std::fabs(m.best_x()[1] - 1.0) < 1e-4,
              "lbfgs rosenbrock x=(%g,%g)", m.best_x()[0], m.best_x()[1]);
        std::printf("lbfgs rosenbrock: f=%.3e evals=%d\n", m.best_f(), m.evaluations());
    }
    {
        // maxeval is a hard cap on evaluations
        const int n = 40;
        Quad q(n);
        LBFGSMinimizer m(n, 7);
        run(m, std::vector<double>(n, 0.0), [&](const double *x, double *g) { return q.fdf(x, g); }, n);
        CHECK(m.evaluations() == 7, "lbfgs maxeval evals=%d", m.evaluations());
    }
    {
        // gdsq trajectory on f = 0.5 (x - 2)^2 from x = 0, worked by hand from
        // the reference loop (gdsq.cc:60-122): x1 = 0.2 accepted, new step
        // 0.5*|s|/aux = 0.5*0.2/0.02 = 5; x = 9.2 and 4.7 rejected (step 2.5,
        // 1.25); x = 2.45 accepted.
        auto f = [](const double *x, double *g) {
            g[0] = x[0] - 2.0;
            return 0.5 * g[0] * g[0];
        };
        GDSQMinimizer m(1, 4);
        std::vector<double> seen;
        std::vector<double> g(1);
        m.start(std::vector<double>{0.0}.data());
        while (!m.done()) {
            seen.push_back(m.point()[0]);
            const double v = f(m.point(), g.data());
            m.tell(v, g.data());
        }
        const double expect[] = {0.0, 0.2, 9.2, 4.7, 2.45};
        CHECK(seen.size() == 5, "gdsq evaluations %zu", seen.size());
        for (size_t k = 0; k < seen.size() && k < 5; ++k)
            CHECK(std::fabs(seen[k] - expect[k]) < 1e-12, "gdsq point %zu: %.15g vs %g", k, seen[k], expect[k]);
        CHECK(std::fabs(m.best_x()[0] - 2.45) < 1e-12, "gdsq best x %g", m.best_x()[0]);
        std::printf("gdsq hand trajectory: %zu points, best x=%.4f\n", seen.size(), m.best_x()[0]);
    }
    {
        // gdsq stops when the step falls below 1e-6 on a non-improving
        // problem: every trial step is rejected and halves the step
        auto f = [](const double *x, double *g) {
            g[0] = 1.0;
            return x[0] == 0.0 ? 0.0 : 1.0;
        };
        GDSQMinimizer m(1, 1000);
        run(m, {0.0}, f, 1);
        // 0.1 / 2^k <= 1e-6 first at k = 17
        CHECK(m.evaluations() == 1 + 17, "gdsq min-step stop evals=%d", m.evaluations());
        CHECK(m.best_x()[0] == 0.0, "gdsq keeps the best point");
    }
    std::printf("%s\n", failures ? "FAILED" : "OK");
    return failures ? 1 : 0;
}

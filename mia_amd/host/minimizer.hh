// Gradient-based minimizers in reverse-communication form: the caller asks
// for the point to evaluate (point()), computes cost and gradient there and
// hands them back (tell()).  This lets many independent registrations step in
// lockstep through one batched GPU evaluation while each keeps exactly the
// trajectory it would have on its own.
//
// The reference's minimizer contract is CMinimizer::Problem::fdf(n, x, g)
// (mia/core/minimizer.hh:81-177, mia/core/minimizer.cc:113-134): the problem
// returns the cost and writes the gradient.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <limits>
#include <string>
#include <vector>

namespace mia_amd {

class ReverseMinimizer {
public:
    virtual ~ReverseMinimizer() = default;
    virtual void start(const double *x0) = 0;
    virtual const double *point() const = 0;
    virtual void tell(double f, const double *g) = 0;
    virtual bool done() const = 0;
    const double *best_x() const { return m_best_x.data(); }
    double best_f() const { return m_best_f; }
    int evaluations() const { return m_evals; }

protected:
    void record(double f, const double *x, size_t n)
    {
        ++m_evals;
        if (std::isfinite(f) && f < m_best_f) {
            m_best_f = f;
            m_best_x.assign(x, x + n);
        }
    }
    std::vector<double> m_best_x;
    double m_best_f = std::numeric_limits<double>::infinity();
    int m_evals = 0;
};

// Gradient descent with quadratic step estimation, MIA's `gdsq` plugin
// (mia/core/minimizer/gdsq.cc:60-122, defaults :125-133: step 0.1, scale 2,
// xtola = gtola = ftolr = 0, maxiter 100, min step 1e-6).  One evaluation at
// the start and one per iteration, as in the reference loop.
class GDSQMinimizer : public ReverseMinimizer {
public:
    GDSQMinimizer(size_t n, unsigned maxiter = 100, double step = 0.1, double scale = 2.0, double xtol = 0.0,
                  double gtol = 0.0, double ftolr = 0.0, double min_step = 1e-6)
        : m_n(n), m_maxiter(maxiter), m_step0(step), m_scale(scale), m_xtol(xtol), m_gtol(gtol), m_ftolr(ftolr),
          m_min_step(min_step), m_x(n), m_xwork(n), m_g(n)
    {
    }

    void start(const double *x0) override
    {
        m_x.assign(x0, x0 + m_n);
        m_xwork = m_x;
        m_state = kInit;
        m_done = false;
    }
    const double *point() const override { return m_xwork.data(); }
    bool done() const override { return m_done; }

    void tell(double f, const double *gt) override
    {
        record(f, m_xwork.data(), m_n);
        if (m_state == kInit) {
            m_g.assign(gt, gt + m_n);
            m_step = m_step0;
            m_f_old = f;
            m_iter = 0;
            m_success = test_tol(m_g, m_xtol);
            m_state = kStep;
            begin_iteration();
            return;
        }
        // evaluation at xwork = x - step * g (gdsq.cc:73-74)
        if (f < m_f_old) {
            double sum = 0.0, nrm = 0.0;
            for (size_t i = 0; i < m_n; ++i) {
                m_g[i] *= -m_step; // cblas_dscal
                sum += m_g[i] * gt[i];
                nrm += m_g[i] * m_g[i];
            }
            const double aux = m_f_old - f + sum;
            if (aux <= 0.0)
                m_step *= m_scale;
            else
                m_step = 0.5 * std::sqrt(nrm) / aux;
            m_x = m_xwork;
            m_g.assign(gt, gt + m_n);
            const double frel = m_f_old - f;
            m_f_old = f;
            m_success |= test_tol(m_g, m_gtol);
            if (frel < m_ftolr)
                m_success |= 1;
        } else {
            m_xwork = m_x;
            m_step /= m_scale;
        }
        m_success |= test_tol(m_g, m_xtol / m_step);
        begin_iteration();
    }

private:
    enum { kInit, kStep };
    bool test_tol(const std::vector<double> &v, double tol) const
    {
        for (double e : v)
            if (std::fabs(e) >= tol)
                return false;
        return true;
    }
    // while (iter++ < maxiter && step > min_step && !success) { xwork -= step*g; fdf }
    void begin_iteration()
    {
        if (m_iter++ < m_maxiter && m_step > m_min_step && !m_success) {
            for (size_t i = 0; i < m_n; ++i)
                m_xwork[i] -= m_step * m_g[i]; // cblas_daxpy
        } else {
            m_done = true;
            m_xwork = m_x;
        }
    }

    size_t m_n;
    unsigned m_maxiter;
    double m_step0, m_scale, m_xtol, m_gtol, m_ftolr, m_min_step;
    std::vector<double> m_x, m_xwork, m_g;
    double m_step = 0.0, m_f_old = 0.0;
    unsigned m_iter = 0;
    bool m_success = false;
    int m_state = kInit;
    bool m_done = false;
};

// Limited-memory BFGS (two-loop recursion, Nocedal & Wright, Numerical
// Optimization, 2nd ed., Alg. 7.4/7.5) with a backtracking Armijo line search
// (sufficient-decrease constant 1e-4, safeguarded quadratic interpolation).
// `maxeval` counts every cost/gradient evaluation, like NLopt's maxeval that
// MIA's `maxiter` maps to (addons/nlopt/nlopt.cc:237,347).  The first step is
// scaled so that no parameter moves by more than `first_step`.
class LBFGSMinimizer : public ReverseMinimizer {
public:
    LBFGSMinimizer(size_t n, int maxeval = 100, int memory = 10, double gtol = 0.0, double ftolr = 0.0,
                   double first_step = 1.0)
        : m_n(n), m_maxeval(maxeval), m_m(memory), m_gtol(gtol), m_ftolr(ftolr), m_first(first_step), m_x(n),
          m_g(n), m_d(n), m_xt(n)
    {
    }

    void start(const double *x0) override
    {
        m_x.assign(x0, x0 + m_n);
        m_xt = m_x;
        m_S.clear();
        m_Y.clear();
        m_rho.clear();
        m_state = kInit;
        m_done = false;
    }
    const double *point() const override { return m_xt.data(); }
    bool done() const override { return m_done; }

    void tell(double f, const double *g) override
    {
        record(f, m_xt.data(), m_n);
        if (m_state == kInit) {
            m_f = f;
            m_g.assign(g, g + m_n);
            if (!std::isfinite(f) || converged_g() || m_evals >= m_maxeval) {
                finish();
                return;
            }
            new_direction(true);
            return;
        }
        const bool armijo = std::isfinite(f) && f <= m_f + 1e-4 * m_alpha * m_gd;
        if (armijo) {
            const double f_old = m_f;
            double sy = 0.0, yy = 0.0, ss = 0.0;
            std::vector<double> s(m_n), y(m_n);
            for (size_t i = 0; i < m_n; ++i) {
                s[i] = m_xt[i] - m_x[i];
                y[i] = g[i] - m_g[i];
                sy += s[i] * y[i];
                yy += y[i] * y[i];
                ss += s[i] * s[i];
            }
            if (sy > 1e-12 * std::sqrt(ss * yy)) {
                if ((int)m_S.size() == m_m) {
                    m_S.erase(m_S.begin());
                    m_Y.erase(m_Y.begin());
                    m_rho.erase(m_rho.begin());
                }
                m_S.push_back(std::move(s));
                m_Y.push_back(std::move(y));
                m_rho.push_back(1.0 / sy);
            }
            m_x = m_xt;
            m_f = f;
            m_g.assign(g, g + m_n);
            if (m_evals >= m_maxeval || converged_g() || std::fabs(f_old - f) <= m_ftolr * std::fabs(f_old)) {
                finish();
                return;
            }
            new_direction(false);
            return;
        }
        if (m_evals >= m_maxeval) {
            finish();
            return;
        }
        // backtrack: minimiser of the quadratic through f(0), f'(0), f(alpha)
        double a = m_alpha;
        double an = 0.5 * a;
        if (std::isfinite(f)) {
            const double den = 2.0 * (f - m_f - a * m_gd);
            if (den > 0.0)
                an = -m_gd * a * a / den;
        }
        an = std::min(std::max(an, 0.1 * a), 0.5 * a);
        if (an < 1e-20) {
            finish();
            return;
        }
        m_alpha = an;
        for (size_t i = 0; i < m_n; ++i)
            m_xt[i] = m_x[i] + m_alpha * m_d[i];
    }

private:
    enum { kInit, kSearch };
    bool converged_g() const
    {
        double gmax = 0.0;
        for (double e : m_g)
            gmax = std::max(gmax, std::fabs(e));
        return gmax <= m_gtol;
    }
    void finish()
    {
        m_done = true;
        m_xt = m_x;
    }
    // d = -H g by the two-loop recursion, H0 = (s'y / y'y) I
    void new_direction(bool first)
    {
        std::vector<double> q(m_g);
        const int k = (int)m_S.size();
        std::vector<double> al(k);
        for (int j = k - 1; j >= 0; --j) {
            double a = 0.0;
            for (size_t i = 0; i < m_n; ++i)
                a += m_S[j][i] * q[i];
            a *= m_rho[j];
            al[j] = a;
            for (size_t i = 0; i < m_n; ++i)
                q[i] -= a * m_Y[j][i];
        }
        if (k > 0) {
            double yy = 0.0;
            for (size_t i = 0; i < m_n; ++i)
                yy += m_Y[k - 1][i] * m_Y[k - 1][i];
            const double gamma = 1.0 / (m_rho[k - 1] * yy);
            for (size_t i = 0; i < m_n; ++i)
                q[i] *= gamma;
        }
        for (int j = 0; j < k; ++j) {
            double b = 0.0;
            for (size_t i = 0; i < m_n; ++i)
                b += m_Y[j][i] * q[i];
            b *= m_rho[j];
            for (size_t i = 0; i < m_n; ++i)
                q[i] += (al[j] - b) * m_S[j][i];
        }
        double gd = 0.0;
        for (size_t i = 0; i < m_n; ++i) {
            m_d[i] = -q[i];
            gd += m_g[i] * m_d[i];
        }
        if (!(gd < 0.0)) { // not a descent direction: restart from steepest descent
            m_S.clear();
            m_Y.clear();
            m_rho.clear();
            gd = 0.0;
            for (size_t i = 0; i < m_n; ++i) {
                m_d[i] = -m_g[i];
                gd += m_g[i] * m_d[i];
            }
            first = true;
        }
        m_gd = gd;
        m_alpha = 1.0;
        if (first) {
            double dmax = 0.0;
            for (double e : m_d)
                dmax = std::max(dmax, std::fabs(e));
            if (dmax > 0.0)
                m_alpha = std::min(1.0, m_first / dmax);
        }
        for (size_t i = 0; i < m_n; ++i)
            m_xt[i] = m_x[i] + m_alpha * m_d[i];
        m_state = kSearch;
    }

    size_t m_n;
    int m_maxeval, m_m;
    double m_gtol, m_ftolr, m_first;
    std::vector<double> m_x, m_g, m_d, m_xt;
    std::vector<std::vector<double>> m_S, m_Y;
    std::vector<double> m_rho;
    double m_f = 0.0, m_gd = 0.0, m_alpha = 1.0;
    int m_state = kInit;
    bool m_done = false;
};

// "lbfgs:maxiter=100" / "gdsq:maxiter=100,step=0.1" (the plugin descriptor
// grammar name:key=value,..., mia/core/optionparser.cc:38).  Returns nullptr
// and sets `err` for unknown names or keys.
ReverseMinimizer *produce_minimizer(const std::string &descr, size_t n, std::string &err);

// Whether `descr` is an lbfgs the on-device L-BFGS (mia_hip_optimize) runs:
// keys maxiter, memory, step, gtol = 0, ftolr = 0 and host = 0 (host=1 asks
// for the host minimizer).  Sets maxeval, memory and step; false otherwise
// (err set only for a malformed descriptor or a bad value).
bool device_lbfgs_params(const std::string &descr, int &maxeval, int &memory, double &step, std::string &err);

} // namespace mia_amd

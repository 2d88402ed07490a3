// mia_cpu_port.cc -- FP32 CPU restatement of MIA's own data flow for one
// image:cost=ssd + spline cost/gradient evaluation, used ONLY as the bench's
// CPU baseline (timed beside the GPU, never a product path).  The reference
// itself cannot be built here (Boost, GSL and NLopt are absent), so this port
// keeps its structure and its thread decomposition:
//   * C3DSplineTransformation::set_parameters -> init_grid: separable x, y, z
//     passes over slices, single-threaded (mia/3d/transform/spline.cc:344-358,
//     541-653)
//   * the interpolator is constructed (copy + 3-axis prefilter, pfor over z
//     and y) on every warp (mia/3d/transform.cc:306, interpolator.cxx:98-164)
//   * warp: pfor over z slabs, per-voxel 4x4x4 float contraction with clamp
//     (mia/3d/transform.cc:308-353, mia/3d/interpolator.cxx:250-287)
//   * get_gradient, SSD force, translate and the second SSD pass serial
//     (mia/3d/image.cc:347-477, mia/template/ssd.hh:54-181,
//     mia/3d/transform/spline.cc:731-791, mia/3d/fullcost/image.cc:67-78)
// pfor uses `nthreads` std::threads (the reference uses hardware_concurrency
// tasks, mia/core/parallelcxx11.cc:25-31).
//
// The 2D evaluation follows the 2D classes: per-pixel 4x4 displacement sums
// (mia/2d/transform/spline.cc:640-666), the interpolator prefiltered on every
// warp (mia/2d/interpolator.cxx:66-132), central-difference gradient, force
// with the float SSD sum that the 2D full cost returns
// (mia/2d/fullcost/image.cc:81-105) and the 2D translate
// (mia/2d/transform/spline.cc:509-576).  The series baseline registers
// frames in a pfor over frames, one frame per task, as
// mia-2dmany2one-nonrigid does (src/2dmany2one-nonrigid.cc:102-118,185),
// with the same L-BFGS as the GPU host path (NLopt is not available).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <atomic>
#include <thread>
#include <vector>

#include "../mia_amd/host/minimizer.hh"

namespace {

struct V3 {
    float x, y, z;
};

template <typename F>
void pfor(int n, int nthreads, F f)
{
    if (nthreads <= 1 || n < 2) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        const int b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e)
            break;
        th.emplace_back([=]() { f(b, e); });
    }
    for (auto &t : th)
        t.join();
}

int fastfloor(double v)
{
    return (int)std::floor(std::nearbyint(v * 65536.0) / 65536.0);
}

void bspline3(double x, double w[4])
{
    const double xm1 = 1.0 - x;
    w[3] = x * x * x / 6.0;
    w[0] = xm1 * xm1 * xm1 / 6.0;
    w[2] = x - 2.0 * w[3] + w[0];
    w[1] = xm1 - 2.0 * w[0] + w[3];
}

double bspline3_value(double x)
{
    x = std::fabs(x);
    if (x <= 1.0)
        return 2.0 / 3.0 - x * x * (1 - 0.5 * x);
    if (x < 2.0) {
        const double d = 2.0 - x;
        return d * d * d / 6.0;
    }
    return 0.0;
}

int mirror(int i, int n)
{
    if (n == 1)
        return 0;
    const int n2 = 2 * n - 2;
    int idx = i < 0 ? -i : i;
    idx = idx < n2 ? idx : idx % n2;
    if (idx >= n)
        idx = n2 - idx;
    return idx;
}

void filter_line(std::vector<double> &c)
{
    const int n = (int)c.size();
    if (n < 2)
        return;
    const double p = std::sqrt(3.0) - 2.0;
    const double lambda = 2.0 - p - 1.0 / p;
    for (auto &v : c)
        v *= lambda;
    double zn = p, iz = 1.0 / p, z2n = std::pow(p, (double)(n - 1));
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

struct Axis {
    std::vector<int> start;
    std::vector<std::vector<double>> w;
    std::vector<int> row_first;
    std::vector<std::vector<double>> row;
};

Axis make_axis(int n, int cs)
{
    Axis a;
    const float scale = (float)(cs - 3) / (float)(n - 1);
    a.start.resize(n);
    a.w.resize(n);
    for (int i = 0; i < n; ++i) {
        const float pos = (float)i * scale + 1.0f;
        const int f = fastfloor(pos);
        double w[4];
        bspline3((double)pos - f, w);
        a.start[i] = f - 1;
        a.w[i].assign(w, w + (i == n - 1 ? 3 : 4));
    }
    // adjoint rows, get_derivative_row (spline.cc:697-729)
    a.row_first.assign(cs, -1);
    a.row.resize(cs);
    for (int o = 0; o < cs; ++o)
        for (int i = 0; i < n; ++i) {
            const double x = i * (double)scale + 1.0 - o;
            if (std::fabs(x) < 2.0) {
                if (a.row_first[o] < 0)
                    a.row_first[o] = i;
                const double y = bspline3_value(x);
                if (y > 0)
                    a.row[o].push_back(y);
                else
                    break;
            }
        }
    return a;
}

void saxpy(int n, float a, const float *x, float *y)
{
    for (int i = 0; i < n; ++i)
        y[i] += a * x[i];
}

} // namespace

extern "C" {

// One full evaluation; returns the cost, writes grad (3 * cx*cy*cz doubles).
// timings (optional, 7 entries, seconds): init_grid, prefilter, warp,
// gradient, force, translate, value.
double mcp_cost_grad_3d(const float *mov, const float *fix, int nx, int ny, int nz, int cx, int cy, int cz,
                        const double *params, int nthreads, double *grad, double *timings)
{
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    const size_t N = (size_t)nx * ny * nz;
    auto t0 = clk::now();
    // --- set_parameters + init_grid -------------------------------------------
    std::vector<V3> coef((size_t)cx * cy * cz);
    for (size_t k = 0; k < coef.size(); ++k)
        coef[k] = {(float)params[3 * k], (float)params[3 * k + 1], (float)params[3 * k + 2]};
    Axis ax = make_axis(nx, cx), ay = make_axis(ny, cy), az = make_axis(nz, cz);
    // x pass: tmp(x, ky, kz) = sum_i wx_i coef(sx+i, ky, kz)
    std::vector<V3> tmp((size_t)nx * cy * cz);
    {
        std::vector<float> slice((size_t)cy * cz * 3), out((size_t)cy * cz * 3);
        for (int x = 0; x < nx; ++x) {
            std::fill(out.begin(), out.end(), 0.0f);
            for (size_t i = 0; i < ax.w[x].size(); ++i) {
                const int kx = ax.start[x] + (int)i;
                if (kx >= cx)
                    break;
                for (int kz = 0; kz < cz; ++kz)
                    for (int ky = 0; ky < cy; ++ky) {
                        const V3 &c = coef[((size_t)kz * cy + ky) * cx + kx];
                        float *o = &slice[3 * ((size_t)kz * cy + ky)];
                        o[0] = c.x;
                        o[1] = c.y;
                        o[2] = c.z;
                    }
                saxpy((int)out.size(), (float)ax.w[x][i], slice.data(), out.data());
            }
            for (int kz = 0; kz < cz; ++kz)
                for (int ky = 0; ky < cy; ++ky) {
                    const float *o = &out[3 * ((size_t)kz * cy + ky)];
                    tmp[((size_t)kz * cy + ky) * nx + x] = {o[0], o[1], o[2]};
                }
        }
    }
    // y pass: tmp2(x, y, kz)
    std::vector<V3> tmp2((size_t)nx * ny * cz);
    {
        std::vector<float> out((size_t)nx * cz * 3), slice((size_t)nx * cz * 3);
        for (int y = 0; y < ny; ++y) {
            std::fill(out.begin(), out.end(), 0.0f);
            for (size_t i = 0; i < ay.w[y].size(); ++i) {
                const int ky = ay.start[y] + (int)i;
                if (ky >= cy)
                    break;
                for (int kz = 0; kz < cz; ++kz)
                    std::memcpy(&slice[3 * (size_t)kz * nx], &tmp[((size_t)kz * cy + ky) * nx], sizeof(V3) * nx);
                saxpy((int)out.size(), (float)ay.w[y][i], slice.data(), out.data());
            }
            for (int kz = 0; kz < cz; ++kz)
                std::memcpy(&tmp2[((size_t)kz * ny + y) * nx], &out[3 * (size_t)kz * nx], sizeof(V3) * nx);
        }
    }
    // z pass into the position field T(x) = x - u(x)
    std::vector<V3> grid(N);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                grid[((size_t)z * ny + y) * nx + x] = {(float)x, (float)y, (float)z};
    for (int z = 0; z < nz; ++z)
        for (size_t i = 0; i < az.w[z].size(); ++i) {
            const int kz = az.start[z] + (int)i;
            if (kz >= cz)
                break;
            saxpy((int)(3 * (size_t)nx * ny), -(float)az.w[z][i], &tmp2[(size_t)kz * nx * ny].x,
                  &grid[(size_t)z * nx * ny].x);
        }
    auto t1 = clk::now();
    // --- interpolator construction (every warp) --------------------------------
    std::vector<float> cf(mov, mov + N);
    float vmin = mov[0], vmax = mov[0];
    for (size_t i = 0; i < N; ++i) {
        vmin = std::min(vmin, mov[i]);
        vmax = std::max(vmax, mov[i]);
    }
    if (0.0f < vmin)
        vmin = 0.0f;
    pfor(nz, nthreads, [&](int zb, int ze) {
        std::vector<double> buf(nx);
        for (int z = zb; z < ze; ++z)
            for (int y = 0; y < ny; ++y) {
                float *row = &cf[((size_t)z * ny + y) * nx];
                for (int x = 0; x < nx; ++x)
                    buf[x] = row[x];
                filter_line(buf);
                for (int x = 0; x < nx; ++x)
                    row[x] = (float)buf[x];
            }
    });
    pfor(nz, nthreads, [&](int zb, int ze) {
        std::vector<double> buf(ny);
        for (int z = zb; z < ze; ++z)
            for (int x = 0; x < nx; ++x) {
                for (int y = 0; y < ny; ++y)
                    buf[y] = cf[((size_t)z * ny + y) * nx + x];
                filter_line(buf);
                for (int y = 0; y < ny; ++y)
                    cf[((size_t)z * ny + y) * nx + x] = (float)buf[y];
            }
    });
    pfor(ny, nthreads, [&](int yb, int ye) {
        std::vector<double> buf(nz);
        for (int y = yb; y < ye; ++y)
            for (int x = 0; x < nx; ++x) {
                for (int z = 0; z < nz; ++z)
                    buf[z] = cf[((size_t)z * ny + y) * nx + x];
                filter_line(buf);
                for (int z = 0; z < nz; ++z)
                    cf[((size_t)z * ny + y) * nx + x] = (float)buf[z];
            }
    });
    auto t2 = clk::now();
    // --- warp, pfor over z ----------------------------------------------------------
    std::vector<float> W(N);
    pfor(nz, nthreads, [&](int zb, int ze) {
        for (int z = zb; z < ze; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const size_t idx = ((size_t)z * ny + y) * nx + x;
                    const V3 p = grid[idx];
                    int s[3];
                    double wd[3][4];
                    const float pc[3] = {p.x, p.y, p.z};
                    for (int a = 0; a < 3; ++a) {
                        const int f = fastfloor(pc[a]);
                        bspline3((double)pc[a] - f, wd[a]);
                        s[a] = f - 1;
                    }
                    int ix[4], iy[4], iz[4];
                    for (int k = 0; k < 4; ++k) {
                        ix[k] = mirror(s[0] + k, nx);
                        iy[k] = mirror(s[1] + k, ny);
                        iz[k] = mirror(s[2] + k, nz);
                    }
                    float r = 0.0f;
                    for (int c = 0; c < 4; ++c) {
                        float ry = 0.0f;
                        for (int b = 0; b < 4; ++b) {
                            const float *row = &cf[((size_t)iz[c] * ny + iy[b]) * nx];
                            float rx = 0.0f;
                            for (int a = 0; a < 4; ++a)
                                rx += (float)wd[0][a] * row[ix[a]];
                            ry += (float)wd[1][b] * rx;
                        }
                        r += (float)wd[2][c] * ry;
                    }
                    W[idx] = r >= vmin ? (r <= vmax ? r : vmax) : vmin;
                }
    });
    auto t3 = clk::now();
    // --- get_gradient (serial) --------------------------------------------------------
    std::vector<V3> g(N);
    const size_t sy = nx, sz = (size_t)nx * ny;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t i = ((size_t)z * ny + y) * nx + x;
                g[i].x = (x > 0 && x < nx - 1) ? 0.5f * (W[i + 1] - W[i - 1]) : 0.0f;
                g[i].y = (y > 0 && y < ny - 1) ? 0.5f * (W[i + sy] - W[i - sy]) : 0.0f;
                g[i].z = (z > 0 && z < nz - 1) ? 0.5f * (W[i + sz] - W[i - sz]) : 0.0f;
            }
    auto t4 = clk::now();
    // --- force (serial, float) ---------------------------------------------------------
    std::vector<V3> force(N);
    float fcost = 0.0f;
    for (size_t i = 0; i < N; ++i) {
        const float d = W[i] - fix[i];
        force[i] = {g[i].x * d, g[i].y * d, g[i].z * d};
        fcost += d * d;
    }
    (void)fcost;
    auto t5 = clk::now();
    // --- translate (serial, separable z -> y -> x) --------------------------------------
    std::vector<V3> t1v((size_t)nx * ny * cz, V3{0, 0, 0});
    for (int oz = 0; oz < cz; ++oz) {
        int i = az.row_first[oz];
        for (double w : az.row[oz]) {
            saxpy((int)(3 * (size_t)nx * ny), (float)w, &force[(size_t)i * nx * ny].x, &t1v[(size_t)oz * nx * ny].x);
            ++i;
        }
    }
    std::vector<V3> t2v((size_t)nx * cy * cz);
    {
        std::vector<V3> in(ny);
        for (int oz = 0; oz < cz; ++oz)
            for (int x = 0; x < nx; ++x) {
                for (int y = 0; y < ny; ++y)
                    in[y] = t1v[((size_t)oz * ny + y) * nx + x];
                for (int oy = 0; oy < cy; ++oy) {
                    V3 acc{0, 0, 0};
                    int j = ay.row_first[oy];
                    for (double w : ay.row[oy]) {
                        acc.x += (float)w * in[j].x;
                        acc.y += (float)w * in[j].y;
                        acc.z += (float)w * in[j].z;
                        ++j;
                    }
                    t2v[((size_t)oz * cy + oy) * nx + x] = acc;
                }
            }
    }
    for (int oz = 0; oz < cz; ++oz)
        for (int oy = 0; oy < cy; ++oy) {
            const V3 *in = &t2v[((size_t)oz * cy + oy) * nx];
            for (int ox = 0; ox < cx; ++ox) {
                V3 acc{0, 0, 0};
                int j = ax.row_first[ox];
                for (double w : ax.row[ox]) {
                    acc.x += (float)w * in[j].x;
                    acc.y += (float)w * in[j].y;
                    acc.z += (float)w * in[j].z;
                    ++j;
                }
                double *r = grad + 3 * (((size_t)oz * cy + oy) * cx + ox);
                r[0] = -acc.x;
                r[1] = -acc.y;
                r[2] = -acc.z;
            }
        }
    auto t6 = clk::now();
    // --- value (second pass, double) ---------------------------------------------------
    double sum = 0.0;
    for (size_t i = 0; i < N; ++i) {
        const double d = (double)W[i] - (double)fix[i];
        sum += d * d;
    }
    auto t7 = clk::now();
    if (timings) {
        timings[0] = secs(t0, t1);
        timings[1] = secs(t1, t2);
        timings[2] = secs(t2, t3);
        timings[3] = secs(t3, t4);
        timings[4] = secs(t4, t5);
        timings[5] = secs(t5, t6);
        timings[6] = secs(t6, t7);
    }
    return 0.5 * sum;
}


// One 2D evaluation (single thread); returns the float-accumulated cost the
// 2D full cost reports, writes grad (2 * cx * cy doubles).
double mcp_cost_grad_2d(const float *mov, const float *fix, int nx, int ny, int cx, int cy, const double *params,
                        double *grad)
{
    const size_t N = (size_t)nx * ny;
    Axis ax = make_axis(nx, cx), ay = make_axis(ny, cy);
    // displacement per pixel, 4x4 sums of the float coefficients
    std::vector<float> px(N), py(N);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            float ux = 0.0f, uy = 0.0f;
            for (size_t j = 0; j < ay.w[y].size(); ++j) {
                const int ky = ay.start[y] + (int)j;
                if (ky >= cy)
                    break;
                float hx = 0.0f, hy = 0.0f;
                for (size_t i = 0; i < ax.w[x].size(); ++i) {
                    const int kx = ax.start[x] + (int)i;
                    if (kx >= cx)
                        break;
                    const size_t k = (size_t)ky * cx + kx;
                    hx += (float)ax.w[x][i] * (float)params[2 * k];
                    hy += (float)ax.w[x][i] * (float)params[2 * k + 1];
                }
                ux += (float)ay.w[y][j] * hx;
                uy += (float)ay.w[y][j] * hy;
            }
            px[(size_t)y * nx + x] = (float)x - ux;
            py[(size_t)y * nx + x] = (float)y - uy;
        }
    // interpolator: copy + prefilter along x and y
    std::vector<float> cf(mov, mov + N);
    float vmin = mov[0], vmax = mov[0];
    for (size_t i = 0; i < N; ++i) {
        vmin = std::min(vmin, mov[i]);
        vmax = std::max(vmax, mov[i]);
    }
    if (0.0f < vmin)
        vmin = 0.0f;
    {
        std::vector<double> buf(nx);
        for (int y = 0; y < ny; ++y) {
            for (int x = 0; x < nx; ++x)
                buf[x] = cf[(size_t)y * nx + x];
            filter_line(buf);
            for (int x = 0; x < nx; ++x)
                cf[(size_t)y * nx + x] = (float)buf[x];
        }
        buf.resize(ny);
        for (int x = 0; x < nx; ++x) {
            for (int y = 0; y < ny; ++y)
                buf[y] = cf[(size_t)y * nx + x];
            filter_line(buf);
            for (int y = 0; y < ny; ++y)
                cf[(size_t)y * nx + x] = (float)buf[y];
        }
    }
    std::vector<float> W(N);
    for (size_t idx = 0; idx < N; ++idx) {
        const float pc[2] = {px[idx], py[idx]};
        int s2[2];
        double wd[2][4];
        for (int a = 0; a < 2; ++a) {
            const int f = fastfloor(pc[a]);
            bspline3((double)pc[a] - f, wd[a]);
            s2[a] = f - 1;
        }
        float r = 0.0f;
        for (int b = 0; b < 4; ++b) {
            const float *row = &cf[(size_t)mirror(s2[1] + b, ny) * nx];
            float rx = 0.0f;
            for (int a = 0; a < 4; ++a)
                rx += (float)wd[0][a] * row[mirror(s2[0] + a, nx)];
            r += (float)wd[1][b] * rx;
        }
        W[idx] = r >= vmin ? (r <= vmax ? r : vmax) : vmin;
    }
    // gradient, force, float cost
    std::vector<float> fxv(N), fyv(N);
    float fcost = 0.0f;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t i = (size_t)y * nx + x;
            const float gx = (x > 0 && x < nx - 1) ? 0.5f * (W[i + 1] - W[i - 1]) : 0.0f;
            const float gy = (y > 0 && y < ny - 1) ? 0.5f * (W[i + nx] - W[i - nx]) : 0.0f;
            const float d = W[i] - fix[i];
            fxv[i] = gx * d;
            fyv[i] = gy * d;
            fcost += d * d;
        }
    // translate: y rows, then x
    std::vector<float> tx((size_t)cy * nx * 2, 0.0f);
    for (int oy = 0; oy < cy; ++oy) {
        int j = ay.row_first[oy];
        for (double w : ay.row[oy]) {
            for (int x = 0; x < nx; ++x) {
                tx[2 * ((size_t)oy * nx + x)] += (float)w * fxv[(size_t)j * nx + x];
                tx[2 * ((size_t)oy * nx + x) + 1] += (float)w * fyv[(size_t)j * nx + x];
            }
            ++j;
        }
    }
    for (int oy = 0; oy < cy; ++oy)
        for (int ox = 0; ox < cx; ++ox) {
            float ax2 = 0.0f, ay2 = 0.0f;
            int j = ax.row_first[ox];
            for (double w : ax.row[ox]) {
                ax2 += (float)w * tx[2 * ((size_t)oy * nx + j)];
                ay2 += (float)w * tx[2 * ((size_t)oy * nx + j) + 1];
                ++j;
            }
            grad[2 * ((size_t)oy * cx + ox)] = -ax2;
            grad[2 * ((size_t)oy * cx + ox) + 1] = -ay2;
        }
    return 0.5 * (double)fcost;
}

} // extern "C"

// oracle restatements used by the multiresolution baseline (same directory,
// test infrastructure): the `scale` filter (mia/3d/filter/scale.cc) and the
// divcurl matrix cells (mia/3d/ppmatrix.cc:178-262)
extern "C" void mo_scale_image(const float *img, const int in_size[3], const int out_size[3], float *out);
extern "C" void mo_pp_cells(const int csize[3], const float range[3], double wdiv, double wrot, double *cells);

namespace {

// C3DPPDivcurlMatrixImpl::evaluate (mia/3d/ppmatrix.cc:300-597) with its pfor
// over z slabs: value sum_i c_i P c_j over j in [i-4, i+3] per axis, and its
// derivative; per-slice sums added in slice order.
double divcurl_eval(const int cs[3], const std::vector<double> &cells, const std::vector<float> &coef,
                    std::vector<double> &grad, int nthreads)
{
    const int nx = cs[0], ny = cs[1], nz = cs[2];
    std::vector<double> part(nz, 0.0);
    grad.assign(coef.size(), 0.0);
    pfor(nz, nthreads, [&](int zb, int ze) {
        for (int zi = zb; zi < ze; ++zi) {
            double r[6] = {0, 0, 0, 0, 0, 0};
            for (int yi = 0; yi < ny; ++yi)
                for (int xi = 0; xi < nx; ++xi) {
                    const size_t i = ((size_t)zi * ny + yi) * nx + xi;
                    const double cix = coef[3 * i], ciy = coef[3 * i + 1], ciz = coef[3 * i + 2];
                    double g[3] = {0, 0, 0};
                    for (int zj = std::max(zi - 4, 0); zj < std::min(zi + 4, nz); ++zj)
                        for (int yj = std::max(yi - 4, 0); yj < std::min(yi + 4, ny); ++yj)
                            for (int xj = std::max(xi - 4, 0); xj < std::min(xi + 4, nx); ++xj) {
                                const size_t j = ((size_t)zj * ny + yj) * nx + xj;
                                const double *pc =
                                    &cells[6 * (((size_t)(zi - zj + 4) * 9 + (yi - yj + 4)) * 9 + (xi - xj + 4))];
                                const double cx = coef[3 * j], cy = coef[3 * j + 1], cz = coef[3 * j + 2];
                                r[0] += cix * cx * pc[0];
                                r[1] += cix * cy * pc[2];
                                r[2] += ciy * cy * pc[3];
                                r[3] += ciz * cx * pc[1];
                                r[4] += ciy * cz * pc[4];
                                r[5] += ciz * cz * pc[5];
                                g[0] += 2 * cx * pc[0] + cy * pc[2] + cz * pc[1];
                                g[1] += 2 * cy * pc[3] + cz * pc[4] + cx * pc[2];
                                g[2] += 2 * cz * pc[5] + cx * pc[1] + cy * pc[4];
                            }
                    grad[3 * i] = g[0];
                    grad[3 * i + 1] = g[1];
                    grad[3 * i + 2] = g[2];
                }
            part[zi] = r[0] + r[1] + r[2] + r[3] + r[4] + r[5];
        }
    });
    double total = 0.0;
    for (double v : part)
        total += v;
    return total;
}

// C3DSplineTransformation::refine (mia/3d/transform/spline.cc:360-397): the
// coefficient field, prefiltered per component, sampled at the new grid's
// knots with cubic weights and mirrored indices, clamped per component to the
// field's range (the lower bound 0 when every component is positive).
std::vector<float> refine_field(const std::vector<float> &coef, const int cs[3], const int ncs[3])
{
    std::vector<double> c(coef.begin(), coef.end());
    double vmin[3], vmax[3];
    for (int k = 0; k < 3; ++k) {
        vmin[k] = 1e300;
        vmax[k] = -1e300;
    }
    const size_t n = (size_t)cs[0] * cs[1] * cs[2];
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            vmin[k] = std::min(vmin[k], c[3 * i + k]);
            vmax[k] = std::max(vmax[k], c[3 * i + k]);
        }
    if (vmin[0] > 0.0 && vmin[1] > 0.0 && vmin[2] > 0.0)
        vmin[0] = vmin[1] = vmin[2] = 0.0;
    for (int axis = 0; axis < 3; ++axis) {
        std::vector<double> line(cs[axis]);
        const size_t stride = axis == 0 ? 1 : axis == 1 ? (size_t)cs[0] : (size_t)cs[0] * cs[1];
        const int o1 = axis == 0 ? cs[1] : cs[0], o2 = axis == 2 ? cs[1] : cs[2];
        for (int b = 0; b < o2; ++b)
            for (int a = 0; a < o1; ++a) {
                const size_t base = axis == 0   ? ((size_t)b * cs[1] + a) * cs[0]
                                    : axis == 1 ? (size_t)b * cs[0] * cs[1] + a
                                                : (size_t)b * cs[0] + a;
                for (int comp = 0; comp < 3; ++comp) {
                    for (int k = 0; k < cs[axis]; ++k)
                        line[k] = c[(base + k * stride) * 3 + comp];
                    filter_line(line);
                    for (int k = 0; k < cs[axis]; ++k)
                        c[(base + k * stride) * 3 + comp] = line[k];
                }
            }
    }
    float dx[3];
    for (int a = 0; a < 3; ++a)
        dx[a] = (float)(cs[a] - 3) / (float)(ncs[a] - 3);
    std::vector<float> out((size_t)ncs[0] * ncs[1] * ncs[2] * 3);
    size_t o = 0;
    for (int z = 0; z < ncs[2]; ++z)
        for (int y = 0; y < ncs[1]; ++y)
            for (int x = 0; x < ncs[0]; ++x, ++o) {
                const int X[3] = {x, y, z};
                double w[3][4];
                int idx[3][4];
                for (int a = 0; a < 3; ++a) {
                    const float p = dx[a] * ((float)X[a] - 1.0f) + 1.0f;
                    const int fl = fastfloor(p);
                    bspline3((double)p - fl, w[a]);
                    for (int k = 0; k < 4; ++k)
                        idx[a][k] = mirror(fl - 1 + k, cs[a]);
                }
                for (int comp = 0; comp < 3; ++comp) {
                    double v = 0.0;
                    for (int cz = 0; cz < 4; ++cz) {
                        double vy = 0.0;
                        for (int b = 0; b < 4; ++b) {
                            double vx = 0.0;
                            for (int a = 0; a < 4; ++a)
                                vx += w[0][a] *
                                      c[(((size_t)idx[2][cz] * cs[1] + idx[1][b]) * cs[0] + idx[0][a]) * 3 + comp];
                            vy += w[1][b] * vx;
                        }
                        v += w[2][cz] * vy;
                    }
                    out[o * 3 + comp] = (float)std::min(std::max(v, vmin[comp]), vmax[comp]);
                }
            }
    return out;
}

} // namespace

extern "C" {

// The 3D multiresolution registration on the host, as TNonrigidRegisterImpl::
// run drives it (mia/template/nonrigidregister.cxx:192-296): src and ref
// mapped jointly to mean 0 / sigma 1 (:147-184); for each of `levels` levels
// (image size / 2^k) the images resampled with the `scale` filter, the
// transformation created (`spline:rate=rate`, zero coefficients) or upscaled
// (coefficients times size / range, spline.cc:399-412), the cost
// image:cost=[ssd] plus the divcurl penalty (weight `penalty_weight`, div =
// curl = 1, value weight * c^T P c, gradient -weight * its derivative,
// mia/3d/splinetransformpenalty.cc:52-71) minimised by L-BFGS (maxeval
// evaluations, memory 10, first step 1) from the level's start, then the grid
// refined to the rate (spline.cc:360-397) and minimised again if it grew.
// Outputs (may be NULL): per level seconds, evaluations and best cost, the final cost
// (costs + penalty at the last level's best point) and coefficient grid.
// Returns the total seconds.
double mcp_register_3d(const float *mov, const float *fix, int nx, int ny, int nz, float rate, int levels,
                       int maxeval, float penalty_weight, int nthreads, double *level_seconds, int *level_evals,
                       double *level_costs, double *final_cost, int *csize_out)
{
    using clk = std::chrono::steady_clock;
    const auto t_all = clk::now();
    const size_t N = (size_t)nx * ny * nz;
    std::vector<float> a(mov, mov + N), b(fix, fix + N);
    {
        double sum = 0.0, sum2 = 0.0;
        for (size_t k = 0; k < N; ++k) {
            sum += (double)a[k] + b[k];
            sum2 += (double)a[k] * a[k] + (double)b[k] * b[k];
        }
        const double m = 2.0 * N, mean = sum / m, sigma = std::sqrt((sum2 - sum * sum / m) / (m - 1));
        if (sigma > 0.0)
            for (size_t k = 0; k < N; ++k) {
                a[k] = (float)(a[k] / sigma - mean / sigma);
                b[k] = (float)(b[k] / sigma - mean / sigma);
            }
    }
    const int full[3] = {nx, ny, nz};
    const double w = (double)penalty_weight;
    std::vector<float> coef;
    int cs[3] = {0, 0, 0}, range[3] = {0, 0, 0};
    double best = 0.0;
    for (int l = 0; l < levels; ++l) {
        const auto t0 = clk::now();
        const int shift = levels - 1 - l;
        int sz[3];
        for (int k = 0; k < 3; ++k)
            sz[k] = full[k] >> shift;
        const size_t n = (size_t)sz[0] * sz[1] * sz[2];
        std::vector<float> la(n), lb(n);
        if (shift == 0) {
            la = a;
            lb = b;
        } else {
            mo_scale_image(a.data(), full, sz, la.data());
            mo_scale_image(b.data(), full, sz, lb.data());
        }
        if (l == 0) {
            for (int k = 0; k < 3; ++k)
                cs[k] = (int)(((float)(sz[k] - 1) + rate) / rate) + 2;
            coef.assign((size_t)cs[0] * cs[1] * cs[2] * 3, 0.0f);
        } else {
            float mx[3];
            for (int k = 0; k < 3; ++k)
                mx[k] = (float)sz[k] / (float)range[k];
            for (size_t k = 0; k < coef.size(); ++k)
                coef[k] = mx[k % 3] * coef[k];
        }
        for (int k = 0; k < 3; ++k)
            range[k] = sz[k];
        int evals = 0;
        auto minimize = [&]() {
            const float rng[3] = {(float)range[0], (float)range[1], (float)range[2]};
            std::vector<double> cells(9 * 9 * 9 * 6);
            mo_pp_cells(cs, rng, 1.0, 1.0, cells.data());
            const size_t np = coef.size();
            mia_amd::LBFGSMinimizer opt(np, maxeval, 10, 0.0, 0.0, 1.0);
            std::vector<double> x0(coef.begin(), coef.end()), g(np), gp;
            std::vector<float> cf(np);
            opt.start(x0.data());
            while (!opt.done()) {
                const double *x = opt.point();
                double f = mcp_cost_grad_3d(la.data(), lb.data(), sz[0], sz[1], sz[2], cs[0], cs[1], cs[2], x,
                                            nthreads, g.data(), nullptr);
                for (size_t k = 0; k < np; ++k)
                    cf[k] = (float)x[k];
                f += w * divcurl_eval(cs, cells, cf, gp, nthreads);
                for (size_t k = 0; k < np; ++k)
                    g[k] += -w * gp[k];
                opt.tell(f, g.data());
            }
            evals += opt.evaluations();
            best = opt.best_f();
            if (opt.evaluations() > 0 && std::isfinite(opt.best_f()))
                for (size_t k = 0; k < np; ++k)
                    coef[k] = (float)opt.best_x()[k];
        };
        minimize();
        int ncs[3];
        bool grow = false;
        for (int k = 0; k < 3; ++k) {
            ncs[k] = (int)(((float)range[k] + rate - 1.0f) / rate) + 2;
            grow |= ncs[k] > cs[k];
        }
        if (grow) {
            coef = refine_field(coef, cs, ncs);
            for (int k = 0; k < 3; ++k)
                cs[k] = ncs[k];
            minimize();
        }
        if (level_seconds)
            level_seconds[l] = std::chrono::duration<double>(clk::now() - t0).count();
        if (level_evals)
            level_evals[l] = evals;
        if (level_costs)
            level_costs[l] = best;
    }
    if (final_cost)
        *final_cost = best;
    if (csize_out)
        for (int k = 0; k < 3; ++k)
            csize_out[k] = cs[k];
    return std::chrono::duration<double>(clk::now() - t_all).count();
}

// The series baseline: frames (sizes 2 per frame, images concatenated) each
// normalised jointly to mean 0 / sigma 1 (nonrigidregister.cxx:147-184) and
// registered from the identity with L-BFGS (maxeval evaluations, memory 10),
// frames taken by nthreads workers one at a time.  Returns the wall time in
// seconds; costs_out / evals_out per frame (may be NULL).
double mcp_series_2d(int n, const int *sizes, const float *movs, const float *fixs, float rate, int maxeval,
                     int nthreads, double *costs_out, int *evals_out)
{
    std::vector<size_t> off(n + 1, 0);
    for (int i = 0; i < n; ++i)
        off[i + 1] = off[i] + (size_t)sizes[2 * i] * sizes[2 * i + 1];
    auto t0 = std::chrono::steady_clock::now();
    std::atomic<int> next(0);
    auto work = [&]() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n)
                return;
            const int nx = sizes[2 * i], ny = sizes[2 * i + 1];
            const size_t N = (size_t)nx * ny;
            std::vector<float> a(movs + off[i], movs + off[i] + N), b(fixs + off[i], fixs + off[i] + N);
            double sum = 0.0, sum2 = 0.0;
            for (size_t k = 0; k < N; ++k) {
                sum += (double)a[k] + b[k];
                sum2 += (double)a[k] * a[k] + (double)b[k] * b[k];
            }
            const double m = 2.0 * N, mean = sum / m, sigma = std::sqrt((sum2 - sum * sum / m) / (m - 1));
            if (sigma > 0.0)
                for (size_t k = 0; k < N; ++k) {
                    a[k] = (float)(a[k] / sigma - mean / sigma);
                    b[k] = (float)(b[k] / sigma - mean / sigma);
                }
            const int cx = (int)(((float)nx + rate - 1.0f) / rate) + 2;
            const int cy = (int)(((float)ny + rate - 1.0f) / rate) + 2;
            const size_t np = 2 * (size_t)cx * cy;
            mia_amd::LBFGSMinimizer opt(np, maxeval, 10, 0.0, 0.0, 1.0);
            std::vector<double> x0(np, 0.0), g(np);
            opt.start(x0.data());
            while (!opt.done()) {
                const double f = mcp_cost_grad_2d(a.data(), b.data(), nx, ny, cx, cy, opt.point(), g.data());
                opt.tell(f, g.data());
            }
            if (costs_out)
                costs_out[i] = opt.best_f();
            if (evals_out)
                evals_out[i] = opt.evaluations();
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < std::max(1, nthreads); ++t)
        th.emplace_back(work);
    for (auto &t : th)
        t.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // extern "C"

/*
 * mia_oracle.c -- FP64 CPU restatement of MIA's B-spline registration
 * cost + gradient path (spline transformation, ssd cost, image full cost).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is compiled into
 * oracle/lib/libmia_oracle.so and loaded only by tests/, by the smoke check
 * and by the bench's CPU-baseline leg.  The product library (mia_amd/csrc)
 * never links, loads or calls it, and there is no CPU fallback through it.
 *
 * Every function cites the reference file:line whose algorithm it restates.
 * All arithmetic is double precision.  Where the reference fixes a quantity in
 * single precision as part of its *definition* (the float grid scale of the
 * spline transformation and the float grid positions derived from it), the
 * oracle reproduces that float value and continues in double.
 *
 * Pinning (what checks each function independently of this restatement) is
 * listed in tests/test_oracle.py next to each test; nothing here is pinned by
 * re-typing its own formula.
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MO_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* a1: kernel weights and the 16.16 fixed point floor                        */
/* ------------------------------------------------------------------------ */

/* fastfloor: floor after rounding to a multiple of 2^-16 (round half even),
 * mia/core/splinekernel.cc:164-192 (the 1.5*2^36 magic-number trick). */
MO_EXPORT int mo_fastfloor(double v)
{
    double r = nearbyint(v * 65536.0);
    return (int)floor(r / 65536.0);
}

/* Cubic B-spline weights at fractional offset f,
 * mia/core/splinekernel/bspline.cc:258-276 (CBSplineKernel3::get_weights). */
MO_EXPORT void mo_bspline3_weights(double x, double w[4])
{
    const double xm1 = 1.0 - x;
    const double onebysix = 1.0 / 6.0;
    w[3] = onebysix * x * x * x;
    w[0] = onebysix * xm1 * xm1 * xm1;
    w[2] = x - 2.0 * w[3] + w[0];
    w[1] = xm1 - 2.0 * w[0] + w[3];
}

/* First derivative weights, mia/core/splinekernel/bspline.cc:279-285. */
MO_EXPORT void mo_bspline3_dweights(double x, double w[4])
{
    w[3] = 0.5 * x * x;
    w[0] = x - 0.5 - w[3];
    w[2] = 1.0 + w[0] - 2.0 * w[3];
    w[1] = -w[0] - w[2] - w[3];
}

/* Kernel value beta_3(x), mia/core/splinekernel/bspline.cc:328-345. */
MO_EXPORT double mo_bspline3_value(double x)
{
    x = fabs(x);
    if (x <= 1.0)
        return 2.0 / 3.0 - x * x * (1.0 - 0.5 * x);
    if (x < 2.0) {
        const double dx = 2.0 - x;
        return dx * dx * dx / 6.0;
    }
    return 0.0;
}

/* Start index + weights of the image interpolation kernel for degree 1 or 3,
 * mia/core/splinekernel.cc:250-255 (get_start_idx_and_value_weights) with
 * CBSplineKernel1 (bspline.cc:90-107, shift 0, half degree 0) and
 * CBSplineKernel3 (bspline.cc:240-276, shift 0, half degree 1). */
static int kernel_weights(int degree, double x, double w[4], double dw[4])
{
    const int ix = mo_fastfloor(x);
    const double f = x - ix;
    if (degree == 1) {
        w[0] = 1.0 - f;
        w[1] = f;
        if (dw) {
            dw[0] = -1.0;
            dw[1] = 1.0;
        }
        return ix;
    }
    mo_bspline3_weights(f, w);
    if (dw)
        mo_bspline3_dweights(f, dw);
    return ix - 1;
}

/* Mirror boundary index map, mia/core/splinebc/bc.cc:52-64. */
MO_EXPORT int mo_mirror_index(int i, int width)
{
    const int width2 = 2 * width - 2;
    int idx = i < 0 ? -i : i;
    if (width == 1)
        return 0;
    idx = idx < width2 ? idx : idx % width2;
    if (width <= idx)
        idx = width2 - idx;
    return idx;
}

/* ------------------------------------------------------------------------ */
/* a2: recursive prefilter with mirror boundaries                             */
/* ------------------------------------------------------------------------ */

/* Causal initial value for mirror boundaries, mia/core/splinebc/bc.cc:67-83. */
static double mirror_initial_coeff(const double *c, int n, double pole)
{
    double zn = pole;
    const double iz = 1.0 / pole;
    double z2n = pow(pole, (double)(n - 1));
    double sum = c[0] + z2n * c[n - 1];
    z2n *= z2n * iz;
    for (int k = 1; k < n - 1; ++k) {
        sum += (zn + z2n) * c[k];
        zn *= pole;
        z2n *= iz;
    }
    return sum / (1.0 - zn * zn);
}

/* Anticausal initial value, mia/core/splinebc/bc.cc:85-88. */
static double mirror_initial_anti_coeff(const double *c, int n, double pole)
{
    return (pole / (pole * pole - 1.0)) * (pole * c[n - 2] + c[n - 1]);
}

/* One line of the B-spline prefilter: gain, causal and anticausal passes,
 * mia/core/boundary_conditions.cc:62-103.  degree 1 has no poles (no-op). */
MO_EXPORT void mo_filter_line(double *c, int n, int degree)
{
    if (n < 2 || degree != 3)
        return;
    const double p = sqrt(3.0) - 2.0; /* bspline.cc:248 */
    const double lambda = 2.0 - p - 1.0 / p;
    for (int k = 0; k < n; ++k)
        c[k] *= lambda;
    c[0] = mirror_initial_coeff(c, n, p);
    for (int k = 1; k < n; ++k)
        c[k] += p * c[k - 1];
    c[n - 1] = mirror_initial_anti_coeff(c, n, p);
    for (int k = n - 2; k >= 0; --k)
        c[k] = p * (c[k + 1] - c[k]);
}

/* Separable prefilter of an image (x, then y, then z lines), as the
 * interpolator constructors do: mia/3d/interpolator.cxx:98-164,
 * mia/2d/interpolator.cxx:91-132.  Also returns the clamp range
 * [min(0, min), max] (3d/interpolator.cxx:108-111, 2d/interpolator.cxx:99-103).
 * nz == 1 gives the 2D case. */
MO_EXPORT void mo_prefilter(const float *img, int nx, int ny, int nz, int degree,
                            double *coeff, double *vmin, double *vmax)
{
    const size_t n = (size_t)nx * ny * nz;
    double mn = img[0], mx = img[0];
    for (size_t i = 0; i < n; ++i) {
        coeff[i] = img[i];
        if (img[i] < mn) mn = img[i];
        if (img[i] > mx) mx = img[i];
    }
    if (0.0 < mn)
        mn = 0.0;
    *vmin = mn;
    *vmax = mx;
    if (degree != 3)
        return;
    int maxn = nx > ny ? nx : ny;
    if (nz > maxn) maxn = nz;
    double *line = (double *)malloc(sizeof(double) * maxn);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            double *row = coeff + ((size_t)z * ny + y) * nx;
            mo_filter_line(row, nx, degree);
        }
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x) {
            for (int y = 0; y < ny; ++y) line[y] = coeff[((size_t)z * ny + y) * nx + x];
            mo_filter_line(line, ny, degree);
            for (int y = 0; y < ny; ++y) coeff[((size_t)z * ny + y) * nx + x] = line[y];
        }
    if (nz > 1)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                for (int z = 0; z < nz; ++z) line[z] = coeff[((size_t)z * ny + y) * nx + x];
                mo_filter_line(line, nz, degree);
                for (int z = 0; z < nz; ++z) coeff[((size_t)z * ny + y) * nx + x] = line[z];
            }
    free(line);
}

/* ------------------------------------------------------------------------ */
/* a3/a5: interpolation of the prefiltered image with clamp                   */
/* ------------------------------------------------------------------------ */

/* Value (and, if grad != NULL, the analytic gradient) of the interpolated
 * image at position p, mia/3d/interpolator.cxx:250-287 (z->y->x contraction,
 * add_3d :187-210), mirror indices via the weight cache
 * (mia/core/splinekernel.cc:145-162, mia/3d/interpolator.cc:43-51), clamp to
 * [vmin, vmax] (:177-185, :284).  The gradient of a clamped value is zero.
 * nz == 1 with p[2] == 0 is the 2D interpolator (2d/interpolator.cxx:277-311):
 * a cubic z kernel at z == 0 on a 1-wide mirror axis has weights summing to 1
 * on index 0, so the caller uses mo_interp2 for exact 2D semantics instead. */
MO_EXPORT double mo_interp3(const double *c, int nx, int ny, int nz, int degree,
                            const double p[3], double vmin, double vmax, double grad[3])
{
    double wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
    const int ns = degree == 1 ? 2 : 4;
    const int sx = kernel_weights(degree, p[0], wx, dx);
    const int sy = kernel_weights(degree, p[1], wy, dy);
    const int sz = kernel_weights(degree, p[2], wz, dz);
    int ix[4], iy[4], iz[4];
    for (int k = 0; k < ns; ++k) {
        ix[k] = mo_mirror_index(sx + k, nx);
        iy[k] = mo_mirror_index(sy + k, ny);
        iz[k] = mo_mirror_index(sz + k, nz);
    }
    double v = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (int a = 0; a < ns; ++a) {
        double vy = 0.0, gyx = 0.0, gyy = 0.0;
        for (int b = 0; b < ns; ++b) {
            const double *row = c + ((size_t)iz[a] * ny + iy[b]) * nx;
            double vx = 0.0, gxx = 0.0;
            for (int k = 0; k < ns; ++k) {
                vx += wx[k] * row[ix[k]];
                gxx += dx[k] * row[ix[k]];
            }
            vy += wy[b] * vx;
            gyx += wy[b] * gxx;
            gyy += dy[b] * vx;
        }
        v += wz[a] * vy;
        gx += wz[a] * gyx;
        gy += wz[a] * gyy;
        gz += dz[a] * vy;
    }
    if (v < vmin || v > vmax) {
        gx = gy = gz = 0.0;
        v = v < vmin ? vmin : vmax;
    }
    if (grad) {
        grad[0] = gx;
        grad[1] = gy;
        grad[2] = gz;
    }
    return v;
}

/* 2D interpolation, mia/2d/interpolator.cxx:277-311 (y->x contraction),
 * clamp :153-160,308. */
MO_EXPORT double mo_interp2(const double *c, int nx, int ny, int degree,
                            const double p[2], double vmin, double vmax, double grad[2])
{
    double wx[4], wy[4], dx[4], dy[4];
    const int ns = degree == 1 ? 2 : 4;
    const int sx = kernel_weights(degree, p[0], wx, dx);
    const int sy = kernel_weights(degree, p[1], wy, dy);
    int ix[4], iy[4];
    for (int k = 0; k < ns; ++k) {
        ix[k] = mo_mirror_index(sx + k, nx);
        iy[k] = mo_mirror_index(sy + k, ny);
    }
    double v = 0.0, gx = 0.0, gy = 0.0;
    for (int b = 0; b < ns; ++b) {
        const double *row = c + (size_t)iy[b] * nx;
        double vx = 0.0, gxx = 0.0;
        for (int k = 0; k < ns; ++k) {
            vx += wx[k] * row[ix[k]];
            gxx += dx[k] * row[ix[k]];
        }
        v += wy[b] * vx;
        gx += wy[b] * gxx;
        gy += dy[b] * vx;
    }
    if (v < vmin || v > vmax) {
        gx = gy = 0.0;
        v = v < vmin ? vmin : vmax;
    }
    if (grad) {
        grad[0] = gx;
        grad[1] = gy;
    }
    return v;
}

/* ------------------------------------------------------------------------ */
/* a4/a10: spline transformation geometry                                      */
/* ------------------------------------------------------------------------ */

/* Coefficient grid size per axis for the cubic transformation kernel
 * (shift 1, enlarge 2).  3D: mia/3d/transform/spline.cc:130-143 computes
 * (float(n-1) + rate) / rate in float and truncates; 2D:
 * mia/2d/transform/spline.cc:103-116 computes (n + rate - 1) / rate.
 * Returns -1 where the reference throws invalid_argument (grid <= 1). */
MO_EXPORT int mo_spline_csize(int n, float rate, int dim)
{
    float v;
    if (dim == 3)
        v = ((float)(n - 1) + rate) / rate;
    else
        v = (((float)n + rate) - 1.0f) / rate;
    const int cs = (int)v;
    if (cs <= 1)
        return -1;
    return cs + 2;
}

/* Float grid scale, spline.cc:237-238 (3D) and 2D spline.cc:181-182:
 * (csize - 1 - enlarge) / (n - 1) evaluated in float. */
MO_EXPORT float mo_spline_scale(int csize, int n)
{
    return (float)(csize - 3) / (float)(n - 1);
}

/* Per-axis forward weight table, spline.cc:249-273 (3D) / 2D :187-200:
 * position i*scale + 1 evaluated in float, start = fastfloor(pos) - 1,
 * weights at pos - fastfloor(pos); the last sample keeps only 3 weights
 * (the fourth is dropped, it multiplies a coefficient beyond the grid).
 * w is n x 4, nw receives the number of weights per sample. */
MO_EXPORT void mo_spline_axis_table(int n, int csize, int *start, double *w, int *nw)
{
    const float scale = mo_spline_scale(csize, n);
    for (int i = 0; i < n; ++i) {
        const float pos = (float)i * scale + 1.0f;
        const int f = mo_fastfloor((double)pos);
        start[i] = f - 1;
        mo_bspline3_weights((double)pos - f, w + 4 * i);
        nw[i] = (i == n - 1) ? 3 : 4;
    }
}

/* Adjoint row of one axis, spline.cc:697-729 (3D) / 2D :509-538: for output
 * coefficient o, the weights beta_3(i*scale + 1 - o) (scale promoted from its
 * float value) over the contiguous input range where |.| < 2.  Written as a
 * dense n x csize matrix M[o*n + i] (zero outside the row). */
MO_EXPORT void mo_spline_derivative_rows(int n, int csize, double *M)
{
    const double scale = (double)mo_spline_scale(csize, n);
    memset(M, 0, sizeof(double) * (size_t)n * csize);
    for (int o = 0; o < csize; ++o) {
        for (int i = 0; i < n; ++i) {
            const double x = i * scale + 1.0 - o;
            if (fabs(x) < 2.0) {
                const double y = mo_bspline3_value(x);
                if (y > 0)
                    M[(size_t)o * n + i] = y;
                else
                    break;
            }
        }
    }
}

typedef struct {
    int n, cs;
    int *start;
    int *nw;
    double *w;
} axis_t;

static void axis_init(axis_t *a, int n, int cs)
{
    a->n = n;
    a->cs = cs;
    a->start = (int *)malloc(sizeof(int) * n);
    a->nw = (int *)malloc(sizeof(int) * n);
    a->w = (double *)malloc(sizeof(double) * 4 * n);
    mo_spline_axis_table(n, cs, a->start, a->w, a->nw);
}

static void axis_free(axis_t *a)
{
    free(a->start);
    free(a->nw);
    free(a->w);
}

/* Dense displacement field u (interleaved, dim components per voxel) of the
 * coefficient grid `params` (interleaved per coefficient, x fastest), i.e.
 * the tensor-product sum of spline.cc:433-466 / init_grid :541-653 and 2D
 * on_grid :640-666; coefficient indices beyond the grid are skipped.  The
 * transformed position is T(x) = x - u(x) (spline.cc:310-314, 631-651). */
MO_EXPORT void mo_displacement(int dim, const int size[3], const int csize[3],
                               const double *params, double *u)
{
    axis_t ax, ay, az;
    const int nz = dim == 3 ? size[2] : 1;
    axis_init(&ax, size[0], csize[0]);
    axis_init(&ay, size[1], csize[1]);
    if (dim == 3)
        axis_init(&az, size[2], csize[2]);
    const int csz = dim == 3 ? csize[2] : 1;
    size_t idx = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < size[1]; ++y)
            for (int x = 0; x < size[0]; ++x, ++idx) {
                double acc[3] = {0, 0, 0};
                const int nwz = dim == 3 ? az.nw[z] : 1;
                for (int c = 0; c < nwz; ++c) {
                    const int kz = dim == 3 ? az.start[z] + c : 0;
                    const double wz = dim == 3 ? az.w[4 * z + c] : 1.0;
                    if (kz >= csz) break;
                    for (int b = 0; b < ay.nw[y]; ++b) {
                        const int ky = ay.start[y] + b;
                        if (ky >= csize[1]) break;
                        const double wy = ay.w[4 * y + b];
                        for (int a = 0; a < ax.nw[x]; ++a) {
                            const int kx = ax.start[x] + a;
                            if (kx >= csize[0]) break;
                            const double wgt = wz * wy * ax.w[4 * x + a];
                            const double *cc = params + dim * (((size_t)kz * csize[1] + ky) * csize[0] + kx);
                            for (int d = 0; d < dim; ++d)
                                acc[d] += wgt * cc[d];
                        }
                    }
                }
                for (int d = 0; d < dim; ++d)
                    u[dim * idx + d] = acc[d];
            }
    axis_free(&ax);
    axis_free(&ay);
    if (dim == 3)
        axis_free(&az);
}

/* Displacement at an arbitrary position x (get_displacement_at,
 * spline.cc:414-418 -> scale() :448-453 in float -> interpolate :420-431 ->
 * sum :433-466; 2D spline.cc:205-215,675-680). */
MO_EXPORT void mo_displacement_at(int dim, const int size[3], const int csize[3], const double *params,
                                  const double *pos, double *out)
{
    int start[3] = {0, 0, 0};
    double w[3][4] = {{1, 0, 0, 0}, {1, 0, 0, 0}, {1, 0, 0, 0}};
    for (int a = 0; a < dim; ++a) {
        const float s = mo_spline_scale(csize[a], size[a]);
        const float p = (float)pos[a] * s + 1.0f;
        const int f = mo_fastfloor((double)p);
        start[a] = f - 1;
        mo_bspline3_weights((double)p - f, w[a]);
    }
    const int nzw = dim == 3 ? 4 : 1;
    const int csz = dim == 3 ? csize[2] : 1;
    double acc[3] = {0, 0, 0};
    for (int c = 0; c < nzw; ++c) {
        const int kz = start[2] + c;
        if (kz < 0 || kz >= csz) continue;
        for (int b = 0; b < 4; ++b) {
            const int ky = start[1] + b;
            if (ky < 0 || ky >= csize[1]) continue;
            for (int a = 0; a < 4; ++a) {
                const int kx = start[0] + a;
                if (kx < 0 || kx >= csize[0]) continue;
                const double wgt = w[2][c] * w[1][b] * w[0][a];
                const double *cc = params + dim * (((size_t)kz * csize[1] + ky) * csize[0] + kx);
                for (int d = 0; d < dim; ++d)
                    acc[d] += wgt * cc[d];
            }
        }
    }
    for (int d = 0; d < dim; ++d)
        out[d] = acc[d];
}

/* ------------------------------------------------------------------------ */
/* a5: warp                                                                    */
/* ------------------------------------------------------------------------ */

/* W(x) = M~(x - u(x)) for every voxel, mia/3d/transform.cc:280-353,
 * mia/2d/transform.cc:258-294.  coeff is the prefiltered moving image.  When
 * mgrad != NULL it also receives the analytic gradient of the interpolated
 * (clamped) moving image at x - u(x) (used by the exact-gradient mode). */
MO_EXPORT void mo_warp(int dim, const int size[3], const double *coeff, int degree,
                       double vmin, double vmax, const double *u, double *W, double *mgrad)
{
    const int nz = dim == 3 ? size[2] : 1;
    size_t idx = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < size[1]; ++y)
            for (int x = 0; x < size[0]; ++x, ++idx) {
                if (dim == 3) {
                    double p[3] = {x - u[3 * idx], y - u[3 * idx + 1], z - u[3 * idx + 2]};
                    double g[3];
                    W[idx] = mo_interp3(coeff, size[0], size[1], size[2], degree, p, vmin, vmax, g);
                    if (mgrad) {
                        mgrad[3 * idx] = g[0];
                        mgrad[3 * idx + 1] = g[1];
                        mgrad[3 * idx + 2] = g[2];
                    }
                } else {
                    double p[2] = {x - u[2 * idx], y - u[2 * idx + 1]};
                    double g[2];
                    W[idx] = mo_interp2(coeff, size[0], size[1], degree, p, vmin, vmax, g);
                    if (mgrad) {
                        mgrad[2 * idx] = g[0];
                        mgrad[2 * idx + 1] = g[1];
                    }
                }
            }
}

/* ------------------------------------------------------------------------ */
/* a6/a7: image gradient, SSD value and force                                  */
/* ------------------------------------------------------------------------ */

/* Central-difference gradient with zero components on the border of that
 * axis, mia/3d/image.cc:347-477 and mia/2d/image.cc:273-279,351-398. */
MO_EXPORT void mo_image_gradient(int dim, const int size[3], const double *W, double *g)
{
    const int nx = size[0], ny = size[1], nz = dim == 3 ? size[2] : 1;
    const size_t sx = 1, sy = nx, sz = (size_t)nx * ny;
    size_t idx = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x, ++idx) {
                double gx = (x > 0 && x < nx - 1) ? 0.5 * (W[idx + sx] - W[idx - sx]) : 0.0;
                double gy = (y > 0 && y < ny - 1) ? 0.5 * (W[idx + sy] - W[idx - sy]) : 0.0;
                g[dim * idx] = gx;
                g[dim * idx + 1] = gy;
                if (dim == 3)
                    g[dim * idx + 2] = (z > 0 && z < nz - 1) ? 0.5 * (W[idx + sz] - W[idx - sz]) : 0.0;
            }
}

/* SSD value, mia/template/ssd.hh:54-99: 0.5*sum d^2 (times 1/N with norm);
 * with autothresh > 0 only voxels with W > thresh count and the sum is divided
 * by their number (FLT_MAX if there are none). */
MO_EXPORT double mo_ssd_value(size_t n, const double *W, const float *R, int norm, float autothresh)
{
    double sum = 0.0;
    if (autothresh == 0.0f) {
        for (size_t i = 0; i < n; ++i) {
            const double d = W[i] - (double)R[i];
            sum += d * d;
        }
        return (norm ? 0.5 / n : 0.5) * sum;
    }
    long cnt = 0;
    for (size_t i = 0; i < n; ++i)
        if (W[i] > autothresh) {
            const double d = W[i] - (double)R[i];
            sum += d * d;
            ++cnt;
        }
    return cnt > 0 ? 0.5 * sum / cnt : (double)FLT_MAX;
}

/* SSD force field, mia/template/ssd.hh:125-181: force = grad * (W - R) * s
 * with s = 1/N (norm) or 1; with autothresh only masked voxels get a force
 * and s = 1/count.  grad is the per-voxel gradient (dim components); returns
 * the cost 0.5*sum d^2*s as FEvalForce does (FLT_MAX if the mask is empty). */
MO_EXPORT double mo_ssd_force(int dim, size_t n, const double *W, const float *R, const double *grad,
                              int norm, float autothresh, double *force)
{
    double cost = 0.0;
    memset(force, 0, sizeof(double) * dim * n);
    if (autothresh == 0.0f) {
        const double s = norm ? 1.0 / n : 1.0;
        for (size_t i = 0; i < n; ++i) {
            const double d = W[i] - (double)R[i];
            for (int k = 0; k < dim; ++k)
                force[dim * i + k] = grad[dim * i + k] * d * s;
            cost += d * d * s;
        }
        return 0.5 * cost;
    }
    long cnt = 0;
    for (size_t i = 0; i < n; ++i)
        if (W[i] > autothresh) {
            const double d = W[i] - (double)R[i];
            for (int k = 0; k < dim; ++k)
                force[dim * i + k] = grad[dim * i + k] * d;
            cost += d * d;
            ++cnt;
        }
    if (cnt == 0)
        return (double)FLT_MAX;
    const double s = 1.0 / cnt;
    for (size_t i = 0; i < dim * n; ++i)
        force[i] *= s;
    return 0.5 * cost * s;
}

/* ------------------------------------------------------------------------ */
/* a8: translate (adjoint of the displacement w.r.t. the coefficients)        */
/* ------------------------------------------------------------------------ */

/* params = -sum_x B_k(x) force(x), separable over the axes.
 * mode 0 ("reference"): the adjoint rows of spline.cc:697-791 (3D) /
 *   2D spline.cc:509-576, i.e. beta_3(i*scale + 1 - o) in double;
 * mode 1 ("forward transpose"): the exact transpose of mo_displacement's
 *   forward tables (float positions, fastfloor, trimmed last weight), which is
 *   the derivative FD of the cost measures. */
MO_EXPORT void mo_translate(int dim, const int size[3], const int csize[3], const double *force,
                            int mode, double *params)
{
    const int nx = size[0], ny = size[1], nz = dim == 3 ? size[2] : 1;
    const int cx = csize[0], cy = csize[1], cz = dim == 3 ? csize[2] : 1;
    double *Mx = (double *)calloc((size_t)nx * cx, sizeof(double));
    double *My = (double *)calloc((size_t)ny * cy, sizeof(double));
    double *Mz = (double *)calloc((size_t)nz * cz, sizeof(double));
    if (mode == 0) {
        mo_spline_derivative_rows(nx, cx, Mx);
        mo_spline_derivative_rows(ny, cy, My);
        if (dim == 3)
            mo_spline_derivative_rows(nz, cz, Mz);
        else
            Mz[0] = 1.0;
    } else {
        const int ns[3] = {nx, ny, nz};
        const int cs[3] = {cx, cy, cz};
        double *Ms[3] = {Mx, My, Mz};
        for (int a = 0; a < dim; ++a) {
            axis_t t;
            axis_init(&t, ns[a], cs[a]);
            for (int i = 0; i < ns[a]; ++i)
                for (int k = 0; k < t.nw[i]; ++k) {
                    const int o = t.start[i] + k;
                    if (o < cs[a])
                        Ms[a][(size_t)o * ns[a] + i] += t.w[4 * i + k];
                }
            axis_free(&t);
        }
        if (dim == 2)
            Mz[0] = 1.0;
    }
    /* contract z, then y, then x (the order of spline.cc:731-791) */
    double *t1 = (double *)calloc((size_t)dim * nx * ny * cz, sizeof(double));
    for (int oz = 0; oz < cz; ++oz)
        for (int z = 0; z < nz; ++z) {
            const double w = Mz[(size_t)oz * nz + z];
            if (w == 0.0) continue;
            const double *src = force + (size_t)dim * nx * ny * z;
            double *dst = t1 + (size_t)dim * nx * ny * oz;
            for (size_t i = 0; i < (size_t)dim * nx * ny; ++i)
                dst[i] += w * src[i];
        }
    double *t2 = (double *)calloc((size_t)dim * nx * cy * cz, sizeof(double));
    for (int oz = 0; oz < cz; ++oz)
        for (int oy = 0; oy < cy; ++oy)
            for (int y = 0; y < ny; ++y) {
                const double w = My[(size_t)oy * ny + y];
                if (w == 0.0) continue;
                const double *src = t1 + (size_t)dim * nx * (y + (size_t)ny * oz);
                double *dst = t2 + (size_t)dim * nx * (oy + (size_t)cy * oz);
                for (size_t i = 0; i < (size_t)dim * nx; ++i)
                    dst[i] += w * src[i];
            }
    for (int oz = 0; oz < cz; ++oz)
        for (int oy = 0; oy < cy; ++oy)
            for (int ox = 0; ox < cx; ++ox) {
                double acc[3] = {0, 0, 0};
                const double *src = t2 + (size_t)dim * nx * (oy + (size_t)cy * oz);
                for (int x = 0; x < nx; ++x) {
                    const double w = Mx[(size_t)ox * nx + x];
                    if (w == 0.0) continue;
                    for (int d = 0; d < dim; ++d)
                        acc[d] += w * src[dim * x + d];
                }
                double *dst = params + (size_t)dim * (((size_t)oz * cy + oy) * cx + ox);
                for (int d = 0; d < dim; ++d)
                    dst[d] = -acc[d];
            }
    free(t1);
    free(t2);
    free(Mx);
    free(My);
    free(Mz);
}

/* ------------------------------------------------------------------------ */
/* a9: the image full cost, evaluate(t, gradient)                              */
/* ------------------------------------------------------------------------ */

/* Cost and gradient of image:cost=ssd under the spline transformation:
 * warp -> force -> translate -> value, mia/3d/fullcost/image.cc:67-78
 * (returns the double SSD value) and mia/2d/fullcost/image.cc:81-105 (returns
 * the cost from evaluate_force; in FP64 both are the same number), scaled by
 * the full-cost weight (mia/template/fullcost.cxx:36-45).
 * grad_mode 0: reference semantics (central differences of the warped image,
 *              reference adjoint rows);
 * grad_mode 1: exact derivative of the returned cost (analytic gradient of the
 *              interpolated moving image at T(x), forward-transpose adjoint).
 * Optional outputs: warped (N doubles), force (dim*N doubles). */
MO_EXPORT double mo_cost_grad(int dim, const int size[3], const float *moving, const float *fixed,
                              const int csize[3], const double *params, int degree, int norm,
                              float autothresh, int grad_mode, double weight, double *grad,
                              double *warped, double *force_out)
{
    const size_t n = (size_t)size[0] * size[1] * (dim == 3 ? size[2] : 1);
    double vmin, vmax;
    double *coeff = (double *)malloc(sizeof(double) * n);
    mo_prefilter(moving, size[0], size[1], dim == 3 ? size[2] : 1, degree, coeff, &vmin, &vmax);
    double *u = (double *)malloc(sizeof(double) * dim * n);
    mo_displacement(dim, size, csize, params, u);
    double *W = (double *)malloc(sizeof(double) * n);
    double *g = (double *)malloc(sizeof(double) * dim * n);
    mo_warp(dim, size, coeff, degree, vmin, vmax, u, W, grad_mode == 1 ? g : NULL);
    if (grad_mode == 0)
        mo_image_gradient(dim, size, W, g);
    double *force = (double *)malloc(sizeof(double) * dim * n);
    mo_ssd_force(dim, n, W, fixed, g, norm, autothresh, force);
    const double cost = mo_ssd_value(n, W, fixed, norm, autothresh);
    if (grad) {
        const size_t np = (size_t)dim * csize[0] * csize[1] * (dim == 3 ? csize[2] : 1);
        mo_translate(dim, size, csize, force, grad_mode == 1 ? 1 : 0, grad);
        for (size_t i = 0; i < np; ++i)
            grad[i] *= weight;
    }
    if (warped)
        memcpy(warped, W, sizeof(double) * n);
    if (force_out)
        memcpy(force_out, force, sizeof(double) * dim * n);
    free(coeff);
    free(u);
    free(W);
    free(g);
    free(force);
    return weight * cost;
}

/* Cost only (the value FD differences are taken of). */
MO_EXPORT double mo_cost(int dim, const int size[3], const float *moving, const float *fixed,
                         const int csize[3], const double *params, int degree, int norm,
                         float autothresh)
{
    return mo_cost_grad(dim, size, moving, fixed, csize, params, degree, norm, autothresh, 0, 1.0,
                        NULL, NULL, NULL);
}

/* The part of the unnormalised SSD 0.5 * sum (W - R)^2 (no autothresh) that
 * comes from the voxels of the box [lo, hi) (x, y, z), for the prefiltered
 * moving image `coeff` with clamp range [vmin, vmax].  A change of one
 * coefficient moves W only inside its support (the voxels whose four spline
 * weights along every axis include it), so the finite difference of mo_cost
 * in that coefficient equals the difference of this sum over a box holding
 * the support: the full-size finite-difference check costs one support per
 * coefficient instead of two whole-image evaluations.  Per voxel it repeats
 * mo_displacement's sum and mo_interp3 / mo_interp2 (the arithmetic of
 * mo_cost_grad). */
MO_EXPORT double mo_cost_box(int dim, const int size[3], const double *coeff, double vmin, double vmax,
                             const float *fixed, const int csize[3], const double *params, int degree,
                             const int lo[3], const int hi[3])
{
    axis_t ax, ay, az;
    const int nz = dim == 3 ? size[2] : 1;
    const int csz = dim == 3 ? csize[2] : 1;
    axis_init(&ax, size[0], csize[0]);
    axis_init(&ay, size[1], csize[1]);
    if (dim == 3)
        axis_init(&az, size[2], csize[2]);
    const int z0 = dim == 3 ? lo[2] : 0, z1 = dim == 3 ? hi[2] : 1;
    double sum = 0.0;
    for (int z = z0; z < z1 && z < nz; ++z)
        for (int y = lo[1]; y < hi[1] && y < size[1]; ++y)
            for (int x = lo[0]; x < hi[0] && x < size[0]; ++x) {
                double u[3] = {0, 0, 0};
                const int nwz = dim == 3 ? az.nw[z] : 1;
                for (int c = 0; c < nwz; ++c) {
                    const int kz = dim == 3 ? az.start[z] + c : 0;
                    const double wz = dim == 3 ? az.w[4 * z + c] : 1.0;
                    if (kz >= csz) break;
                    for (int b = 0; b < ay.nw[y]; ++b) {
                        const int ky = ay.start[y] + b;
                        if (ky >= csize[1]) break;
                        const double wy = ay.w[4 * y + b];
                        for (int a = 0; a < ax.nw[x]; ++a) {
                            const int kx = ax.start[x] + a;
                            if (kx >= csize[0]) break;
                            const double wgt = wz * wy * ax.w[4 * x + a];
                            const double *cc = params + dim * (((size_t)kz * csize[1] + ky) * csize[0] + kx);
                            for (int d = 0; d < dim; ++d)
                                u[d] += wgt * cc[d];
                        }
                    }
                }
                const size_t idx = ((size_t)z * size[1] + y) * size[0] + x;
                double w;
                if (dim == 3) {
                    const double p[3] = {x - u[0], y - u[1], z - u[2]};
                    w = mo_interp3(coeff, size[0], size[1], size[2], degree, p, vmin, vmax, NULL);
                } else {
                    const double p[2] = {x - u[0], y - u[1]};
                    w = mo_interp2(coeff, size[0], size[1], degree, p, vmin, vmax, NULL);
                }
                const double d = w - (double)fixed[idx];
                sum += d * d;
            }
    axis_free(&ax);
    axis_free(&ay);
    if (dim == 3)
        axis_free(&az);
    return 0.5 * sum;
}

/* ------------------------------------------------------------------------ */
/* f3: image resampling (the `scale` filter)                                 */
/* ------------------------------------------------------------------------ */

/* Binomial filter of half width F over a line, with the boundary handling of
 * C1DGaussFilterKernel::do_apply (mia/core/spacialkernel/gauss.cc:76-140):
 * the taps that fall before the start read data[F-i-1 .. 0] reflected, the
 * ones after the end read the line backwards from its last sample.  Kernel:
 * row 2F of Pascal's triangle / 2^(2F) (gauss.cc:25-55). */
MO_EXPORT void mo_gauss_line(const double *data, int n, int F, double *out)
{
    const int K = 2 * F + 1;
    double *g = (double *)malloc(sizeof(double) * K);
    for (int k = 0; k < K; ++k)
        g[k] = 0.0;
    g[0] = 1.0;
    for (int i = 1; i < K; ++i)
        for (int j = i; j > 0; --j)
            g[j] += g[j - 1];
    const double norm = pow(2.0, K - 1);
    for (int k = 0; k < K; ++k)
        g[k] /= norm;
    for (int k = 0; k < n; ++k)
        out[k] = 0.0;
    int trgt = 0, pos = n;
    for (int i = 0; i < F && pos >= 0 && trgt < n; ++i, --pos, ++trgt) {
        double ht = 0.0;
        int ik = 0;
        for (int k = F - i; k > 0; --k, ++ik)
            ht += data[(k - 1) < n ? (k - 1) : (n - 1)] * g[ik];
        for (int s = 0; ik < K && s < n; ++s, ++ik)
            ht += data[s] * g[ik];
        out[trgt] = ht;
    }
    int start = 0;
    while (pos-- > F) {
        double ht = 0.0;
        for (int j = 0; j < K; ++j)
            ht += g[j] * data[start + j];
        out[trgt++] = ht;
        ++start;
    }
    for (int i = 0; i < F && pos >= 0 && trgt < n && start < n; ++i, --pos, ++trgt, ++start) {
        double ht = 0.0;
        int ik = 0;
        for (int s = start; s < n && ik < K; ++s, ++ik)
            ht += data[s] * g[ik];
        for (int r = n - 1; ik < K && r >= 0; --r, ++ik)
            ht += data[r] * g[ik];
        out[trgt] = ht;
    }
    free(g);
}

/* C1DScalar with the cubic kernel (mia/core/scaler1d.cc:26-230): scale =
 * (n_out - 1) / (n_in - 1); upscale prefilters the line and samples at x
 * advanced by 1/scale per output sample (repeated addition); downscale first
 * applies the binomial filter of half width floor(1/scale + 0.5), then
 * prefilters and samples at i / scale; indices mirrored at the line ends. */
MO_EXPORT void mo_scale_line(const double *in, int n_in, double *out, int n_out)
{
    if (n_in == 1) {
        for (int i = 0; i < n_out; ++i)
            out[i] = in[0];
        return;
    }
    const double scale = (double)(n_out - 1) / (double)(n_in - 1);
    if (scale == 1.0) {
        for (int i = 0; i < n_out; ++i)
            out[i] = in[i];
        return;
    }
    double *c = (double *)malloc(sizeof(double) * n_in);
    const double dx = 1.0 / scale;
    if (scale > 1.0) {
        for (int k = 0; k < n_in; ++k)
            c[k] = in[k];
    } else {
        mo_gauss_line(in, n_in, (int)floor(1.0 / scale + 0.5), c);
    }
    mo_filter_line(c, n_in, 3);
    double x = 0.0;
    for (int i = 0; i < n_out; ++i, x += dx) {
        const double xi = scale > 1.0 ? x : i * dx;
        double w[4];
        const int st = mo_fastfloor(xi);
        mo_bspline3_weights(xi - st, w);
        double v = 0.0;
        for (int k = 0; k < 4; ++k)
            v += w[k] * c[mo_mirror_index(st - 1 + k, n_in)];
        out[i] = v;
    }
    free(c);
}

/* Separable resampling of a float image, x then y then z, double between the
 * axes, float result (mia/3d/filter/scale.cc:38-104; the 2D filter is the
 * same on two axes).  Axes whose size does not change are copied. */
MO_EXPORT void mo_scale_image(const float *img, const int in_size[3], const int out_size[3], float *out)
{
    int cur[3] = {in_size[0], in_size[1], in_size[2]};
    size_t n = (size_t)cur[0] * cur[1] * cur[2];
    double *a = (double *)malloc(sizeof(double) * n);
    for (size_t i = 0; i < n; ++i)
        a[i] = img[i];
    for (int ax = 0; ax < 3; ++ax) {
        if (cur[ax] == out_size[ax])
            continue;
        int nxt[3] = {cur[0], cur[1], cur[2]};
        nxt[ax] = out_size[ax];
        const size_t m = (size_t)nxt[0] * nxt[1] * nxt[2];
        double *b = (double *)malloc(sizeof(double) * m);
        double *li = (double *)malloc(sizeof(double) * cur[ax]);
        double *lo = (double *)malloc(sizeof(double) * nxt[ax]);
        const int o1 = ax == 0 ? cur[1] : cur[0];
        const int o2 = ax == 2 ? cur[1] : cur[2];
        for (int q = 0; q < o2; ++q)
            for (int p = 0; p < o1; ++p) {
                for (int k = 0; k < cur[ax]; ++k) {
                    const int x = ax == 0 ? k : p, y = ax == 1 ? k : (ax == 0 ? p : q), z = ax == 2 ? k : q;
                    li[k] = a[((size_t)z * cur[1] + y) * cur[0] + x];
                }
                mo_scale_line(li, cur[ax], lo, nxt[ax]);
                for (int k = 0; k < nxt[ax]; ++k) {
                    const int x = ax == 0 ? k : p, y = ax == 1 ? k : (ax == 0 ? p : q), z = ax == 2 ? k : q;
                    b[((size_t)z * nxt[1] + y) * nxt[0] + x] = lo[k];
                }
            }
        free(li);
        free(lo);
        free(a);
        a = b;
        cur[ax] = nxt[ax];
    }
    n = (size_t)cur[0] * cur[1] * cur[2];
    for (size_t i = 0; i < n; ++i)
        out[i] = (float)a[i];
    free(a);
}

/* ------------------------------------------------------------------------ */
/* f2: the divcurl penalty                                                   */
/* ------------------------------------------------------------------------ */

/* Cubic B-spline and its derivatives, bspline<3,0..2>::apply
 * (mia/core/splinekernel/bspline.cc:328-375). */
MO_EXPORT double mo_bspline3_deg(double x, int deg)
{
    if (deg == 0) {
        x = fabs(x);
        if (x <= 1.0)
            return 2.0 / 3.0 - x * x * (1 - 0.5 * x);
        if (x < 2.0)
            return (2.0 - x) * (2.0 - x) * (2.0 - x) / 6.0;
        return 0.0;
    }
    if (deg == 1) {
        const double ax = fabs(x);
        if (ax <= 1.0)
            return (1.5 * ax - 2) * x;
        if (ax <= 2.0)
            return (x > 0.0 ? -0.5 : 0.5) * (2 - ax) * (2 - ax);
        return 0.0;
    }
    x = fabs(x);
    if (x > 2.0)
        return 0.0;
    if (x > 1.0)
        return 2.0 - x;
    return 3.0 * x - 2.0;
}

/* integrate2 (mia/core/splinekernel.cc:287-315) with n = 1: composite
 * Simpson (mia/core/simpson.hh:36-52), 8 intervals per unit, of
 * B^(d1)(x - s1) B^(d2)(x - s2) over the overlap of both supports and [x0, L]. */
MO_EXPORT double mo_integrate2(double s1, double s2, int d1, int d2, double x0, double L)
{
    double a = s1 - 2.0, b = s1 + 2.0;
    if (a < s2 - 2.0) a = s2 - 2.0;
    if (a < x0) a = x0;
    if (b > s2 + 2.0) b = s2 + 2.0;
    if (b > L) b = L;
    if (b <= a)
        return 0.0;
    const size_t iv = (size_t)(8 * (b - a));
    const double dx = (b - a) / iv;
    double sum = 0.5 * (mo_bspline3_deg(a - s1, d1) * mo_bspline3_deg(a - s2, d2) +
                        mo_bspline3_deg(b - s1, d1) * mo_bspline3_deg(b - s2, d2));
    double x = a + dx;
    for (size_t k = 1; k < iv; ++k, x += dx)
        sum += mo_bspline3_deg(x - s1, d1) * mo_bspline3_deg(x - s2, d2);
    x = a + 0.5 * dx;
    for (size_t k = 0; k < iv; ++k, x += dx)
        sum += 2 * mo_bspline3_deg(x - s1, d1) * mo_bspline3_deg(x - s2, d2);
    return sum * dx / 3.0;
}

/* The 9^3 matrix cells of C3DPPDivcurlMatrixImpl::reset
 * (mia/3d/ppmatrix.cc:178-262), 6 doubles each in the reference's field
 * order vxx, vxz, vxy, vyy, vyz, vzz; h = (csize - 1) / range in float. */
MO_EXPORT void mo_pp_cells(const int csize[3], const float range[3], double wdiv, double wrot, double *cells)
{
    float h1[3], h2[3], h3[3], h4[3];
    for (int a = 0; a < 3; ++a) {
        h1[a] = (float)(csize[a] - 1) / range[a];
        h2[a] = h1[a] * h1[a];
        h3[a] = h1[a] * h2[a];
        h4[a] = h1[a] * h3[a];
    }
    const double gs = 1.0 / (h1[0] * h1[1] * h1[2]);
    const double wsum = wdiv + wrot, wdelta = 2 * (wdiv - wrot);
    size_t k = 0;
#define MO_I(d, p, q) mo_integrate2(4.0, (d) + 4.0, p, q, 0.0, 8.0)
    for (int z = -4; z <= 4; ++z)
        for (int y = -4; y <= 4; ++y)
            for (int x = -4; x <= 4; ++x, ++k) {
                const double r00z = MO_I(z, 0, 0), r01z = h1[2] * MO_I(z, 0, 1), r11z = h2[2] * MO_I(z, 1, 1);
                const double r12z = h3[2] * MO_I(z, 1, 2), r22z = h4[2] * MO_I(z, 2, 2);
                const double r00y = MO_I(y, 0, 0), r01y = h1[1] * MO_I(y, 0, 1), r10y = h1[1] * MO_I(y, 1, 0);
                const double r11y = h2[1] * MO_I(y, 1, 1), r12y = h3[1] * MO_I(y, 1, 2);
                const double r21y = h3[1] * MO_I(y, 2, 1), r22y = h4[1] * MO_I(y, 2, 2);
                const double r00x = MO_I(x, 0, 0), r01x = h1[0] * MO_I(x, 0, 1), r10x = h1[0] * MO_I(x, 1, 0);
                const double r11x = h2[0] * MO_I(x, 1, 1), r21x = h3[0] * MO_I(x, 2, 1);
                const double r22x = h4[0] * MO_I(x, 2, 2);
                const double r200200 = r22x * r00y * r00z, r020020 = r00x * r22y * r00z;
                const double r002002 = r00x * r00y * r22z, r110110 = r11x * r11y * r00z;
                const double r101101 = r11x * r00y * r11z, r011011 = r00x * r11y * r11z;
                double *c = cells + 6 * k;
                c[0] = gs * (r002002 * wrot + r011011 * 2.0 * wrot + r020020 * wrot + r101101 * wsum +
                             r110110 * wsum + r200200 * wdiv);
                c[3] = gs * (r002002 * wrot + r011011 * wsum + r020020 * wdiv + r101101 * 2 * wrot +
                             r110110 * wsum + r200200 * wrot);
                c[5] = gs * (r002002 * wdiv + r011011 * wsum + r020020 * wrot + r101101 * wsum +
                             r110110 * wrot * 2 + r200200 * wrot);
                c[1] = c[2] = c[4] = 0.0;
                if (wdelta != 0.0) {
                    c[2] = gs * wdelta * (r21x * r01y * r00z + r10x * r12y * r00z + r01x * r10y * r11z);
                    c[1] = gs * wdelta * (r21x * r00y * r01z + r10x * r11y * r01z + r10x * r00y * r12z);
                    c[4] = gs * wdelta * (r11x * r10y * r01z + r00x * r21y * r01z + r00x * r10y * r12z);
                }
            }
#undef MO_I
}

/* C3DPPDivcurlMatrixImpl::evaluate (mia/3d/ppmatrix.cc:300-505): value
 * sum_i sum_j c_i (P_(i-j)) c_j over j in [i-4, i+3] per axis and, if grad
 * is not NULL, its derivative per coefficient; coefficients 3 floats each. */
MO_EXPORT double mo_divcurl(const int cs[3], const double *cells, const float *coef, double *grad)
{
    const int nx = cs[0], ny = cs[1], nz = cs[2];
    double total = 0.0;
    for (int zi = 0; zi < nz; ++zi) {
        double r[6] = {0, 0, 0, 0, 0, 0};
        for (int yi = 0; yi < ny; ++yi)
            for (int xi = 0; xi < nx; ++xi) {
                const size_t i = ((size_t)zi * ny + yi) * nx + xi;
                const double cix = coef[3 * i], ciy = coef[3 * i + 1], ciz = coef[3 * i + 2];
                double g[3] = {0, 0, 0};
                for (int zj = (zi - 4 > 0 ? zi - 4 : 0); zj < (zi + 4 < nz ? zi + 4 : nz); ++zj)
                    for (int yj = (yi - 4 > 0 ? yi - 4 : 0); yj < (yi + 4 < ny ? yi + 4 : ny); ++yj)
                        for (int xj = (xi - 4 > 0 ? xi - 4 : 0); xj < (xi + 4 < nx ? xi + 4 : nx); ++xj) {
                            const size_t j = ((size_t)zj * ny + yj) * nx + xj;
                            const double *p = cells + 6 * (((size_t)(zi - zj + 4) * 9 + (yi - yj + 4)) * 9 +
                                                           (xi - xj + 4));
                            const double cx = coef[3 * j], cy = coef[3 * j + 1], cz = coef[3 * j + 2];
                            const double vxx = p[0], vxz = p[1], vxy = p[2], vyy = p[3], vyz = p[4], vzz = p[5];
                            r[0] += cix * cx * vxx;
                            r[1] += cix * cy * vxy;
                            r[2] += ciy * cy * vyy;
                            r[3] += ciz * cx * vxz;
                            r[4] += ciy * cz * vyz;
                            r[5] += ciz * cz * vzz;
                            g[0] += 2 * cx * vxx + cy * vxy + cz * vxz;
                            g[1] += 2 * cy * vyy + cz * vyz + cx * vxy;
                            g[2] += 2 * cz * vzz + cx * vxz + cy * vyz;
                        }
                if (grad) {
                    grad[3 * i] = g[0];
                    grad[3 * i + 1] = g[1];
                    grad[3 * i + 2] = g[2];
                }
            }
        total += r[0] + r[1] + r[2] + r[3] + r[4] + r[5];
    }
    return total;
}

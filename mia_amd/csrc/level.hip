// Per-level work of the multiresolution registration on the GPU:
//   * image resampling with the reference's `scale` filter semantics
//     (mia/3d/filter/scale.cc:38-104, mia/2d/filter/scale.cc, mia/core/scaler1d.cc):
//     per axis and line, an optional binomial ("gauss:w=round(1/scale)")
//     prefilter (mia/core/spacialkernel/gauss.cc:25-55,76-140), the cubic
//     B-spline prefilter with mirror boundaries and resampling with cubic
//     B-spline weights and mirrored indices, in double between the passes;
//   * the divcurl spline penalty, a banded quadratic form in the coefficients
//     (mia/3d/ppmatrix.cc, mia/3d/splinepenalty/divcurl.cc) with its gradient.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mia_hip.h"
#include "common.hh"

using mia_hip_rt::fail;

namespace {

constexpr int kLB = 256;

// Per-device state of the entry points that hold no context, kept until the
// process ends: one stream (each entry point waits for its own work on it
// before it returns) and a scratch allocation for the resampling that grows
// to the largest call so far (used under the mutex).
struct DeviceCache {
    std::mutex m;
    hipStream_t stream = nullptr;
    void *scratch = nullptr;
    size_t cap = 0;
};
constexpr int kMaxCacheDevices = 64;

DeviceCache *device_cache(int device)
{
    static DeviceCache caches[kMaxCacheDevices];
    return device >= 0 && device < kMaxCacheDevices ? &caches[device] : nullptr;
}

// the device's shared stream, taken on first use (call with the mutex held
// and the device current)
hipError_t cache_stream(DeviceCache &c, hipStream_t *s)
{
    if (!c.stream) {
        int device = 0;
        hipError_t e = hipGetDevice(&device);
        if (e == hipSuccess)
            e = mia_hip_rt::stream_acquire(device, &c.stream);
        if (e != hipSuccess) {
            c.stream = nullptr;
            return e;
        }
    }
    *s = c.stream;
    return hipSuccess;
}

__device__ __forceinline__ int mirror_line(int i, int n)
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

// ---------------------------------------------------------------------------
// scale filter
// ---------------------------------------------------------------------------

enum ScaleStrategy { SC_COPY = 0, SC_FILL, SC_UP, SC_DOWN };

struct ScalePass {
    int axis;          // 0 x, 1 y, 2 z
    int nx, ny;        // input volume extent in x and y (the line grid)
    int n_in, n_out;   // line length before / after
    int strategy;
    double dx;         // sample step on the input line
    int F;             // binomial half width (downscale)
};

// One line per thread.  Input is float (first pass) or double, output double
// or float (last pass).  scratch holds the line's coefficients, element k of
// line l at k * nlines + l (coalesced across threads).
__global__ void __launch_bounds__(kLB) k_scale_lines(const float *__restrict__ in_f, const double *__restrict__ in_d,
                                                     double *__restrict__ out_d, float *__restrict__ out_f,
                                                     double *__restrict__ scratch,
                                                     const double *__restrict__ gauss, ScalePass P,
                                                     long long nlines)
{
    const long long l = (long long)blockIdx.x * kLB + threadIdx.x;
    if (l >= nlines)
        return;
    long long ib, ob, is, os;
    if (P.axis == 0) {
        ib = l * P.n_in;
        ob = l * P.n_out;
        is = os = 1;
    } else if (P.axis == 1) {
        const long long z = l / P.nx, x = l % P.nx;
        ib = z * P.nx * P.n_in + x;
        ob = z * P.nx * P.n_out + x;
        is = os = P.nx;
    } else {
        ib = ob = l;
        is = os = (long long)P.nx * P.ny;
    }
    auto rd = [&](int k) -> double { return in_f ? (double)in_f[ib + k * is] : in_d[ib + k * is]; };
    auto wr = [&](int k, double v) {
        if (out_f)
            out_f[ob + k * os] = (float)v;
        else
            out_d[ob + k * os] = v;
    };
    const int N = P.n_in;
    auto C = [&](int k) -> double & { return scratch[(long long)k * nlines + l]; };
    if (P.strategy == SC_COPY) {
        for (int k = 0; k < N; ++k)
            wr(k, rd(k));
        return;
    }
    if (P.strategy == SC_FILL) {
        const double v = rd(0);
        for (int k = 0; k < P.n_out; ++k)
            wr(k, v);
        return;
    }
    if (P.strategy == SC_UP) {
        for (int k = 0; k < N; ++k)
            C(k) = rd(k);
    } else {
        // binomial filter with the reference's boundary handling
        // (gauss.cc:76-140): reflected taps at both ends, no write beyond N
        const int F = P.F, K = 2 * F + 1;
        int trgt = 0, pos = N;
        for (int i = 0; i < F && pos >= 0 && trgt < N; ++i, --pos, ++trgt) {
            double ht = 0.0;
            int ik = 0;
            for (int k = F - i; k > 0; --k, ++ik)
                ht += rd(min(k - 1, N - 1)) * gauss[ik];
            for (int s = 0; ik < K && s < N; ++s, ++ik)
                ht += rd(s) * gauss[ik];
            C(trgt) = ht;
        }
        int start = 0;
        while (pos-- > F) {
            double ht = 0.0;
            for (int j = 0; j < K; ++j)
                ht += gauss[j] * rd(start + j);
            C(trgt++) = ht;
            ++start;
        }
        for (int i = 0; i < F && pos >= 0 && trgt < N && start < N; ++i, --pos, ++trgt, ++start) {
            double ht = 0.0;
            int ik = 0;
            for (int s = start; s < N && ik < K; ++s, ++ik)
                ht += rd(s) * gauss[ik];
            for (int r = N - 1; ik < K && r >= 0; --r, ++ik)
                ht += rd(r) * gauss[ik];
            C(trgt) = ht;
        }
    }
    // cubic B-spline prefilter, mirror boundaries (boundary_conditions.cc:62-103)
    if (N >= 2) {
        const double p = 1.7320508075688772 - 2.0;
        const double lambda = 2.0 - p - 1.0 / p;
        for (int k = 0; k < N; ++k)
            C(k) *= lambda;
        double zn = p, z2n = pow(p, (double)(N - 1));
        const double iz = 1.0 / p;
        double sum = C(0) + z2n * C(N - 1);
        z2n *= z2n * iz;
        for (int k = 1; k < N - 1; ++k) {
            sum += (zn + z2n) * C(k);
            zn *= p;
            z2n *= iz;
        }
        C(0) = sum / (1.0 - zn * zn);
        for (int k = 1; k < N; ++k)
            C(k) += p * C(k - 1);
        C(N - 1) = (p / (p * p - 1.0)) * (p * C(N - 2) + C(N - 1));
        for (int k = N - 2; k >= 0; --k)
            C(k) = p * (C(k + 1) - C(k));
    }
    // resampling: upscale steps x by repeated addition, downscale uses i * dx
    // (scaler1d.cc:49-91)
    double x = 0.0;
    for (int i = 0; i < P.n_out; ++i) {
        const double xi = P.strategy == SC_UP ? x : i * P.dx;
        const double fl = floor(rint(xi * 65536.0) / 65536.0);
        const double f = xi - fl;
        const double fm1 = 1.0 - f;
        double w[4];
        w[3] = f * f * f / 6.0;
        w[0] = fm1 * fm1 * fm1 / 6.0;
        w[2] = f - 2.0 * w[3] + w[0];
        w[1] = fm1 - 2.0 * w[0] + w[3];
        const int st = (int)fl - 1;
        double v = 0.0;
        for (int k = 0; k < 4; ++k)
            v += w[k] * C(mirror_line(st + k, N));
        wr(i, v);
        x += P.dx;
    }
}

// ---------------------------------------------------------------------------
// divcurl penalty
// ---------------------------------------------------------------------------

struct PPCell {
    double vxx, vxz, vxy, vyy, vyz, vzz; // the reference's field order (ppmatrix.cc:34-41)
};
constexpr int kPK = 4;                  // kernel support of bspline:d=3
constexpr int kPW = 2 * kPK + 1;        // cells per axis

// A coefficient component as the penalty reads it: the transformation keeps
// float coefficients (the double parameters rounded to float,
// mia/3d/transform/spline.cc:344-358), used in double.
__device__ __forceinline__ double coef_value(const float *c, long long k)
{
    return (double)c[k];
}

__device__ __forceinline__ double coef_value(const double *c, long long k)
{
    return (double)(float)c[k];
}

// Threads per coefficient (one per z plane of its neighbourhood) and
// coefficients per block of the divcurl kernel.
constexpr int kDcPlanes = 2 * kPK;
constexpr int kDcCoef = kLB / kDcPlanes;

// Coefficient i's gradient g_i and row sum of the quadratic form over the
// neighbours j with i - 4 <= j <= i + 3 per axis (the reference's loop
// bounds, ppmatrix.cc:330-345): thread (i, q) takes the z plane
// zj = zi - 4 + q, the eight planes' sums are added in ascending order,
// and the block's coefficients' row sums in a fixed order into part[block].
// SCALED: grad[3i + k] = -w * g_k (the penalty's weighted gradient with the
// reference's sign, splinetransformpenalty.cc:60-71, the product rounded on
// its own, to be added to the image cost's gradient); otherwise grad = g
// (unweighted, may be NULL).
template <typename T, bool SCALED>
__global__ void __launch_bounds__(kLB) k_divcurl(const T *__restrict__ coef, const PPCell *__restrict__ cells,
                                                 int nx, int ny, int nz, double w, double *__restrict__ grad,
                                                 double *__restrict__ part)
{
    __shared__ PPCell P[kPW * kPW * kPW];
    __shared__ double acc[kDcCoef][kDcPlanes][10];
    __shared__ double red[kLB / 64];
    for (int k = threadIdx.x; k < kPW * kPW * kPW; k += kLB)
        P[k] = cells[k];
    __syncthreads();
    const long long n = (long long)nx * ny * nz;
    const int cl = threadIdx.x / kDcPlanes, q = threadIdx.x % kDcPlanes;
    const long long i = (long long)blockIdx.x * kDcCoef + cl;
    double r0 = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0, r5 = 0, g0 = 0, g1 = 0, g2 = 0;
    double cix = 0.0, ciy = 0.0, ciz = 0.0;
    if (i < n) {
        const int xi = (int)(i % nx), yi = (int)((i / nx) % ny), zi = (int)(i / ((long long)nx * ny));
        cix = coef_value(coef, 3 * i);
        ciy = coef_value(coef, 3 * i + 1);
        ciz = coef_value(coef, 3 * i + 2);
        const int zj = zi - kPK + q;
        if (zj >= 0 && zj < nz) {
            const int dz = zi - zj + kPK;
            for (int yj = max(0, yi - kPK); yj < min(yi + kPK, ny); ++yj) {
                const int dy = yi - yj + kPK;
                const int xs = max(0, xi - kPK), xe = min(xi + kPK, nx);
                const long long j0 = 3 * (((long long)zj * ny + yj) * nx + xs);
                const PPCell *p = P + (dz * kPW + dy) * kPW + (xi - xs + kPK);
                for (int xj = xs; xj < xe; ++xj, --p) {
                    const long long j = j0 + 3LL * (xj - xs);
                    const double cx = coef_value(coef, j), cy = coef_value(coef, j + 1),
                                 cz = coef_value(coef, j + 2);
                    const double xx = cx * p->vxx, yy = cy * p->vyy, zz = cz * p->vzz;
                    const double yxy = cy * p->vxy, xxz = cx * p->vxz, zyz = cz * p->vyz;
                    r0 += cix * xx;
                    r1 += cix * yxy;
                    r2 += ciy * yy;
                    r3 += ciz * xxz;
                    r4 += ciy * zyz;
                    r5 += ciz * zz;
                    g0 += 2.0 * xx + yxy + cz * p->vxz;
                    g1 += 2.0 * yy + zyz + cx * p->vxy;
                    g2 += 2.0 * zz + xxz + cy * p->vyz;
                }
            }
        }
    }
    double *a = acc[cl][q];
    a[0] = r0;
    a[1] = r1;
    a[2] = r2;
    a[3] = r3;
    a[4] = r4;
    a[5] = r5;
    a[6] = g0;
    a[7] = g1;
    a[8] = g2;
    __syncthreads();
    double v = 0.0;
    if (q == 0 && i < n) {
        double t[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            t[k] = acc[cl][0][k];
#pragma unroll
            for (int pl = 1; pl < kDcPlanes; ++pl)
                t[k] += acc[cl][pl][k];
        }
        if (SCALED) {
            grad[3 * i] = -w * t[6];
            grad[3 * i + 1] = -w * t[7];
            grad[3 * i + 2] = -w * t[8];
        } else if (grad) {
            grad[3 * i] = t[6];
            grad[3 * i + 1] = t[7];
            grad[3 * i + 2] = t[8];
        }
        v = t[0] + t[1] + t[2] + t[3] + t[4] + t[5];
    }
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int wv = 0; wv < kLB / 64; ++wv)
            sum += red[wv];
        part[blockIdx.x] = sum;
    }
}

// Sum of the block partials in a fixed order into *out.
__global__ void __launch_bounds__(kLB) k_sum_parts(const double *__restrict__ part, int n, double *__restrict__ out)
{
    __shared__ double red[kLB / 64];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += kLB)
        s += part[k];
    for (int o = 32; o > 0; o >>= 1)
        s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int wv = 0; wv < kLB / 64; ++wv)
            t += red[wv];
        *out = t;
    }
}

// cubic B-spline and its first two derivatives (mia/core/splinekernel/bspline.cc:328-375)
double bspline3_deg(double x, int deg)
{
    if (deg == 0) {
        x = std::fabs(x);
        if (x <= 1.0)
            return 2.0 / 3.0 - x * x * (1 - 0.5 * x);
        if (x < 2.0) {
            const double dx = 2.0 - x;
            return dx * dx * dx * (1.0 / 6.0);
        }
        return 0.0;
    }
    if (deg == 1) {
        const double ax = std::fabs(x);
        if (ax <= 1.0)
            return (1.5 * ax - 2) * x;
        if (ax <= 2.0)
            return (x > 0.0 ? -0.5 : 0.5) * (2 - ax) * (2 - ax);
        return 0.0;
    }
    x = std::fabs(x);
    if (x > 2.0)
        return 0.0;
    if (x > 1.0)
        return 2.0 - x;
    return 3.0 * x - 2.0;
}

// integral of B^(d1)(x - s1) B^(d2)(x - s2) over [0, L] by composite Simpson
// with 8 intervals per unit (mia/core/splinekernel.cc:287-315, mia/core/simpson.hh:36-52)
double integrate2(double s1, double s2, int d1, int d2, double L)
{
    const double radius = 2.0;
    double a = s1 - radius, b = s1 + radius;
    if (a < s2 - radius)
        a = s2 - radius;
    if (a < 0.0)
        a = 0.0;
    if (b > s2 + radius)
        b = s2 + radius;
    if (b > L)
        b = L;
    if (b <= a)
        return 0.0;
    const size_t intervals = size_t(8 * (b - a));
    auto f = [&](double x) { return bspline3_deg(x - s1, d1) * bspline3_deg(x - s2, d2); };
    const double dx = (b - a) / intervals;
    double sum = 0.5 * (f(a) + f(b));
    double x = a + dx;
    for (size_t k = 1; k < intervals; ++k, x += dx)
        sum += f(x);
    x = a + 0.5 * dx;
    for (size_t k = 0; k < intervals; ++k, x += dx)
        sum += 2 * f(x);
    return sum * dx / 3.0;
}

// P cells (ppmatrix.cc:178-262): tensor products of the 1D integrals scaled
// by powers of the knot spacing h = (csize - 1) / range
void make_pp_cells(const uint32_t cs[3], const float range[3], double wdiv, double wrot, std::vector<PPCell> &out)
{
    const float h1[3] = {(float)(cs[0] - 1) / range[0], (float)(cs[1] - 1) / range[1], (float)(cs[2] - 1) / range[2]};
    float h2[3], h3[3], h4[3];
    for (int a = 0; a < 3; ++a) {
        h2[a] = h1[a] * h1[a];
        h3[a] = h1[a] * h2[a];
        h4[a] = h1[a] * h3[a];
    }
    const double global_scale = 1.0 / (h1[0] * h1[1] * h1[2]);
    auto I = [](int delta, int d1, int d2) { return integrate2(kPK, delta + kPK, d1, d2, 2 * kPK); };
    const double wsum = wdiv + wrot;
    const double wdelta = 2 * (wdiv - wrot);
    out.assign(kPW * kPW * kPW, PPCell{0, 0, 0, 0, 0, 0});
    size_t k = 0;
    for (int z = -kPK; z <= kPK; ++z) {
        const double r00z = I(z, 0, 0), r01z = h1[2] * I(z, 0, 1), r11z = h2[2] * I(z, 1, 1);
        const double r12z = h3[2] * I(z, 1, 2), r22z = h4[2] * I(z, 2, 2);
        for (int y = -kPK; y <= kPK; ++y) {
            const double r00y = I(y, 0, 0), r01y = h1[1] * I(y, 0, 1), r10y = h1[1] * I(y, 1, 0);
            const double r11y = h2[1] * I(y, 1, 1), r12y = h3[1] * I(y, 1, 2), r21y = h3[1] * I(y, 2, 1);
            const double r22y = h4[1] * I(y, 2, 2);
            for (int x = -kPK; x <= kPK; ++x, ++k) {
                const double r00x = I(x, 0, 0), r01x = h1[0] * I(x, 0, 1), r10x = h1[0] * I(x, 1, 0);
                const double r11x = h2[0] * I(x, 1, 1), r21x = h3[0] * I(x, 2, 1), r22x = h4[0] * I(x, 2, 2);
                const double r200200 = r22x * r00y * r00z;
                const double r020020 = r00x * r22y * r00z;
                const double r002002 = r00x * r00y * r22z;
                const double r110110 = r11x * r11y * r00z;
                const double r101101 = r11x * r00y * r11z;
                const double r011011 = r00x * r11y * r11z;
                PPCell c{0, 0, 0, 0, 0, 0};
                c.vxx = global_scale * (r002002 * wrot + r011011 * 2.0 * wrot + r020020 * wrot + r101101 * wsum +
                                        r110110 * wsum + r200200 * wdiv);
                c.vyy = global_scale * (r002002 * wrot + r011011 * wsum + r020020 * wdiv + r101101 * 2 * wrot +
                                        r110110 * wsum + r200200 * wrot);
                c.vzz = global_scale * (r002002 * wdiv + r011011 * wsum + r020020 * wrot + r101101 * wsum +
                                        r110110 * wrot * 2 + r200200 * wrot);
                if (wdelta != 0.0) {
                    c.vxy = global_scale * wdelta * (r21x * r01y * r00z + r10x * r12y * r00z + r01x * r10y * r11z);
                    c.vxz = global_scale * wdelta * (r21x * r00y * r01z + r10x * r11y * r01z + r10x * r00y * r12z);
                    c.vyz = global_scale * wdelta * (r11x * r10y * r01z + r00x * r21y * r01z + r00x * r10y * r12z);
                }
                out[k] = c;
            }
        }
    }
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

namespace {
// The resampling of one image on `device`: input from host memory (h_in) or
// the device (d_in), output to host memory (h_out) or the device (d_out);
// synchronous.  Validates the sizes as mia_hip_scale_image documents.
int scale_impl(int device, int dim, const uint32_t *in_size, const float *h_in, const float *d_in_ext,
               const uint32_t *out_size, float *h_out, float *d_out_ext)
{
    if ((dim != 2 && dim != 3) || !in_size || !out_size || !(h_in || d_in_ext) || !(h_out || d_out_ext))
        return fail(MIA_HIP_ERR_ARG, "scale_image: bad arguments");
    int n_in[3], n_out[3];
    for (int a = 0; a < 3; ++a) {
        n_in[a] = a < dim ? (int)in_size[a] : 1;
        n_out[a] = a < dim ? (int)out_size[a] : 1;
        if (n_in[a] < 1 || n_out[a] < 1)
            return fail(MIA_HIP_ERR_SIZE, "scale_image: empty size");
        // a single input sample has no scale: only fill is defined
        if (n_in[a] > 1 && n_out[a] == 1)
            return fail(MIA_HIP_ERR_SIZE, "scale_image: cannot scale an axis to one sample");
    }
    HIP_TRY(hipSetDevice(device));
    const size_t vin = (size_t)n_in[0] * n_in[1] * n_in[2];
    const size_t vout = (size_t)n_out[0] * n_out[1] * n_out[2];
    // the device's shared stream and scratch (input, output, two double
    // volumes, the line coefficients and the binomial taps, 256-byte aligned)
    DeviceCache *dc = device_cache(device);
    if (!dc)
        return fail(MIA_HIP_ERR_ARG, "scale_image: device index out of range");
    std::lock_guard<std::mutex> lock(dc->m);
    hipStream_t s = nullptr;
    HIP_TRY(cache_stream(*dc, &s));
    if (n_in[0] == n_out[0] && n_in[1] == n_out[1] && n_in[2] == n_out[2]) {
        const size_t b = vin * sizeof(float);
        if (h_in && h_out)
            std::copy(h_in, h_in + vin, h_out);
        else
            HIP_TRY(hipMemcpyAsync(h_out ? (void *)h_out : (void *)d_out_ext, h_in ? (const void *)h_in : d_in_ext, b,
                                   h_in ? hipMemcpyHostToDevice
                                        : (h_out ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice),
                                   s));
        HIP_TRY(hipStreamSynchronize(s));
        return MIA_HIP_OK;
    }
    // passes over the axes that change size, x then y then z (scale.cc:44-101);
    // the buffers hold the largest volume any pass reads or writes
    std::vector<int> axes;
    for (int a = 0; a < dim; ++a)
        if (n_in[a] != n_out[a])
            axes.push_back(a);
    size_t vmax = vin;
    {
        size_t v[3] = {(size_t)n_in[0], (size_t)n_in[1], (size_t)n_in[2]};
        for (int a : axes) {
            v[a] = (size_t)n_out[a];
            vmax = std::max(vmax, v[0] * v[1] * v[2]);
        }
    }
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_in = d_in_ext ? 0 : al(vin * sizeof(float)), b_out = d_out_ext ? 0 : al(vout * sizeof(float));
    const size_t b_v = al(vmax * sizeof(double));
    const size_t need = b_in + b_out + 3 * b_v + al(256 * sizeof(double));
    if (dc->cap < need) {
        if (dc->scratch)
            HIP_TRY(hipFree(dc->scratch));
        dc->scratch = nullptr;
        dc->cap = 0;
        HIP_TRY(hipMalloc(&dc->scratch, need));
        dc->cap = need;
    }
    char *base = (char *)dc->scratch;
    const float *d_in = d_in_ext ? d_in_ext : (const float *)base;
    float *d_out = d_out_ext ? d_out_ext : (float *)(base + b_in);
    double *d_a = (double *)(base + b_in + b_out);
    double *d_b = (double *)(base + b_in + b_out + b_v);
    double *d_scr = (double *)(base + b_in + b_out + 2 * b_v);
    double *d_g = (double *)(base + b_in + b_out + 3 * b_v);
    if (!d_in_ext)
        HIP_TRY(hipMemcpyAsync((void *)d_in, h_in, vin * sizeof(float), hipMemcpyHostToDevice, s));
    int cur[3] = {n_in[0], n_in[1], n_in[2]};
    const float *src_f = d_in;
    const double *src_d = nullptr;
    std::vector<double> gauss;
    for (size_t ai = 0; ai < axes.size(); ++ai) {
        const int a = axes[ai];
        ScalePass P{};
        P.axis = a;
        P.n_in = n_in[a];
        P.n_out = n_out[a];
        const double sc = (double)(n_out[a] - 1) / (double)(n_in[a] - 1);
        P.F = 0;
        if (n_in[a] == 1) {
            P.strategy = SC_FILL;
        } else if (sc > 1.0) {
            P.strategy = SC_UP;
            P.dx = 1.0 / sc;
        } else {
            P.strategy = SC_DOWN;
            P.dx = 1.0 / sc;
            P.F = (int)std::floor(1.0 / sc + 0.5);
            if (P.F < 1 || P.F > 127) {
                (void)hipStreamSynchronize(s);
                return fail(MIA_HIP_ERR_SIZE, "scale_image: downscale factor out of range");
            }
            // binomial kernel of 2F+1 taps normalised by 2^(2F) (gauss.cc:25-55)
            const int K = 2 * P.F + 1;
            gauss.assign(K, 0.0);
            gauss[0] = 1.0;
            for (int i = 1; i < K; ++i)
                for (int j = i; j > 0; --j)
                    gauss[j] += gauss[j - 1];
            const double norm = std::pow(2.0, K - 1);
            for (double &g : gauss)
                g /= norm;
            HIP_TRY(hipMemcpyAsync(d_g, gauss.data(), K * sizeof(double), hipMemcpyHostToDevice, s));
            // the taps are read by the pass below before `gauss` changes
            HIP_TRY(hipStreamSynchronize(s));
        }
        // line grid of this pass
        long long nlines;
        if (a == 0)
            nlines = (long long)cur[1] * cur[2];
        else if (a == 1)
            nlines = (long long)cur[0] * cur[2];
        else
            nlines = (long long)cur[0] * cur[1];
        P.nx = cur[0];
        P.ny = cur[1];
        const bool last = ai + 1 == axes.size();
        double *dst_d = last ? nullptr : (src_d == d_a ? d_b : d_a);
        float *dst_f = last ? d_out : nullptr;
        const unsigned blocks = (unsigned)((nlines + kLB - 1) / kLB);
        hipLaunchKernelGGL(k_scale_lines, dim3(blocks), dim3(kLB), 0, s, src_f, src_d, dst_d, dst_f, d_scr, d_g, P,
                           nlines);
        LAUNCH_CHECK("scale");
        cur[a] = n_out[a];
        src_f = nullptr;
        src_d = dst_d;
    }
    if (h_out)
        HIP_TRY(hipMemcpyAsync(h_out, d_out, vout * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return MIA_HIP_OK;
}
} // namespace

extern "C" int mia_hip_scale_image(int device, int dim, const uint32_t *in_size, const float *in,
                                   const uint32_t *out_size, float *out)
{
    if (!in || !out)
        return fail(MIA_HIP_ERR_ARG, "scale_image: bad arguments");
    return scale_impl(device, dim, in_size, in, nullptr, out_size, out, nullptr);
}

// An image held in device memory (float, x fastest).
struct mia_hip_image {
    int device = 0, dim = 0;
    uint32_t size[3] = {1, 1, 1};
    float *d = nullptr;
    size_t count() const { return (size_t)size[0] * size[1] * size[2]; }
};

extern "C" int mia_hip_image_destroy(mia_hip_image *img)
{
    if (!img)
        return MIA_HIP_OK;
    (void)hipSetDevice(img->device);
    if (img->d)
        (void)hipFree(img->d);
    delete img;
    return MIA_HIP_OK;
}

namespace {
int image_alloc(int device, int dim, const uint32_t *size, mia_hip_image **out)
{
    if (!out || !size || (dim != 2 && dim != 3))
        return fail(MIA_HIP_ERR_ARG, "image: bad arguments");
    *out = nullptr;
    std::unique_ptr<mia_hip_image, int (*)(mia_hip_image *)> img(new mia_hip_image, mia_hip_image_destroy);
    img->device = device;
    img->dim = dim;
    for (int a = 0; a < dim; ++a) {
        if (size[a] < 1)
            return fail(MIA_HIP_ERR_SIZE, "image: empty size");
        img->size[a] = size[a];
    }
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipMalloc(&img->d, img->count() * sizeof(float)));
    *out = img.release();
    return MIA_HIP_OK;
}
} // namespace

extern "C" int mia_hip_image_create(int device, int dim, const uint32_t *size, const float *data, mia_hip_image **out)
{
    if (!data)
        return fail(MIA_HIP_ERR_ARG, "image_create: null data");
    mia_hip_image *img = nullptr;
    int st = image_alloc(device, dim, size, &img);
    if (st)
        return st;
    DeviceCache *dc = device_cache(device);
    if (!dc) {
        mia_hip_image_destroy(img);
        return fail(MIA_HIP_ERR_ARG, "image_create: device index out of range");
    }
    {
        std::lock_guard<std::mutex> lock(dc->m);
        hipStream_t s = nullptr;
        hipError_t e = cache_stream(*dc, &s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(img->d, data, img->count() * sizeof(float), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            mia_hip_image_destroy(img);
            return fail(MIA_HIP_ERR_DEVICE, std::string("image_create: ") + hipGetErrorString(e));
        }
    }
    *out = img;
    return MIA_HIP_OK;
}

extern "C" int mia_hip_image_scale(const mia_hip_image *in, const uint32_t *out_size, mia_hip_image **out)
{
    if (!in || !out_size || !out)
        return fail(MIA_HIP_ERR_ARG, "image_scale: null argument");
    mia_hip_image *img = nullptr;
    int st = image_alloc(in->device, in->dim, out_size, &img);
    if (st)
        return st;
    if ((st = scale_impl(in->device, in->dim, in->size, nullptr, in->d, out_size, nullptr, img->d))) {
        mia_hip_image_destroy(img);
        return st;
    }
    *out = img;
    return MIA_HIP_OK;
}

extern "C" int mia_hip_image_read(const mia_hip_image *img, float *out)
{
    if (!img || !out)
        return fail(MIA_HIP_ERR_ARG, "image_read: null argument");
    HIP_TRY(hipSetDevice(img->device));
    HIP_TRY(hipMemcpy(out, img->d, img->count() * sizeof(float), hipMemcpyDeviceToHost));
    return MIA_HIP_OK;
}

extern "C" int mia_hip_set_images_image(mia_hip_ctx *ctx, int inst, const mia_hip_image *moving,
                                        const mia_hip_image *fixed)
{
    if (!ctx || !moving || !fixed)
        return fail(MIA_HIP_ERR_ARG, "set_images_image: null argument");
    uint32_t size[3];
    int dim = 0, device = -1;
    int st = mia_hip_rt::instance_geometry(ctx, inst, &dim, size, &device);
    if (st)
        return st;
    for (const mia_hip_image *img : {moving, fixed}) {
        if (img->device != device)
            return fail(MIA_HIP_ERR_ARG, "set_images_image: image on another device");
        if (img->dim != dim || img->size[0] != size[0] || img->size[1] != size[1] || img->size[2] != size[2])
            return fail(MIA_HIP_ERR_SIZE, "set_images_image: image size differs from the instance's");
    }
    return mia_hip_set_images_device(ctx, inst, moving->d, fixed->d);
}

struct mia_hip_divcurl {
    int device = 0;
    uint32_t cs[3] = {0, 0, 0};
    long long n = 0;
    int nblocks = 0;
    hipStream_t stream = nullptr;
    PPCell *d_cells = nullptr;
    float *d_coef = nullptr;
    double *d_grad = nullptr, *d_part = nullptr, *d_value = nullptr;
    std::vector<PPCell> cells;
};

extern "C" int mia_hip_divcurl_destroy(mia_hip_divcurl *p)
{
    if (!p)
        return MIA_HIP_OK;
    (void)hipSetDevice(p->device);
    (void)hipFree(p->d_cells);
    (void)hipFree(p->d_coef);
    (void)hipFree(p->d_grad);
    (void)hipFree(p->d_part);
    (void)hipFree(p->d_value);
    // (p->stream is the device's shared stream)
    delete p;
    return MIA_HIP_OK;
}

extern "C" int mia_hip_divcurl_create(int device, const uint32_t *csize, const float *range, double wdiv,
                                      double wcurl, mia_hip_divcurl **out)
{
    if (!csize || !range || !out)
        return fail(MIA_HIP_ERR_ARG, "divcurl_create: null argument");
    *out = nullptr;
    for (int a = 0; a < 3; ++a) {
        if (csize[a] < 1)
            return fail(MIA_HIP_ERR_SIZE, "divcurl_create: empty coefficient grid");
        if (!(range[a] > 0.0f))
            return fail(MIA_HIP_ERR_ARG, "divcurl_create: range must be positive");
    }
    if (!(wdiv >= 0.0) || !(wcurl >= 0.0))
        return fail(MIA_HIP_ERR_ARG, "divcurl_create: div and curl weights must be non-negative");
    std::unique_ptr<mia_hip_divcurl, int (*)(mia_hip_divcurl *)> p(new mia_hip_divcurl, mia_hip_divcurl_destroy);
    p->device = device;
    std::copy(csize, csize + 3, p->cs);
    p->n = (long long)csize[0] * csize[1] * csize[2];
    p->nblocks = (int)((p->n + kDcCoef - 1) / kDcCoef);
    make_pp_cells(csize, range, wdiv, wcurl, p->cells);
    HIP_TRY(hipSetDevice(device));
    {
        DeviceCache *dc = device_cache(device);
        if (!dc)
            return fail(MIA_HIP_ERR_ARG, "divcurl_create: device index out of range");
        std::lock_guard<std::mutex> lock(dc->m);
        HIP_TRY(cache_stream(*dc, &p->stream));
    }
    HIP_TRY(hipMalloc(&p->d_cells, p->cells.size() * sizeof(PPCell)));
    HIP_TRY(hipMalloc(&p->d_coef, 3 * p->n * sizeof(float)));
    HIP_TRY(hipMalloc(&p->d_grad, 3 * p->n * sizeof(double)));
    HIP_TRY(hipMalloc(&p->d_part, p->nblocks * sizeof(double)));
    HIP_TRY(hipMalloc(&p->d_value, sizeof(double)));
    HIP_TRY(hipMemcpyAsync(p->d_cells, p->cells.data(), p->cells.size() * sizeof(PPCell), hipMemcpyHostToDevice,
                           p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    *out = p.release();
    return MIA_HIP_OK;
}

extern "C" int mia_hip_divcurl_evaluate(mia_hip_divcurl *p, const float *coeffs, double *value, double *gradient)
{
    if (!p || !coeffs || !value)
        return fail(MIA_HIP_ERR_ARG, "divcurl_evaluate: null argument");
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipMemcpyAsync(p->d_coef, coeffs, 3 * p->n * sizeof(float), hipMemcpyHostToDevice, p->stream));
    hipLaunchKernelGGL((k_divcurl<float, false>), dim3(p->nblocks), dim3(kLB), 0, p->stream,
                       (const float *)p->d_coef, (const PPCell *)p->d_cells, (int)p->cs[0], (int)p->cs[1],
                       (int)p->cs[2], 0.0, gradient ? p->d_grad : nullptr, p->d_part);
    LAUNCH_CHECK("divcurl");
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kLB), 0, p->stream, (const double *)p->d_part, p->nblocks,
                       p->d_value);
    LAUNCH_CHECK("divcurl sum");
    HIP_TRY(hipMemcpyAsync(value, p->d_value, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    if (gradient)
        HIP_TRY(hipMemcpyAsync(gradient, p->d_grad, 3 * p->n * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return MIA_HIP_OK;
}

namespace mia_hip_rt {

int divcurl_grid(const mia_hip_divcurl *p, uint32_t cs[3])
{
    if (!p)
        return fail(MIA_HIP_ERR_ARG, "null divcurl penalty");
    std::copy(p->cs, p->cs + 3, cs);
    return MIA_HIP_OK;
}

int divcurl_enqueue(mia_hip_divcurl *p, hipStream_t s, const double *params, double weight, double *wgrad,
                    double *value)
{
    hipLaunchKernelGGL((k_divcurl<double, true>), dim3(p->nblocks), dim3(kLB), 0, s, params,
                       (const PPCell *)p->d_cells, (int)p->cs[0], (int)p->cs[1], (int)p->cs[2], weight, wgrad,
                       p->d_part);
    LAUNCH_CHECK("divcurl");
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kLB), 0, s, (const double *)p->d_part, p->nblocks, value);
    LAUNCH_CHECK("divcurl sum");
    return MIA_HIP_OK;
}

} // namespace mia_hip_rt

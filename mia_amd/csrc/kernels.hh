// Device kernels of the B-spline SSD cost + gradient path for gfx950.
//
// Data flow of one evaluation (all instances of a batch in each launch):
//   k_coef      params (double, interleaved) -> coefficient vectors (float)
//   k_warp      displacement u(x) (separable through an LDS coefficient
//               window) and warp W(x) = M~(x - u(x)) of the prefiltered moving
//               image with mirror indices and clamp              [reference mode]
//   k_stage1    per (z slice, y knot band): SSD residual, force, FP64 SSD
//               partial with a wave64 shuffle reduction, contraction of the
//               force with the y weights (registers) and the x weights (LDS)
//               -> per-band partial sums on the coefficient grid; in exact mode
//               it also does the warp and the analytic image gradient inline
//   k_finalize  per instance: fixed-order FP64 sum of the SSD partials, cost,
//               gradient scale (norm / autothresh / weight)
//   k_reduce    per coefficient: fixed-order FP64 sum over the four y bands
//               and the z support -> gradient (no atomics: deterministic)
//
// Reference algorithm: mia/3d/transform/spline.cc:541-653 (init_grid),
// mia/3d/interpolator.cxx:250-287 (interpolation), mia/3d/image.cc:347-477
// (gradient), mia/template/ssd.hh:54-181 (value and force),
// mia/3d/transform/spline.cc:731-791 (translate); 2D counterparts in
// mia/2d/transform/spline.cc, mia/2d/interpolator.cxx, mia/2d/image.cc.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

namespace mia_hip {

constexpr int kBlock = 256;
constexpr int kWarpTileX = 64;   // warp kernel tile: 64 (x) by 16 (y) voxels
constexpr int kWarpTileY = 16;
constexpr int kMaxRow = 512;     // longest x row the stage-1 kernel keeps in LDS

// Per-instance description, identical on host and device.
struct InstDesc {
    int nx, ny, nz;           // image size (nz = 1 in 2D)
    int cx, cy, cz;           // coefficient grid (cz = 1 in 2D)
    int ncy, ncz;             // number of y / z knot bands (max start index + 1)
    long long img_off;        // offset of the instance's images (floats)
    long long par_off;        // offset of its parameters / gradient (doubles)
    long long coef_off;       // offset of its coefficient vectors (float4 / float2)
    long long part_off;       // offset of its stage-1 partials (floats)
    int tab_x, tab_y, tab_z;  // offsets into the axis tables (one entry per voxel)
    int cell_x, cell_y, cell_z; // offsets into the band tables (csize + 1 entries)
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Block -> instance by binary search over the per-instance block prefix.
__device__ __forceinline__ int find_inst(const int *__restrict__ prefix, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// The hardware hands workgroup i to XCD (i mod 8).  Remap so that each XCD
// gets a contiguous range of logical tiles: neighbouring tiles share their
// moving-image halo and coefficient windows in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int x = bid & 7, k = bid >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// floor after rounding to a multiple of 2^-16 (mia/core/splinekernel.cc:164-192).
// v * 65536 and the scaling back are exact in float.
__device__ __forceinline__ int fastfloor16(float v)
{
    return (int)floorf(rintf(v * 65536.0f) * (1.0f / 65536.0f));
}

// mirror boundary index map (mia/core/splinebc/bc.cc:52-64)
__device__ __forceinline__ int mirror_idx(int i, int n)
{
    if (n == 1)
        return 0;
    const int n2 = 2 * n - 2;
    int idx = i < 0 ? -i : i;
    if (idx >= n2)
        idx = idx % n2;
    if (idx >= n)
        idx = n2 - idx;
    return idx;
}

// Image-kernel start index and weights (and derivative weights) for degree
// 3 (mia/core/splinekernel/bspline.cc:258-285) or 1 (:90-107).
template <int DEG, bool GRAD>
__device__ __forceinline__ int kweights(float p, float *w, float *dw)
{
    // Keep the float -> int conversion and the mirror arithmetic in range for
    // any input (NaN and +-inf map to the bounds; fmaxf drops a NaN operand).
    p = fminf(fmaxf(p, -1.0e6f), 1.0e6f);
    const int fl = fastfloor16(p);
    const float f = p - (float)fl;
    if constexpr (DEG == 3) {
        const float fm1 = 1.0f - f;
        const float s = 1.0f / 6.0f;
        w[3] = s * f * f * f;
        w[0] = s * fm1 * fm1 * fm1;
        w[2] = f - 2.0f * w[3] + w[0];
        w[1] = fm1 - 2.0f * w[0] + w[3];
        if constexpr (GRAD) {
            dw[3] = 0.5f * f * f;
            dw[0] = f - 0.5f - dw[3];
            dw[2] = 1.0f + dw[0] - 2.0f * dw[3];
            dw[1] = -dw[0] - dw[2] - dw[3];
        }
        return fl - 1;
    } else {
        w[0] = 1.0f - f;
        w[1] = f;
        if constexpr (GRAD) {
            dw[0] = -1.0f;
            dw[1] = 1.0f;
        }
        return fl;
    }
}

template <int NS>
__device__ __forceinline__ void axis_index(int s, int n, int *idx)
{
    if (s >= 0 && s + NS - 1 < n) {
#pragma unroll
        for (int k = 0; k < NS; ++k)
            idx[k] = s + k;
    } else {
#pragma unroll
        for (int k = 0; k < NS; ++k)
            idx[k] = mirror_idx(s + k, n);
    }
}

// Interpolation of the prefiltered image at p with clamp to [vmin, vmax]
// (mia/3d/interpolator.cxx:250-287, z->y->x contraction as add_3d :187-210).
template <int DEG, bool GRAD>
__device__ __forceinline__ float interp3(const float *__restrict__ M, int nx, int ny, int nz, float px,
                                         float py, float pz, float vmin, float vmax, float *g)
{
    constexpr int NS = DEG == 3 ? 4 : 2;
    float wx[NS], wy[NS], wz[NS], dx[NS], dy[NS], dz[NS];
    const int sx = kweights<DEG, GRAD>(px, wx, dx);
    const int sy = kweights<DEG, GRAD>(py, wy, dy);
    const int sz = kweights<DEG, GRAD>(pz, wz, dz);
    int ix[NS], iy[NS], iz[NS];
    axis_index<NS>(sx, nx, ix);
    axis_index<NS>(sy, ny, iy);
    axis_index<NS>(sz, nz, iz);
    float v = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        const float *slab = M + (size_t)iz[a] * nx * ny;
        float vy = 0.0f, gyx = 0.0f, gyy = 0.0f;
#pragma unroll
        for (int b = 0; b < NS; ++b) {
            const float *row = slab + (size_t)iy[b] * nx;
            float vx = 0.0f, gxx = 0.0f;
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                const float c = row[ix[k]];
                vx = fmaf(wx[k], c, vx);
                if constexpr (GRAD)
                    gxx = fmaf(dx[k], c, gxx);
            }
            vy = fmaf(wy[b], vx, vy);
            if constexpr (GRAD) {
                gyx = fmaf(wy[b], gxx, gyx);
                gyy = fmaf(dy[b], vx, gyy);
            }
        }
        v = fmaf(wz[a], vy, v);
        if constexpr (GRAD) {
            gx = fmaf(wz[a], gyx, gx);
            gy = fmaf(wz[a], gyy, gy);
            gz = fmaf(dz[a], vy, gz);
        }
    }
    if constexpr (GRAD) {
        const bool clamped = v < vmin || v > vmax;
        g[0] = clamped ? 0.0f : gx;
        g[1] = clamped ? 0.0f : gy;
        g[2] = clamped ? 0.0f : gz;
    }
    return v >= vmin ? (v <= vmax ? v : vmax) : vmin;
}

// 2D interpolation (mia/2d/interpolator.cxx:277-311)
template <int DEG, bool GRAD>
__device__ __forceinline__ float interp2(const float *__restrict__ M, int nx, int ny, float px, float py,
                                         float vmin, float vmax, float *g)
{
    constexpr int NS = DEG == 3 ? 4 : 2;
    float wx[NS], wy[NS], dx[NS], dy[NS];
    const int sx = kweights<DEG, GRAD>(px, wx, dx);
    const int sy = kweights<DEG, GRAD>(py, wy, dy);
    int ix[NS], iy[NS];
    axis_index<NS>(sx, nx, ix);
    axis_index<NS>(sy, ny, iy);
    float v = 0.0f, gx = 0.0f, gy = 0.0f;
#pragma unroll
    for (int b = 0; b < NS; ++b) {
        const float *row = M + (size_t)iy[b] * nx;
        float vx = 0.0f, gxx = 0.0f;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const float c = row[ix[k]];
            vx = fmaf(wx[k], c, vx);
            if constexpr (GRAD)
                gxx = fmaf(dx[k], c, gxx);
        }
        v = fmaf(wy[b], vx, v);
        if constexpr (GRAD) {
            gx = fmaf(wy[b], gxx, gx);
            gy = fmaf(dy[b], vx, gy);
        }
    }
    if constexpr (GRAD) {
        const bool clamped = v < vmin || v > vmax;
        g[0] = clamped ? 0.0f : gx;
        g[1] = clamped ? 0.0f : gy;
    }
    return v >= vmin ? (v <= vmax ? v : vmax) : vmin;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ unsigned wave_sum_u(unsigned v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// Deterministic block sum (fixed shuffle tree per wave, then waves in order).
__device__ __forceinline__ double block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
        red[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k)
            s += red[k];
    return s;
}

__device__ __forceinline__ unsigned block_sum_u(unsigned v, unsigned *red)
{
    v = wave_sum_u(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
        red[wid] = v;
    __syncthreads();
    unsigned s = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k)
            s += red[k];
    return s;
}

// ---------------------------------------------------------------------------
// set-up kernels (once per level): clamp range and prefilter
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_minmax_partial(const float *__restrict__ img, long long n,
                                                           float2 *__restrict__ part)
{
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
        const float v = img[i];
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, off, 64));
        mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    }
    __shared__ float2 red[kBlock / 64];
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = make_float2(mn, mx);
    __syncthreads();
    if (threadIdx.x == 0) {
        float2 r = red[0];
        for (int k = 1; k < kBlock / 64; ++k) {
            r.x = fminf(r.x, red[k].x);
            r.y = fmaxf(r.y, red[k].y);
        }
        part[blockIdx.x] = r;
    }
}

// clamp range [min(0, min), max] (mia/3d/interpolator.cxx:108-111)
__global__ void k_minmax_final(const float2 *__restrict__ part, int n, float2 *__restrict__ out)
{
    if (threadIdx.x == 0) {
        float2 r = part[0];
        for (int k = 1; k < n; ++k) {
            r.x = fminf(r.x, part[k].x);
            r.y = fmaxf(r.y, part[k].y);
        }
        if (0.0f < r.x)
            r.x = 0.0f;
        *out = r;
    }
}

// One thread per line of the cubic B-spline prefilter with mirror boundaries
// (mia/core/boundary_conditions.cc:62-103, mia/core/splinebc/bc.cc:67-88), in
// double within the line, stored as float (the reference's coefficient type).
// Line l starts at base(l) and steps by `stride`; mode 0: x lines
// (base = l*nx), 1: y lines (x = l % nx, z = l / nx), 2: z lines (base = l).
__global__ void __launch_bounds__(kBlock) k_prefilter_lines(float *__restrict__ img, long long nlines, int n,
                                                            long long stride, int mode, int nx, int ny,
                                                            double *__restrict__ scratch)
{
    const long long l = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (l >= nlines || n < 2)
        return;
    long long base;
    if (mode == 0)
        base = l * nx;
    else if (mode == 1)
        base = (l / nx) * (long long)nx * ny + (l % nx);
    else
        base = l;
    const double p = 1.7320508075688772 - 2.0; // sqrt(3) - 2, bspline.cc:248
    const double lambda = 2.0 - p - 1.0 / p;
    double zn = p;
    const double iz = 1.0 / p;
    double z2n = pow(p, (double)(n - 1));
    double sum = lambda * (double)img[base] + z2n * lambda * (double)img[base + (long long)(n - 1) * stride];
    z2n *= z2n * iz;
    for (int k = 1; k < n - 1; ++k) {
        sum += (zn + z2n) * lambda * (double)img[base + (long long)k * stride];
        zn *= p;
        z2n *= iz;
    }
    double c = sum / (1.0 - zn * zn);
    scratch[l] = c;
    for (int k = 1; k < n; ++k) {
        c = lambda * (double)img[base + (long long)k * stride] + p * c;
        scratch[(long long)k * nlines + l] = c;
    }
    double a = (p / (p * p - 1.0)) * (p * scratch[(long long)(n - 2) * nlines + l] + c);
    img[base + (long long)(n - 1) * stride] = (float)a;
    for (int k = n - 2; k >= 0; --k) {
        a = p * (a - scratch[(long long)k * nlines + l]);
        img[base + (long long)k * stride] = (float)a;
    }
}

// ---------------------------------------------------------------------------
// per-evaluation kernels
// ---------------------------------------------------------------------------

// double parameters -> float coefficient vectors (the reference stores
// C3DFVector / C2DFVector coefficients, spline.cc:344-358)
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_coef(const double *__restrict__ params, void *__restrict__ coef,
                                                 long long first, long long count)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= count)
        return;
    const long long k = first + i;
    if constexpr (DIM == 3) {
        const double *p = params + 3 * k;
        ((float4 *)coef)[k] = make_float4((float)p[0], (float)p[1], (float)p[2], 0.0f);
    } else {
        const double *p = params + 2 * k;
        ((float2 *)coef)[k] = make_float2((float)p[0], (float)p[1]);
    }
}

// Warp kernel, 3D.  Tile = 64 x 16 voxels of one z slice; 256 threads, four
// rows each.  The coefficients are first contracted along z into an LDS
// window covering the tile's x/y knot range (zero padded beyond the grid),
// then each voxel contracts 4x4 window entries with its y/x table weights.
template <int DEG>
__global__ void __launch_bounds__(kBlock)
    k_warp3d(const InstDesc *__restrict__ inst, int n_inst, const int *__restrict__ prefix, int boff, int nblocks,
             const float4 *__restrict__ coef, const float4 *__restrict__ tw, const int *__restrict__ ts,
             const float *__restrict__ M, const float2 *__restrict__ clamp, float *__restrict__ W)
{
    constexpr int WX = kWarpTileX + 4 + 4, WY = kWarpTileY + 4 + 4;
    __shared__ float4 win[WX * WY];
    const int b = boff + xcd_remap(blockIdx.x, nblocks);
    const int i = find_inst(prefix, n_inst, b);
    const InstDesc d = inst[i];
    const int lb = b - prefix[i];
    const int tiles_x = (d.nx + kWarpTileX - 1) / kWarpTileX;
    const int tiles_y = (d.ny + kWarpTileY - 1) / kWarpTileY;
    const int tx = lb % tiles_x;
    const int rest = lb / tiles_x;
    const int ty = rest % tiles_y;
    const int z = rest / tiles_y;
    const int x0 = tx * kWarpTileX, y0 = ty * kWarpTileY;
    const int x1 = min(x0 + kWarpTileX, d.nx) - 1, y1 = min(y0 + kWarpTileY, d.ny) - 1;
    const int kx0 = ts[d.tab_x + x0], ky0 = ts[d.tab_y + y0];
    const int wxn = ts[d.tab_x + x1] + 4 - kx0;
    const int wyn = ts[d.tab_y + y1] + 4 - ky0;
    const int sz = ts[d.tab_z + z];
    const float4 wz = tw[d.tab_z + z];
    const float wzs[4] = {wz.x, wz.y, wz.z, wz.w};
    const float4 *cbase = coef + d.coef_off;
    for (int e = threadIdx.x; e < wxn * wyn; e += kBlock) {
        const int ky = ky0 + e / wxn, kx = kx0 + e % wxn;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        if (kx < d.cx && ky < d.cy) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int kz = sz + c;
                if (kz < d.cz) {
                    const float4 q = cbase[((long long)kz * d.cy + ky) * d.cx + kx];
                    ax = fmaf(wzs[c], q.x, ax);
                    ay = fmaf(wzs[c], q.y, ay);
                    az = fmaf(wzs[c], q.z, az);
                }
            }
        }
        win[e] = make_float4(ax, ay, az, 0.0f);
    }
    __syncthreads();
    const int lx = threadIdx.x & 63;
    const int x = x0 + lx;
    if (x >= d.nx)
        return;
    const int sx = ts[d.tab_x + x] - kx0;
    const float4 wxv = tw[d.tab_x + x];
    const float wxs[4] = {wxv.x, wxv.y, wxv.z, wxv.w};
    const float2 cl = clamp[i];
    const float *Mi = M + d.img_off;
    float *Wi = W + d.img_off;
    for (int ly = threadIdx.x >> 6; ly < kWarpTileY; ly += kBlock / 64) {
        const int y = y0 + ly;
        if (y >= d.ny)
            break;
        const int sy = ts[d.tab_y + y] - ky0;
        const float4 wyv = tw[d.tab_y + y];
        const float wys[4] = {wyv.x, wyv.y, wyv.z, wyv.w};
        float ux = 0.0f, uy = 0.0f, uz = 0.0f;
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            float hx = 0.0f, hy = 0.0f, hz = 0.0f;
            const float4 *wr = win + (sy + bb) * wxn + sx;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const float4 q = wr[a];
                hx = fmaf(wxs[a], q.x, hx);
                hy = fmaf(wxs[a], q.y, hy);
                hz = fmaf(wxs[a], q.z, hz);
            }
            ux = fmaf(wys[bb], hx, ux);
            uy = fmaf(wys[bb], hy, uy);
            uz = fmaf(wys[bb], hz, uz);
        }
        const float v = interp3<DEG, false>(Mi, d.nx, d.ny, d.nz, (float)x - ux, (float)y - uy, (float)z - uz,
                                            cl.x, cl.y, nullptr);
        Wi[((long long)z * d.ny + y) * d.nx + x] = v;
    }
}

// Warp kernel, 2D (same tiling, no z contraction).
template <int DEG>
__global__ void __launch_bounds__(kBlock)
    k_warp2d(const InstDesc *__restrict__ inst, int n_inst, const int *__restrict__ prefix, int boff, int nblocks,
             const float2 *__restrict__ coef, const float4 *__restrict__ tw, const int *__restrict__ ts,
             const float *__restrict__ M, const float2 *__restrict__ clamp, float *__restrict__ W)
{
    constexpr int WX = kWarpTileX + 4 + 4, WY = kWarpTileY + 4 + 4;
    __shared__ float2 win[WX * WY];
    const int b = boff + xcd_remap(blockIdx.x, nblocks);
    const int i = find_inst(prefix, n_inst, b);
    const InstDesc d = inst[i];
    const int lb = b - prefix[i];
    const int tiles_x = (d.nx + kWarpTileX - 1) / kWarpTileX;
    const int tx = lb % tiles_x;
    const int ty = lb / tiles_x;
    const int x0 = tx * kWarpTileX, y0 = ty * kWarpTileY;
    const int x1 = min(x0 + kWarpTileX, d.nx) - 1, y1 = min(y0 + kWarpTileY, d.ny) - 1;
    const int kx0 = ts[d.tab_x + x0], ky0 = ts[d.tab_y + y0];
    const int wxn = ts[d.tab_x + x1] + 4 - kx0;
    const int wyn = ts[d.tab_y + y1] + 4 - ky0;
    const float2 *cbase = coef + d.coef_off;
    for (int e = threadIdx.x; e < wxn * wyn; e += kBlock) {
        const int ky = ky0 + e / wxn, kx = kx0 + e % wxn;
        win[e] = (kx < d.cx && ky < d.cy) ? cbase[(long long)ky * d.cx + kx] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    const int lx = threadIdx.x & 63;
    const int x = x0 + lx;
    if (x >= d.nx)
        return;
    const int sx = ts[d.tab_x + x] - kx0;
    const float4 wxv = tw[d.tab_x + x];
    const float wxs[4] = {wxv.x, wxv.y, wxv.z, wxv.w};
    const float2 cl = clamp[i];
    const float *Mi = M + d.img_off;
    float *Wi = W + d.img_off;
    for (int ly = threadIdx.x >> 6; ly < kWarpTileY; ly += kBlock / 64) {
        const int y = y0 + ly;
        if (y >= d.ny)
            break;
        const int sy = ts[d.tab_y + y] - ky0;
        const float4 wyv = tw[d.tab_y + y];
        const float wys[4] = {wyv.x, wyv.y, wyv.z, wyv.w};
        float ux = 0.0f, uy = 0.0f;
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            float hx = 0.0f, hy = 0.0f;
            const float2 *wr = win + (sy + bb) * wxn + sx;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const float2 q = wr[a];
                hx = fmaf(wxs[a], q.x, hx);
                hy = fmaf(wxs[a], q.y, hy);
            }
            ux = fmaf(wys[bb], hx, ux);
            uy = fmaf(wys[bb], hy, uy);
        }
        Wi[(long long)y * d.nx + x] =
            interp2<DEG, false>(Mi, d.nx, d.ny, (float)x - ux, (float)y - uy, cl.x, cl.y, nullptr);
    }
}

// Stage 1, 3D: block = (z slice, y knot band).  Each thread owns x columns;
// for every row j of the band it forms the residual d = W - R, the force
// f = grad * d (reference mode: central differences of W with zero border
// components, mia/3d/image.cc:347-477; exact mode: analytic gradient of the
// interpolated moving image at x - u(x)), accumulates d^2 in FP64 and
// f * wy_t(j) for the band's four y coefficients in registers.  The rows are
// then contracted along x with the x table weights through LDS and written as
// partial sums P[z][band][t][kx][c] (t = ky - band).
template <int DEG, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_stage1_3d(const InstDesc *__restrict__ inst, int n_inst, const int *__restrict__ prefix, int nblocks,
                const float4 *__restrict__ coef, const float4 *__restrict__ tw, const int *__restrict__ ts,
                const int *__restrict__ cells, const float *__restrict__ M, const float2 *__restrict__ clamp,
                const float *__restrict__ W, const float *__restrict__ R, float thr, float *__restrict__ P,
                double *__restrict__ ssd_part, unsigned *__restrict__ cnt_part)
{
    extern __shared__ float smem[];
    __shared__ double red[kBlock / 64];
    __shared__ unsigned redu[kBlock / 64];
    const int b = xcd_remap(blockIdx.x, nblocks);
    const int i = find_inst(prefix, n_inst, b);
    const InstDesc d = inst[i];
    const int lb = b - prefix[i];
    const int band = lb % d.ncy;
    const int z = lb / d.ncy;
    const int j0 = cells[d.cell_y + band], j1 = cells[d.cell_y + band + 1];
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    float *acc_s = smem;                          // nx * 12 floats
    float4 *win = (float4 *)(smem + 12 * nx);     // exact mode: 4 x (cx + 4) window
    const int wxn = d.cx + 4;
    if constexpr (MODE == 1) {
        const int sz = ts[d.tab_z + z];
        const float4 wz = tw[d.tab_z + z];
        const float wzs[4] = {wz.x, wz.y, wz.z, wz.w};
        const float4 *cbase = coef + d.coef_off;
        for (int e = threadIdx.x; e < 4 * wxn; e += kBlock) {
            const int ky = band + e / wxn, kx = e % wxn;
            float ax = 0.0f, ay = 0.0f, az = 0.0f;
            if (kx < d.cx && ky < d.cy) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int kz = sz + c;
                    if (kz < d.cz) {
                        const float4 q = cbase[((long long)kz * d.cy + ky) * d.cx + kx];
                        ax = fmaf(wzs[c], q.x, ax);
                        ay = fmaf(wzs[c], q.y, ay);
                        az = fmaf(wzs[c], q.z, az);
                    }
                }
            }
            win[e] = make_float4(ax, ay, az, 0.0f);
        }
        __syncthreads();
    }
    const long long plane = (long long)nx * ny;
    const float *Wi = W + d.img_off;
    const float *Ri = R + d.img_off;
    const float *Mi = M + d.img_off;
    const float2 cl = clamp[i];
    double ssd = 0.0;
    unsigned cnt = 0;
    for (int x = threadIdx.x; x < nx; x += kBlock) {
        float acc[12];
#pragma unroll
        for (int k = 0; k < 12; ++k)
            acc[k] = 0.0f;
        const int sx = ts[d.tab_x + x];
        const float4 wxv = tw[d.tab_x + x];
        const float wxs[4] = {wxv.x, wxv.y, wxv.z, wxv.w};
        for (int j = j0; j < j1; ++j) {
            const long long idx = (long long)z * plane + (long long)j * nx + x;
            const float4 wyv = tw[d.tab_y + j];
            const float wys[4] = {wyv.x, wyv.y, wyv.z, wyv.w};
            float w, gx, gy, gz;
            if constexpr (MODE == 0) {
                // neighbour indices stay inside the image on the border
                const bool ix = x > 0 && x < nx - 1, iy = j > 0 && j < ny - 1, iz = z > 0 && z < nz - 1;
                const long long xm = ix ? idx - 1 : idx, xp = ix ? idx + 1 : idx;
                const long long ym = iy ? idx - nx : idx, yp = iy ? idx + nx : idx;
                const long long zm = iz ? idx - plane : idx, zp = iz ? idx + plane : idx;
                w = Wi[idx];
                gx = ix ? 0.5f * (Wi[xp] - Wi[xm]) : 0.0f;
                gy = iy ? 0.5f * (Wi[yp] - Wi[ym]) : 0.0f;
                gz = iz ? 0.5f * (Wi[zp] - Wi[zm]) : 0.0f;
            } else {
                float ux = 0.0f, uy = 0.0f, uz = 0.0f;
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) {
                    float hx = 0.0f, hy = 0.0f, hz = 0.0f;
                    const float4 *wr = win + bb * wxn + sx;
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const float4 q = wr[a];
                        hx = fmaf(wxs[a], q.x, hx);
                        hy = fmaf(wxs[a], q.y, hy);
                        hz = fmaf(wxs[a], q.z, hz);
                    }
                    ux = fmaf(wys[bb], hx, ux);
                    uy = fmaf(wys[bb], hy, uy);
                    uz = fmaf(wys[bb], hz, uz);
                }
                float g[3];
                w = interp3<DEG, true>(Mi, nx, ny, nz, (float)x - ux, (float)j - uy, (float)z - uz, cl.x, cl.y, g);
                gx = g[0];
                gy = g[1];
                gz = g[2];
            }
            const float r = Ri[idx];
            const bool on = (thr == 0.0f) || (w > thr);
            const float dl = on ? w - r : 0.0f;
            ssd += (double)dl * (double)dl;
            cnt += on ? 1u : 0u;
            const float fx = gx * dl, fy = gy * dl, fz = gz * dl;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                acc[3 * t + 0] = fmaf(wys[t], fx, acc[3 * t + 0]);
                acc[3 * t + 1] = fmaf(wys[t], fy, acc[3 * t + 1]);
                acc[3 * t + 2] = fmaf(wys[t], fz, acc[3 * t + 2]);
            }
        }
#pragma unroll
        for (int k = 0; k < 12; ++k)
            acc_s[12 * x + k] = acc[k];
    }
    const double s = block_sum(ssd, red);
    const unsigned c = block_sum_u(cnt, redu);
    if (threadIdx.x == 0) {
        ssd_part[b] = s;
        cnt_part[b] = c;
    }
    __syncthreads();
    // x contraction: output o = (t * cx + kx) * 3 + comp
    const int cx = d.cx;
    const float *twf = (const float *)tw;
    float *Pout = P + d.part_off + ((long long)z * d.ncy + band) * 12 * cx;
    for (int o = threadIdx.x; o < 12 * cx; o += kBlock) {
        const int comp = o % 3;
        const int q = o / 3;
        const int kx = q % cx;
        const int t = q / cx;
        const int xb = cells[d.cell_x + max(kx - 3, 0)], xe = cells[d.cell_x + kx + 1];
        float sum = 0.0f;
        for (int x = xb; x < xe; ++x) {
            const int k = kx - ts[d.tab_x + x];
            sum = fmaf(twf[4 * (d.tab_x + x) + k], acc_s[12 * x + 3 * t + comp], sum);
        }
        Pout[o] = sum;
    }
}

// Stage 1, 2D: block = y knot band.
template <int DEG, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_stage1_2d(const InstDesc *__restrict__ inst, int n_inst, const int *__restrict__ prefix, int nblocks,
                const float2 *__restrict__ coef, const float4 *__restrict__ tw, const int *__restrict__ ts,
                const int *__restrict__ cells, const float *__restrict__ M, const float2 *__restrict__ clamp,
                const float *__restrict__ W, const float *__restrict__ R, float thr, float *__restrict__ P,
                double *__restrict__ ssd_part, unsigned *__restrict__ cnt_part)
{
    extern __shared__ float smem[];
    __shared__ double red[kBlock / 64];
    __shared__ unsigned redu[kBlock / 64];
    const int b = xcd_remap(blockIdx.x, nblocks);
    const int i = find_inst(prefix, n_inst, b);
    const InstDesc d = inst[i];
    const int band = b - prefix[i];
    const int j0 = cells[d.cell_y + band], j1 = cells[d.cell_y + band + 1];
    const int nx = d.nx, ny = d.ny;
    float *acc_s = smem;                       // nx * 8 floats
    float2 *win = (float2 *)(smem + 8 * nx);   // exact mode: 4 x (cx + 4)
    const int wxn = d.cx + 4;
    if constexpr (MODE == 1) {
        const float2 *cbase = coef + d.coef_off;
        for (int e = threadIdx.x; e < 4 * wxn; e += kBlock) {
            const int ky = band + e / wxn, kx = e % wxn;
            win[e] = (kx < d.cx && ky < d.cy) ? cbase[(long long)ky * d.cx + kx] : make_float2(0.0f, 0.0f);
        }
        __syncthreads();
    }
    const float *Wi = W + d.img_off;
    const float *Ri = R + d.img_off;
    const float *Mi = M + d.img_off;
    const float2 cl = clamp[i];
    double ssd = 0.0;
    unsigned cnt = 0;
    for (int x = threadIdx.x; x < nx; x += kBlock) {
        float acc[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            acc[k] = 0.0f;
        const int sx = ts[d.tab_x + x];
        const float4 wxv = tw[d.tab_x + x];
        const float wxs[4] = {wxv.x, wxv.y, wxv.z, wxv.w};
        for (int j = j0; j < j1; ++j) {
            const long long idx = (long long)j * nx + x;
            const float4 wyv = tw[d.tab_y + j];
            const float wys[4] = {wyv.x, wyv.y, wyv.z, wyv.w};
            float w, gx, gy;
            if constexpr (MODE == 0) {
                const bool ix = x > 0 && x < nx - 1, iy = j > 0 && j < ny - 1;
                const long long xm = ix ? idx - 1 : idx, xp = ix ? idx + 1 : idx;
                const long long ym = iy ? idx - nx : idx, yp = iy ? idx + nx : idx;
                w = Wi[idx];
                gx = ix ? 0.5f * (Wi[xp] - Wi[xm]) : 0.0f;
                gy = iy ? 0.5f * (Wi[yp] - Wi[ym]) : 0.0f;
            } else {
                float ux = 0.0f, uy = 0.0f;
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) {
                    float hx = 0.0f, hy = 0.0f;
                    const float2 *wr = win + bb * wxn + sx;
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const float2 q = wr[a];
                        hx = fmaf(wxs[a], q.x, hx);
                        hy = fmaf(wxs[a], q.y, hy);
                    }
                    ux = fmaf(wys[bb], hx, ux);
                    uy = fmaf(wys[bb], hy, uy);
                }
                float g[2];
                w = interp2<DEG, true>(Mi, nx, ny, (float)x - ux, (float)j - uy, cl.x, cl.y, g);
                gx = g[0];
                gy = g[1];
            }
            const float r = Ri[idx];
            const bool on = (thr == 0.0f) || (w > thr);
            const float dl = on ? w - r : 0.0f;
            ssd += (double)dl * (double)dl;
            cnt += on ? 1u : 0u;
            const float fx = gx * dl, fy = gy * dl;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                acc[2 * t + 0] = fmaf(wys[t], fx, acc[2 * t + 0]);
                acc[2 * t + 1] = fmaf(wys[t], fy, acc[2 * t + 1]);
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            acc_s[8 * x + k] = acc[k];
    }
    const double s = block_sum(ssd, red);
    const unsigned c = block_sum_u(cnt, redu);
    if (threadIdx.x == 0) {
        ssd_part[b] = s;
        cnt_part[b] = c;
    }
    __syncthreads();
    const int cx = d.cx;
    const float *twf = (const float *)tw;
    float *Pout = P + d.part_off + (long long)band * 8 * cx;
    for (int o = threadIdx.x; o < 8 * cx; o += kBlock) {
        const int comp = o & 1;
        const int q = o >> 1;
        const int kx = q % cx;
        const int t = q / cx;
        const int xb = cells[d.cell_x + max(kx - 3, 0)], xe = cells[d.cell_x + kx + 1];
        float sum = 0.0f;
        for (int x = xb; x < xe; ++x) {
            const int k = kx - ts[d.tab_x + x];
            sum = fmaf(twf[4 * (d.tab_x + x) + k], acc_s[8 * x + 2 * t + comp], sum);
        }
        Pout[o] = sum;
    }
}

// Per instance: SSD sum in a fixed order, cost and gradient scale
// (mia/template/ssd.hh:54-99,125-181; weight: mia/template/fullcost.cxx:36-45).
__global__ void __launch_bounds__(kBlock)
    k_finalize(const InstDesc *__restrict__ inst, const int *__restrict__ prefix, const double *__restrict__ ssd_part,
               const unsigned *__restrict__ cnt_part, int norm, float thr, double weight, double *__restrict__ costs,
               double *__restrict__ gscale)
{
    __shared__ double red[kBlock / 64];
    __shared__ unsigned redu[kBlock / 64];
    const int i = blockIdx.x;
    const int beg = prefix[i], end = prefix[i + 1];
    double s = 0.0;
    unsigned c = 0;
    for (int k = beg + threadIdx.x; k < end; k += kBlock) {
        s += ssd_part[k];
        c += cnt_part[k];
    }
    s = block_sum(s, red);
    c = block_sum_u(c, redu);
    if (threadIdx.x == 0) {
        const InstDesc d = inst[i];
        const double n = (double)d.nx * d.ny * d.nz;
        double cost, sc;
        if (thr == 0.0f) {
            sc = norm ? 1.0 / n : 1.0;
            cost = 0.5 * s * sc;
        } else if (c > 0) {
            sc = 1.0 / (double)c;
            cost = 0.5 * s * sc;
        } else {
            sc = 0.0;
            cost = (double)FLT_MAX;
        }
        costs[i] = weight * cost;
        gscale[i] = -sc * weight;
    }
}

// Per coefficient: gradient = scale * sum_z wz(z) sum_t P[z][ky - t][t][kx][.]
// in FP64, fixed order (adjoint of spline.cc:731-791 completed along y and z).
__global__ void __launch_bounds__(kBlock)
    k_reduce3d(const InstDesc *__restrict__ inst, int n_inst, const int *__restrict__ prefix,
               const float *__restrict__ P, const float4 *__restrict__ tw, const int *__restrict__ ts,
               const int *__restrict__ cells, const double *__restrict__ gscale, double *__restrict__ grad)
{
    const int b = blockIdx.x;
    const int i = find_inst(prefix, n_inst, b);
    const InstDesc d = inst[i];
    const long long e = (long long)(b - prefix[i]) * kBlock + threadIdx.x;
    const long long ncoef = (long long)d.cx * d.cy * d.cz;
    if (e >= ncoef)
        return;
    const int kx = (int)(e % d.cx);
    const int ky = (int)((e / d.cx) % d.cy);
    const int kz = (int)(e / ((long long)d.cx * d.cy));
    const int zb = cells[d.cell_z + max(kz - 3, 0)], ze = cells[d.cell_z + kz + 1];
    const float *twf = (const float *)tw;
    const long long row = 12LL * d.cx;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int z = zb; z < ze; ++z) {
        const double wz = twf[4 * (d.tab_z + z) + (kz - ts[d.tab_z + z])];
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int band = ky - t;
            if (band < 0 || band >= d.ncy)
                continue;
            const float *p = P + d.part_off + ((long long)z * d.ncy + band) * row + (t * d.cx + kx) * 3;
            s0 += p[0];
            s1 += p[1];
            s2 += p[2];
        }
        g0 += wz * s0;
        g1 += wz * s1;
        g2 += wz * s2;
    }
    const double sc = gscale[i];
    double *out = grad + d.par_off + 3 * e;
    out[0] = sc * g0;
    out[1] = sc * g1;
    out[2] = sc * g2;
}

__global__ void __launch_bounds__(kBlock)
    k_reduce2d(const InstDesc *__restrict__ inst, int n_inst, const int *__restrict__ prefix,
               const float *__restrict__ P, const double *__restrict__ gscale, double *__restrict__ grad)
{
    const int b = blockIdx.x;
    const int i = find_inst(prefix, n_inst, b);
    const InstDesc d = inst[i];
    const long long e = (long long)(b - prefix[i]) * kBlock + threadIdx.x;
    if (e >= (long long)d.cx * d.cy)
        return;
    const int kx = (int)(e % d.cx);
    const int ky = (int)(e / d.cx);
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int band = ky - t;
        if (band < 0 || band >= d.ncy)
            continue;
        const float *p = P + d.part_off + (long long)band * 8 * d.cx + (t * d.cx + kx) * 2;
        s0 += p[0];
        s1 += p[1];
    }
    const double sc = gscale[i];
    double *out = grad + d.par_off + 2 * e;
    out[0] = sc * s0;
    out[1] = sc * s1;
}

} // namespace mia_hip

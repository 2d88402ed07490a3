// Device kernels of the B-spline SSD cost + gradient path for gfx950.
//
// Data flow of one evaluation (every launch covers all instances of a batch):
//   k_warp      per 32x8x8 (3D) / 64x32 (2D) voxel tile: displacement u(x) by a
//               separable contraction of the double parameters through an LDS
//               coefficient window, sample positions x - u(x) kept in
//               registers, the tile's bounding box of interpolation taps
//               loaded once into LDS (mirror indices applied on the load), then
//               the 4^3 (4^2) taps per voxel read from LDS, clamp, store W.
//               A tile whose bounding box does not fit LDS gathers from global
//               memory instead (same arithmetic).                  [reference mode]
//   k_stage1    per (z slice, y knot band): residual d = W - R, force
//               f = grad * d, FP64 SSD partial (wave64 shuffle + fixed-order
//               block sum), contraction of f with the band's y weights in
//               registers and with the x weights through LDS -> partial sums
//               P[z][band][t][kx][c] on the coefficient grid.  In exact mode it
//               computes the warp and the analytic image gradient inline.
//               3D reference mode (rows up to 256 voxels): k_stage1z_3d per
//               (z knot cell, y band) also folds the cell's slices with their z
//               weights -> P[cell][band][k][t][kx][c], read once by
//               k_reducez_3d.
//   k_finalize  (only with autothresh) per instance: SSD sum, mask count, cost
//               and gradient scale.
//   k_reduce    per coefficient component: fixed-order FP64 sum over the four y
//               bands and the z support -> gradient; without autothresh its
//               first n_inst blocks also do the finalize work.  No atomics: the
//               result is deterministic and independent of the batch layout.
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
#include <type_traits>

namespace mia_hip {

constexpr int kBlock = 256;
constexpr int kMaxRow = 512;       // longest x row the stage-1 kernel keeps in LDS
// 3D warp tile and LDS budgets
// Tile shapes of the 3D warp: TX x TY threads (TX * TY = kBlock), TZ voxels
// per thread along z.  The displacement intermediates (phases 1-2) and the box
// of moving-image taps (phases 3-4) share one LDS buffer; the box has
// compile-time pitches PX x PY (x PZ slices) so that the 64 taps of a voxel are
// one base address plus immediate offsets.  LDS serves a wave's two halves
// (32 lanes each) one after the other, 32 banks of 4 bytes; the voxels of a
// half-wave are 8 x columns by 4 rows (G8X4, when PX = 8 mod 32: the four
// rows' taps start a quarter of the banks apart, and a deformation that moves
// a row's taps by a row stays within its quarter), else TX columns of 32 / TX
// rows.  A tap gather then costs about 1.5 bank cycles per half-wave at
// spline-scale deformations, where one 32-column row costs about 2 (a row
// switch inside it shifts its banks by PX).
template <int TX_, int TY_, int TZ_, int PX_, int PY_, int PZ_>
struct W3Tile {
    static constexpr int TX = TX_, TY = TY_, TZ = TZ_, PX = PX_, PY = PY_, PZ = PZ_;
    static_assert(TX * TY == kBlock, "one thread per (x, y) column");
    static constexpr int LOG_TX = TX == 16 ? 4 : TX == 32 ? 5 : 6;
    static constexpr int ROWS = 64 / TX; // tile rows per wave
    static constexpr bool G8X4 = PX % 32 == 8 && TX % 8 == 0 && TY % 4 == 0;
    static constexpr int Buf = PX * PY * PZ;
    static constexpr int Buf4 = (Buf + 3) / 4;
};
// longest z knot window of a tile that the separable displacement keeps in
// registers (TZ = 8 at spline rate >= 2)
constexpr int kW3MaxKz = 8;
// the variants the autotuner chooses from (index = variant id)
using W3T0 = W3Tile<32, 8, 8, 40, 16, 15>;  // 38.4 KB LDS: 4 blocks per CU
using W3T1 = W3Tile<32, 8, 4, 40, 16, 12>;  // 30.7 KB: 5 blocks per CU
using W3T2 = W3Tile<64, 4, 8, 72, 12, 15>;  // 51.8 KB: 3 blocks per CU, one row per wave
using W3T3 = W3Tile<16, 16, 8, 36, 24, 15>; // 51.8 KB: 3 blocks per CU
constexpr int kW3Variants = 4;
__host__ __device__ inline void w3_tile_dims(int v, int &tx, int &ty, int &tz)
{
    switch (v) {
    case 1: tx = W3T1::TX; ty = W3T1::TY; tz = W3T1::TZ; break;
    case 2: tx = W3T2::TX; ty = W3T2::TY; tz = W3T2::TZ; break;
    case 3: tx = W3T3::TX; ty = W3T3::TY; tz = W3T3::TZ; break;
    default: tx = W3T0::TX; ty = W3T0::TY; tz = W3T0::TZ; break;
    }
}
// 2D warp tile variants (x, y pixels per block of kBlock threads; a thread
// takes one x column and kBlock / x rows apart; the taps from an LDS box of
// the tile, or gathered straight from the image: variants 4 and 5) and LDS
// budgets
constexpr int kW2Variants = 6;
__host__ __device__ constexpr int w2_tile_x(int v) { return v == 2 || v == 5 ? 32 : v == 3 ? 128 : 64; }
__host__ __device__ constexpr int w2_tile_y(int v) { return v == 1 || v == 3 || v == 4 ? 16 : 32; }
__host__ __device__ constexpr bool w2_box(int v) { return v < 4; }
// float2, largest coefficient window over the variants (rate 1)
constexpr int kW2Win = (128 + 4) * (16 + 4);
constexpr int kW2Region = 6144;                     // floats: the tap box
// floats ahead of the coefficient window in the dynamic LDS: the tap box, or
// without one the y-contracted displacement of the tile (TY x (TX + 4) float2)
__host__ __device__ constexpr int w2_region(int v)
{
    return w2_box(v) ? kW2Region : w2_tile_y(v) * (w2_tile_x(v) + 4) * 2;
}

// Per-instance description, identical on host and device.
struct InstDesc {
    int nx, ny, nz;           // image size (nz = 1 in 2D)
    int cx, cy, cz;           // coefficient grid (cz = 1 in 2D)
    int ncx, ncy, ncz;        // number of x / y / z knot bands (max start index + 1)
    long long img_off;        // offset of the instance's images (floats)
    long long par_off;        // offset of its parameters / gradient (doubles)
    long long part_off;       // offset of its stage-1 partials (floats)
    int tab_x, tab_y, tab_z;  // offsets into the axis tables (one entry per voxel)
    int cell_x, cell_y, cell_z; // offsets into the band tables (csize + 1 entries)
    int zcell;                // 3D reference gradient: stage 1 per (z knot cell, y band)
};

// Entry i of a spline axis table with the reference's semantics
// (mia/3d/transform/spline.cc:237-273, mia/2d/transform/spline.cc:181-200):
// float scale (csize - 3) / (n - 1), float position i*scale + 1, start index
// fastfloor(pos) - 1 (16.16 rounding, mia/core/splinekernel.cc:164-192),
// cubic weights at pos - fastfloor(pos) in double stored as float, the last
// sample's fourth weight dropped.  Returns the start index.  The host set-up
// and the device set-up of the allocation-free entry point both call it; with
// contraction off every operation rounds as written, so both build the same
// bits.
__host__ __device__ inline int spline_axis_entry(int i, int n, int cs, float w[4])
{
#pragma clang fp contract(off)
    const float scale = (float)(cs - 3) / (float)(n - 1);
    const float si = (float)i * scale;
    const float pos = si + 1.0f;
    const int f = (int)floor(nearbyint((double)pos * 65536.0) / 65536.0);
    const double x = (double)pos - f;
    const double xm1 = 1.0 - x;
    const double w3 = (1.0 / 6.0) * x * x * x;
    const double w0 = (1.0 / 6.0) * xm1 * xm1 * xm1;
    const double w2 = x - 2.0 * w3 + w0;
    const double w1 = xm1 - 2.0 * w0 + w3;
    w[0] = (float)w0;
    w[1] = (float)w1;
    w[2] = (float)w2;
    w[3] = (i == n - 1) ? 0.0f : (float)w3;
    return f - 1;
}

// Where the stage-1 SSD partials of an instance are: per (slice, y band) block
// at s1[i] .. s1[i + 1], or, for an instance with zcell set while zmode is on
// (3D reference gradient), per (z knot cell, y band) block at zoff + s1z[i] ..
// zoff + s1z[i + 1].
struct SsdIndex {
    const int *s1, *s1z;
    int zoff, zmode;
};

// Spline penalties evaluated beside the image cost (on another stream of the
// same evaluation): per instance whether it has one (on[i]), its weight, the
// unweighted value of its quadratic form and, per parameter, its weighted
// gradient -w dP/dc.  The finalize and the reduce add them to the image
// cost's value and gradient as the reference's full cost adds a penalty's:
// each product rounded, then the sum (mia/3d/splinetransformpenalty.cc:52-71).
// A null `on` means no penalties.
struct PenaltyIn {
    const int *on;
    const double *w, *value, *grad;
};

// grad = sc * g (+ the instance's penalty gradient element e)
__device__ __forceinline__ double grad_with_penalty(double sc, double g, const PenaltyIn &pen, int i, long long e)
{
#pragma clang fp contract(off)
    double v = sc * g;
    if (pen.on && pen.on[i])
        v = v + pen.grad[e];
    return v;
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Block -> instance: the largest i < n with prefix[i] <= b (prefix is
// non-decreasing with prefix[0] = 0, so instances without blocks in a launch
// are skipped).  Every lane of the calling wave must be active and pass the
// same b; the result is wave-uniform.  Up to 64 instances one load per lane
// and a ballot, up to 4096 the same over every s-th entry and then within the
// chunk found (two loads in sequence instead of a binary search's log2(n)),
// beyond that a binary search.
__device__ __forceinline__ int find_inst(const int *__restrict__ prefix, int n, int b)
{
    const int lane = threadIdx.x & 63;
    if (n == 1)
        return 0;
    if (n <= 64) {
        const int v = prefix[min(lane, n - 1)];
        const unsigned long long m = __builtin_amdgcn_ballot_w64(lane < n && v <= b);
        return 63 - __builtin_clzll(m);
    }
    if (n <= 64 * 64) {
        const int s = (n + 63) / 64; // chunk length, <= 64
        const int v = prefix[min(lane * s, n - 1)];
        const unsigned long long m = __builtin_amdgcn_ballot_w64(lane * s < n && v <= b);
        const int c0 = (63 - __builtin_clzll(m)) * s;
        const int u = prefix[min(c0 + lane, n - 1)];
        const unsigned long long m2 = __builtin_amdgcn_ballot_w64(lane < s && c0 + lane < n && u <= b);
        return c0 + 63 - __builtin_clzll(m2);
    }
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
__device__ __forceinline__ float fastfloor16f(float v)
{
    return floorf(rintf(v * 65536.0f) * (1.0f / 65536.0f));
}

__device__ __forceinline__ int fastfloor16(float v)
{
    return (int)fastfloor16f(v);
}

// Sample positions are kept in [-1e6, 1e6]: the float -> int conversion and
// the mirror arithmetic stay in range for any input (NaN maps to the lower
// bound because fmaxf drops a NaN operand).
__device__ __forceinline__ float clamp_pos(float p)
{
    return fminf(fmaxf(p, -1.0e6f), 1.0e6f);
}

// The same clamp as one median of three, equal to clamp_pos for any p that
// is not NaN (a NaN p gives an unspecified value; callers that may see one
// apply clamp_pos to the result).
__device__ __forceinline__ float clamp_pos_nn(float p)
{
    float r;
    asm("v_med3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(p), "v"(-1.0e6f), "v"(1.0e6f));
    return r;
}

// e / d for 0 <= e < 2^21 and d >= 1, with inv = 1.0f / d: (e + 0.5) / d is at
// least 0.5 / d away from an integer, more than its float error (two roundings,
// at most about 2^-23 (e + 0.5) / d < 0.25 / d). Callers pass e < 4096.
__device__ __forceinline__ int small_div(int e, float inv)
{
    return (int)(((float)e + 0.5f) * inv);
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
// 3 (mia/core/splinekernel/bspline.cc:258-285) or 1 (:90-107).  SCALED
// (degree 3): the weights times 6, without the two multiplications by 1/6;
// a product of three axes' scaled weights then carries the factor 216
// (kWScale3).
template <int DEG, bool GRAD, bool SCALED = false>
__device__ __forceinline__ int kweights(float p, float *w, float *dw)
{
    // positions are clamped to +-1e6, so the floor is an exact integer in float
    const float flf = fastfloor16f(p);
    const int fl = (int)flf;
    const float f = p - flf;
    if constexpr (DEG == 3 && SCALED) {
        static_assert(!GRAD, "scaled weights carry no derivative");
        const float fm1 = 1.0f - f;
        w[3] = f * f * f;
        w[0] = fm1 * fm1 * fm1;
        w[2] = fmaf(6.0f, f, fmaf(-2.0f, w[3], w[0]));
        w[1] = fmaf(6.0f, fm1, fmaf(-2.0f, w[0], w[3]));
        return fl - 1;
    } else if constexpr (DEG == 3) {
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

// kweights' scaled weights (degree 3; degree 1 as kweights), the floor of p
// returned as a float (an exact integer: positions are clamped to +-1e6)
template <int DEG>
__device__ __forceinline__ float kweights_sf(float p, float *w)
{
    const float flf = fastfloor16f(p);
    const float f = p - flf;
    if constexpr (DEG == 3) {
        const float fm1 = 1.0f - f;
        w[3] = f * f * f;
        w[0] = fm1 * fm1 * fm1;
        w[2] = fmaf(6.0f, f, fmaf(-2.0f, w[3], w[0]));
        w[1] = fmaf(6.0f, fm1, fmaf(-2.0f, w[0], w[3]));
    } else {
        w[0] = 1.0f - f;
        w[1] = f;
    }
    return flf;
}

// v clamped to [vmin, vmax] (vmin <= vmax) as one median of three (values
// interpolated from finite coefficients are never NaN).
__device__ __forceinline__ float clamp_val(float v, float vmin, float vmax)
{
    float r;
    asm("v_med3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(vmin), "v"(vmax));
    return r;
}

// 1 / 216: the factor of a trilinear product of scaled cubic weights
template <int DEG>
constexpr float kWScale3 = DEG == 3 ? 1.0f / 216.0f : 1.0f;

// Interpolation of the prefiltered image at p from global memory with mirror
// indices and clamp to [vmin, vmax] (mia/3d/interpolator.cxx:250-287,
// z->y->x contraction as add_3d :187-210); GRAD adds the analytic gradient,
// zero where the value is clamped.  Without GRAD the weights are the scaled
// ones and the sum is multiplied by kWScale3 once, as in the warp's LDS path.
template <int DEG, bool GRAD>
__device__ __forceinline__ float interp3(const float *__restrict__ M, int nx, int ny, int nz, float px,
                                         float py, float pz, float vmin, float vmax, float *g)
{
    constexpr int NS = DEG == 3 ? 4 : 2;
    float wx[NS], wy[NS], wz[NS], dx[NS], dy[NS], dz[NS];
    const int sx = kweights<DEG, GRAD, !GRAD>(px, wx, dx);
    const int sy = kweights<DEG, GRAD, !GRAD>(py, wy, dy);
    const int sz = kweights<DEG, GRAD, !GRAD>(pz, wz, dz);
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
    } else if constexpr (DEG == 3) {
        v *= kWScale3<DEG>;
    }
    return clamp_val(v, vmin, vmax);
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
    return clamp_val(v, vmin, vmax);
}

// The same contractions over an LDS box that already holds the mirrored taps
// (local start indices lx/ly/lz inside an RX x RY x . box); same arithmetic
// order as the global versions.
template <int DEG>
__device__ __forceinline__ float contract_lds3(const float *R, int RX, int RY, int lx, int ly, int lz,
                                               const float *wx, const float *wy, const float *wz)
{
    constexpr int NS = DEG == 3 ? 4 : 2;
    float v = 0.0f;
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        const float *slab = R + (lz + a) * RX * RY;
        float vy = 0.0f;
#pragma unroll
        for (int b = 0; b < NS; ++b) {
            const float *row = slab + (ly + b) * RX + lx;
            float vx = 0.0f;
#pragma unroll
            for (int k = 0; k < NS; ++k)
                vx = fmaf(wx[k], row[k], vx);
            vy = fmaf(wy[b], vx, vy);
        }
        v = fmaf(wz[a], vy, v);
    }
    return v;
}

template <int DEG>
__device__ __forceinline__ float contract_lds2(const float *R, int RX, int lx, int ly, const float *wx,
                                               const float *wy)
{
    constexpr int NS = DEG == 3 ? 4 : 2;
    float v = 0.0f;
#pragma unroll
    for (int b = 0; b < NS; ++b) {
        const float *row = R + (ly + b) * RX + lx;
        float vx = 0.0f;
#pragma unroll
        for (int k = 0; k < NS; ++k)
            vx = fmaf(wx[k], row[k], vx);
        v = fmaf(wy[b], vx, v);
    }
    return v;
}

// Fill an RX x RY x RZ box of LDS with the image values at (rx0 + ix, ...),
// mirrored at the image borders.  Each thread walks the flat box index in
// steps of the block size, updating (ix, iy, iz) incrementally: the step is
// split into (kBlock % RX, row step % RY, row step / RY), so each coordinate
// needs at most one carry per step, for any box shape.
__device__ __forceinline__ void fill_box3(float *box, const float *__restrict__ img, int nx, int ny, int nz,
                                          int rx0, int ry0, int rz0, int RX, int RY, int RZ)
{
    const int vol = RX * RY * RZ;
    const int stx = kBlock % RX, str = kBlock / RX; // step: str rows and stx columns
    int ix = threadIdx.x % RX;
    int row = threadIdx.x / RX;
    int iy = row % RY, iz = row / RY;
    const int sty = str % RY, stz = str / RY;
    const bool xin = rx0 >= 0 && rx0 + RX <= nx;
    const bool yin = ry0 >= 0 && ry0 + RY <= ny;
    const bool zin = rz0 >= 0 && rz0 + RZ <= nz;
    for (int e = threadIdx.x; e < vol; e += kBlock) {
        const int gx = xin ? rx0 + ix : mirror_idx(rx0 + ix, nx);
        const int gy = yin ? ry0 + iy : mirror_idx(ry0 + iy, ny);
        const int gz = zin ? rz0 + iz : mirror_idx(rz0 + iz, nz);
        box[e] = img[((long long)gz * ny + gy) * nx + gx];
        ix += stx;
        iy += sty;
        iz += stz;
        if (ix >= RX) {
            ix -= RX;
            ++iy;
        }
        if (iy >= RY) {
            iy -= RY;
            ++iz;
        }
    }
}

__device__ __forceinline__ void fill_box2(float *box, const float *__restrict__ img, int nx, int ny, int rx0,
                                          int ry0, int RX, int RY)
{
    // eight loads in flight per thread before their LDS stores (a load and
    // its store per iteration would wait out the memory latency each time);
    // vol <= kW2Region, so small_div is exact
    constexpr int U = 8;
    const int vol = RX * RY;
    const float irx = 1.0f / (float)RX;
    const bool xin = rx0 >= 0 && rx0 + RX <= nx;
    const bool yin = ry0 >= 0 && ry0 + RY <= ny;
    for (int e0 = threadIdx.x; e0 < vol; e0 += U * kBlock) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * kBlock;
            v[u] = 0.0f;
            if (e < vol) {
                const int iy = small_div(e, irx), ix = e - iy * RX;
                const int gx = xin ? rx0 + ix : mirror_idx(rx0 + ix, nx);
                const int gy = yin ? ry0 + iy : mirror_idx(ry0 + iy, ny);
                v[u] = img[(long long)gy * nx + gx];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (e0 + u * kBlock < vol)
                box[e0 + u * kBlock] = v[u];
    }
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

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ float lane_f(float v, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// Minimum / maximum of values that are never NaN in one instruction each
// (fminf / fmaxf cost three: IEEE mode quiets both operands before v_min /
// v_max); callers rule out NaN inputs or route them elsewhere.
__device__ __forceinline__ float min2_nn(float a, float b)
{
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

__device__ __forceinline__ float max2_nn(float a, float b)
{
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

__device__ __forceinline__ float min3_nn(float a, float b, float c)
{
    float r;
    asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

__device__ __forceinline__ float max3_nn(float a, float b, float c)
{
    float r;
    asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// minimum / maximum of v[0 .. N-1] (N >= 1) with three-operand instructions
template <int N>
__device__ __forceinline__ float min_n(const float *v)
{
    float m = v[0];
    int k = 1;
#pragma unroll
    for (; k + 1 < N; k += 2)
        m = min3_nn(m, v[k], v[k + 1]);
    if (k < N)
        m = min2_nn(m, v[k]);
    return m;
}

template <int N>
__device__ __forceinline__ float max_n(const float *v)
{
    float m = v[0];
    int k = 1;
#pragma unroll
    for (; k + 1 < N; k += 2)
        m = max3_nn(m, v[k], v[k + 1]);
    if (k < N)
        m = max2_nn(m, v[k]);
    return m;
}

// Minimum of (a, b, c) and maximum of (d, e, f) over the wave, as uniform
// values.  One v_min / v_max per step with the DPP operand folded in: within
// each row of 16 lanes (quad swaps, half-row and row mirrors), then row 0 into
// row 1 and row 2 into row 3 (row_bcast:15), then lane 31 into rows 2 and 3
// (row_bcast:31); lane 63 ends with the result, read into a scalar.  Rows
// masked off keep their value (the destination is also the second source).
// The six reductions interleave, so every DPP read comes six instructions
// after the write of its source (two wait states are needed; "s_nop 1" covers
// the inputs).  No NaN quieting: callers pass values that are never NaN.
__device__ __forceinline__ void wave_min3_max3(float &a, float &b, float &c, float &d, float &e, float &f)
{
    asm volatile("s_nop 1\n\t"
                 "v_min_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %1, %1, %1 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %2, %2, %2 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %3, %3, %3 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %4, %4, %4 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %5, %5, %5 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %0, %0, %0 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %1, %1, %1 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %2, %2, %2 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %3, %3, %3 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %4, %4, %4 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %5, %5, %5 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %0, %0, %0 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %1, %1, %1 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %2, %2, %2 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %3, %3, %3 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %4, %4, %4 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %5, %5, %5 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %0, %0, %0 row_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %1, %1, %1 row_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %2, %2, %2 row_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %3, %3, %3 row_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %4, %4, %4 row_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_max_f32_dpp %5, %5, %5 row_mirror row_mask:0xf bank_mask:0xf\n\t"
                 "v_min_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
                 "v_min_f32_dpp %1, %1, %1 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
                 "v_min_f32_dpp %2, %2, %2 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
                 "v_max_f32_dpp %3, %3, %3 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
                 "v_max_f32_dpp %4, %4, %4 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
                 "v_max_f32_dpp %5, %5, %5 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
                 "v_min_f32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
                 "v_min_f32_dpp %1, %1, %1 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
                 "v_min_f32_dpp %2, %2, %2 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
                 "v_max_f32_dpp %3, %3, %3 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
                 "v_max_f32_dpp %4, %4, %4 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
                 "v_max_f32_dpp %5, %5, %5 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
                 "s_nop 1"
                 : "+v"(a), "+v"(b), "+v"(c), "+v"(d), "+v"(e), "+v"(f));
    a = lane_f(a, 63);
    b = lane_f(b, 63);
    c = lane_f(c, 63);
    d = lane_f(d, 63);
    e = lane_f(e, 63);
    f = lane_f(f, 63);
}

// Deterministic block sum (fixed shuffle tree per wave, then waves in order);
// the result is valid in thread 0.
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

// Spline coefficient k of an instance, converted to float as the reference
// stores it (C3DFVector / C2DFVector, spline.cc:344-358).
__device__ __forceinline__ float3 coef3(const double *__restrict__ p, long long k)
{
    const double *q = p + 3 * k;
    return make_float3((float)q[0], (float)q[1], (float)q[2]);
}

__device__ __forceinline__ float2 coef2(const double *__restrict__ p, long long k)
{
    const double *q = p + 2 * k;
    return make_float2((float)q[0], (float)q[1]);
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
    mn = wave_min(mn);
    mx = wave_max(mx);
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

// clamp range [min(0, min), max] (mia/3d/interpolator.cxx:108-111), one
// wave over the block partials (min and max are exact in any order)
__global__ void __launch_bounds__(64) k_minmax_final(const float2 *__restrict__ part, int n, float2 *__restrict__ out)
{
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int k = threadIdx.x; k < n; k += 64) {
        const float2 v = part[k];
        mn = fminf(mn, v.x);
        mx = fmaxf(mx, v.y);
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    if (threadIdx.x == 0) {
        float2 r = make_float2(mn, mx);
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
    // each walk loads kU elements (clamped indices, no branch) before it uses
    // them, then runs the recursion over them in order: with one wave per
    // SIMD a load at a time would wait out the memory latency every element
    constexpr int kU = 8;
    for (int k0 = 1; k0 < n - 1; k0 += kU) {
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            v[u] = (double)img[base + (long long)min(k0 + u, n - 2) * stride];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (k0 + u < n - 1) {
                sum += (zn + z2n) * lambda * v[u];
                zn *= p;
                z2n *= iz;
            }
    }
    double c = sum / (1.0 - zn * zn);
    scratch[l] = c;
    for (int k0 = 1; k0 < n; k0 += kU) {
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            v[u] = (double)img[base + (long long)min(k0 + u, n - 1) * stride];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (k0 + u < n) {
                c = lambda * v[u] + p * c;
                scratch[(long long)(k0 + u) * nlines + l] = c;
            }
    }
    double a = (p / (p * p - 1.0)) * (p * scratch[(long long)(n - 2) * nlines + l] + c);
    img[base + (long long)(n - 1) * stride] = (float)a;
    for (int k0 = n - 2; k0 >= 0; k0 -= kU) {
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            v[u] = scratch[(long long)max(k0 - u, 0) * nlines + l];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (k0 - u >= 0) {
                a = p * (a - v[u]);
                img[base + (long long)(k0 - u) * stride] = (float)a;
            }
    }
}

// ---------------------------------------------------------------------------
// warp
// ---------------------------------------------------------------------------

// Buffer resource for loads that the kernel keeps in bounds by construction
// (mirrored image indices, clamped table entries, masked window columns).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void *base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, 0x7fffffff, 0x00020000);
}

// One dword per active lane from base + voff + soff into LDS at lds + 4 * lane
// (lds must be wave-uniform).
__device__ __forceinline__ void dma_dword(__amdgpu_buffer_rsrc_t r, void *lds, unsigned voff, unsigned soff)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, lds, 4, voff, soff, 0, 0);
}


// The NS^3 taps of one voxel from an LDS box with compile-time pitches
// (o0: byte offset of the first tap), z -> y -> x contraction as
// contract_lds3.
template <int DEG, int PX, int PY>
__device__ __forceinline__ float box_contract(const float *box, unsigned o0, const float *wa, const float *wb,
                                              const float *wc)
{
    constexpr int NS = DEG == 3 ? 4 : 2;
    float v = 0.0f;
#pragma unroll
    for (int c = 0; c < NS; ++c) {
        // an opaque copy of the slice's byte offset: the slice's taps are
        // then immediate offsets from one register (left alone, the compiler
        // forms a new address for most pairs of taps beyond the first slice)
        unsigned oc = o0 + (unsigned)(c * PY * PX * 4);
        asm("" : "+v"(oc));
        float vy = 0.0f;
#pragma unroll
        for (int bb = 0; bb < NS; ++bb) {
            const float *row = (const float *)((const char *)box + oc) + bb * PX;
            float vx = 0.0f;
#pragma unroll
            for (int a = 0; a < NS; ++a)
                vx = fmaf(wa[a], row[a], vx);
            vy = fmaf(wb[bb], vx, vy);
        }
        v = fmaf(wc[c], vy, v);
    }
    return v;
}

// Fill an RX x RY x RZ box (pitches PX x PY) with the image values at
// (rx0 + i, ry0 + j, rz0 + k), mirrored at the image borders, straight from
// global memory into LDS: one wave instruction per box row and 64 columns,
// the block's waves taking rows round robin.  A box inside the image in y and
// z walks its rows with scalar offsets updated per row (no branches, no vector
// work); otherwise lane j of the wave first works out the mirrored source and
// LDS offsets of the wave's j-th row and the loop reads them from the lanes.
// The caller waits (vmcnt) and synchronises before reading the box.  Images
// are below 2^31 bytes (32-bit byte offsets).
template <int PX, int PY>
__device__ __forceinline__ void fill_box_rows(float *box, const float *__restrict__ img, int nx, int ny, int nz,
                                              int rx0, int ry0, int rz0, int RX, int RY, int RZ)
{
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int nw = kBlock / 64;
    const __amdgpu_buffer_rsrc_t rm = raw_rsrc(img);
    const int nrows = RY * RZ;
    const bool yzin = ry0 >= 0 && ry0 + RY <= ny && rz0 >= 0 && rz0 + RZ <= nz;
#pragma unroll
    for (int cb = 0; cb < PX; cb += 64) {
        // columns cb .. cb + 63 of the box
        const bool xact = lane + cb < RX;
        int gx = rx0 + cb + lane;
        if (rx0 < 0 || rx0 + RX > nx)
            gx = mirror_idx(xact ? gx : 0, nx);
        const unsigned vo = (unsigned)gx * 4u;
        if (yzin) {
            // row r = wv + nw k at (iy, iz) = (r % RY, r / RY); RY >= 2, so a
            // step of nw = 4 rows wraps at most twice
            int iy = wv, iz = 0;
            while (iy >= RY) {
                iy -= RY;
                ++iz;
            }
            unsigned lo = (unsigned)((iz * PY + iy) * PX + cb) * 4u;
            unsigned so = (unsigned)(((rz0 + iz) * ny + ry0 + iy) * nx) * 4u;
            const unsigned lstep = nw * PX * 4, lwrap = (unsigned)((PY - RY) * PX * 4);
            const unsigned sstep = (unsigned)(nw * nx * 4), swrap = (unsigned)((ny - RY) * nx * 4);
            if (xact)
                for (int r = wv; r < nrows; r += nw) {
                    dma_dword(rm, (char *)box + lo, vo, so);
                    iy += nw;
                    lo += lstep;
                    so += sstep;
                    const bool w1 = iy >= RY;
                    iy -= w1 ? RY : 0;
                    lo += w1 ? lwrap : 0u;
                    so += w1 ? swrap : 0u;
                    const bool w2 = iy >= RY;
                    iy -= w2 ? RY : 0;
                    lo += w2 ? lwrap : 0u;
                    so += w2 ? swrap : 0u;
                }
        } else {
            // 64 rows at a time: lane j holds the offsets of row wv + nw (k0 + j)
            for (int k0 = 0; wv + nw * k0 < nrows; k0 += 64) {
                const int nk = min(64, (nrows - wv - nw * k0 + nw - 1) / nw);
                const int r = min(wv + nw * (k0 + lane), nrows - 1);
                const int iz = small_div(r, 1.0f / (float)RY), iy = r - iz * RY;
                const int gy = mirror_idx(ry0 + iy, ny), gz = mirror_idx(rz0 + iz, nz);
                int so = (gz * ny + gy) * nx * 4;
                int lo = ((iz * PY + iy) * PX + cb) * 4;
                // every lane computes its row's offsets before the branch below:
                // the loop reads them from lanes that the branch switches off,
                // and left alone the compiler moves this arithmetic into the
                // branch
                asm volatile("" : "+v"(so), "+v"(lo));
                if (xact)
                    for (int k = 0; k < nk; ++k)
                        dma_dword(rm, (char *)box + __builtin_amdgcn_readlane(lo, k), vo,
                                  (unsigned)__builtin_amdgcn_readlane(so, k));
            }
        }
    }
}

// 3D warp.  Block = TX x TY x TZ voxels (32 x 8 x 8 by default), thread = one
// (x, y) column of TZ voxels.
//   phase 1: the tile's coefficient window (double params -> float, zero
//            beyond the grid) and the tile's x/y/z spline tables into LDS;
//   phase 2: displacement, contracted along y in LDS (T1[kz][y][kx]), then
//            along x into the thread's registers (one value per z knot), then
//            along z per voxel; when the window does not fit, each thread
//            contracts its column directly from global memory;
//   phase 3: the bounding box of all interpolation taps of the tile's voxels
//            (block min/max of the sample positions) is loaded into LDS with
//            mirror indices applied on the load;
//   phase 4: every voxel reads its 4^3 taps from LDS at one base address plus
//            compile-time offsets, clamps and stores W.  A tile whose box
//            exceeds the PX x PY x PZ pitches gathers from global memory with
//            the same arithmetic.
template <int DEG, class TL>
__global__ void __launch_bounds__(kBlock)
    k_warp3d(const InstDesc *__restrict__ inst, const int *__restrict__ ids, int n_inst,
             const int *__restrict__ prefix, int boff, int nblocks,
             const double *__restrict__ params, const float4 *__restrict__ tw, const int *__restrict__ ts,
             const float *__restrict__ M, const float2 *__restrict__ clamp, float *__restrict__ W)
{
    constexpr int NS = DEG == 3 ? 4 : 2, LO = DEG == 3 ? 1 : 0;
    __shared__ float4 buf4[TL::Buf4];
    __shared__ __align__(16) float red[6][kBlock / 64];
    __shared__ float4 wx_s[TL::TX], wy_s[TL::TY], wz_s[TL::TZ];
    __shared__ int sx_s[TL::TX], sy_s[TL::TY], sz_s[TL::TZ];
    float *region = (float *)buf4;
    const int b = boff + xcd_remap(blockIdx.x, nblocks);
    const int gi = find_inst(prefix, n_inst, b); // position in this launch's instance list
    const int i = ids[gi];
    const InstDesc d = inst[i];
    const int lb = b - prefix[gi];
    const int tiles_x = (d.nx + TL::TX - 1) / TL::TX;
    const int tiles_y = (d.ny + TL::TY - 1) / TL::TY;
    const int tx = lb % tiles_x;
    const int rest = lb / tiles_x;
    const int x0 = tx * TL::TX, y0 = (rest % tiles_y) * TL::TY, z0 = (rest / tiles_y) * TL::TZ;
    // the tile's (x, y) column of this thread: half-wave h holds 8 columns
    // by 4 rows (G8X4), else a wave holds tile rows w, w + 4, ...
    int xl, yl;
    if constexpr (TL::G8X4) {
        const int h = threadIdx.x >> 5, hl = threadIdx.x & 31;
        xl = (h % (TL::TX / 8)) * 8 + (hl & 7);
        yl = (h / (TL::TX / 8)) * 4 + (hl >> 3);
    } else {
        xl = threadIdx.x & (TL::TX - 1);
        yl = ((threadIdx.x >> TL::LOG_TX) & (TL::ROWS - 1)) * 4 + (threadIdx.x >> 6);
    }
    const int x = x0 + xl, y = y0 + yl;
    const bool valid = x < d.nx && y < d.ny;
    const int nzl = min(TL::TZ, d.nz - z0);
    const int kx0 = ts[d.tab_x + x0], ky0 = ts[d.tab_y + y0], kz0 = ts[d.tab_z + z0];
    const int wxn = ts[d.tab_x + min(x0 + TL::TX, d.nx) - 1] + 4 - kx0;
    const int wyn = ts[d.tab_y + min(y0 + TL::TY, d.ny) - 1] + 4 - ky0;
    const int wzn = ts[d.tab_z + z0 + nzl - 1] + 4 - kz0;
    const double *P = params + d.par_off;
    // phase 1: spline tables of the tile (clamped to the image for the
    // partial tiles' unused lanes) and the coefficient window
    if (threadIdx.x < TL::TX) {
        const int xx = min(x0 + (int)threadIdx.x, d.nx - 1);
        wx_s[threadIdx.x] = tw[d.tab_x + xx];
        sx_s[threadIdx.x] = ts[d.tab_x + xx] - kx0;
    } else if (threadIdx.x < TL::TX + TL::TY) {
        const int k = threadIdx.x - TL::TX;
        const int yy = min(y0 + k, d.ny - 1);
        wy_s[k] = tw[d.tab_y + yy];
        sy_s[k] = ts[d.tab_y + yy] - ky0;
    } else if (threadIdx.x < TL::TX + TL::TY + TL::TZ) {
        const int k = threadIdx.x - TL::TX - TL::TY;
        const int zz = min(z0 + k, d.nz - 1);
        wz_s[k] = tw[d.tab_z + zz];
        sz_s[k] = ts[d.tab_z + zz] - kz0;
    }
    const int nwin = wxn * wyn * wzn, nt1 = wzn * TL::TY * wxn;
    const bool separable = nwin + nt1 <= TL::Buf4 && wzn <= kW3MaxKz; // block-uniform
    float4 *cwin = buf4, *T1 = buf4 + nwin;
    // every coefficient this thread reads is finite (then so is the
    // displacement, a convex combination of them)
    bool fin = true;
    if (separable) {
        const float iwx = 1.0f / (float)wxn, iwy = 1.0f / (float)wyn;
        for (int e = threadIdx.x; e < nwin; e += kBlock) {
            const int t = small_div(e, iwx), tz = small_div(t, iwy);
            const int kx = kx0 + e - t * wxn;
            const int ky = ky0 + t - tz * wyn, kz = kz0 + tz;
            float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (kx < d.cx && ky < d.cy && kz < d.cz) {
                const float3 c = coef3(P, ((long long)kz * d.cy + ky) * d.cx + kx);
                fin = fin && __builtin_isfinite(c.x) && __builtin_isfinite(c.y) && __builtin_isfinite(c.z);
                q = make_float4(c.x, c.y, c.z, 0.0f);
            }
            cwin[e] = q;
        }
    }
    __syncthreads();
    float ux[TL::TZ], uy[TL::TZ], uz[TL::TZ];
#pragma unroll
    for (int zl = 0; zl < TL::TZ; ++zl)
        ux[zl] = uy[zl] = uz[zl] = 0.0f;
    if (separable) {
        // phase 2a: T1[kz][yl][kx] = sum_b wy_b(yl) * C[kz][sy(yl) + b][kx]
        const float iwx = 1.0f / (float)wxn;
        for (int e = threadIdx.x; e < nt1; e += kBlock) {
            const int t = small_div(e, iwx);
            const int kx = e - t * wxn;
            const int r = t % TL::TY, kz = t / TL::TY;
            const float4 w = wy_s[r];
            const float ws[4] = {w.x, w.y, w.z, w.w};
            const float4 *src = cwin + (kz * wyn + sy_s[r]) * wxn + kx;
            float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll
            for (int bb = 0; bb < 4; ++bb) {
                const float4 q = src[bb * wxn];
                ax = fmaf(ws[bb], q.x, ax);
                ay = fmaf(ws[bb], q.y, ay);
                az = fmaf(ws[bb], q.z, az);
            }
            T1[e] = make_float4(ax, ay, az, 0.0f);
        }
        __syncthreads();
        // phase 2b: this thread's own column of the x contraction, in
        // registers: t[kz] = sum_a wx_a(xl) * T1[kz][yl][sx(xl) + a], kz < wzn
        float tx[kW3MaxKz], ty[kW3MaxKz], tz[kW3MaxKz];
        {
            const float4 w = wx_s[xl];
            const float ws[4] = {w.x, w.y, w.z, w.w};
            const float4 *src = T1 + yl * wxn + sx_s[xl];
#pragma unroll
            for (int kz = 0; kz < kW3MaxKz; ++kz) {
                float ax = 0.0f, ay = 0.0f, az = 0.0f;
                if (kz < wzn) {
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const float4 q = src[kz * TL::TY * wxn + a];
                        ax = fmaf(ws[a], q.x, ax);
                        ay = fmaf(ws[a], q.y, ay);
                        az = fmaf(ws[a], q.z, az);
                    }
                }
                tx[kz] = ax;
                ty[kz] = ay;
                tz[kz] = az;
            }
        }
        // phase 2c: z contraction per voxel from those registers, also for
        // the slices beyond a partial tile (their table entries repeat the
        // last slice's; the results are not stored).  The first knot of a
        // slice is the same for the whole block, so picking the registers is
        // a scalar branch.
#pragma unroll
        for (int zl = 0; zl < TL::TZ; ++zl) {
            const float4 w = wz_s[zl];
            const float ws[4] = {w.x, w.y, w.z, w.w};
            const int o = __builtin_amdgcn_readfirstlane(sz_s[zl]);
#pragma unroll
            for (int oo = 0; oo + 4 <= kW3MaxKz; ++oo) {
                if (o == oo) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        ux[zl] = fmaf(ws[c], tx[oo + c], ux[zl]);
                        uy[zl] = fmaf(ws[c], ty[oo + c], uy[zl]);
                        uz[zl] = fmaf(ws[c], tz[oo + c], uz[zl]);
                    }
                }
            }
        }
    } else {
        // the window does not fit: contract this column's coefficients from
        // global memory, x then y per kz, then scatter along z
        const float4 wxv = wx_s[xl], wyv = wy_s[yl];
        const float wxs[4] = {wxv.x, wxv.y, wxv.z, wxv.w};
        const float wys[4] = {wyv.x, wyv.y, wyv.z, wyv.w};
        const int sx = sx_s[xl], sy = sy_s[yl];
        for (int kzl = 0; kzl < wzn; ++kzl) {
            const int kz = kz0 + kzl;
            float vx = 0.0f, vy = 0.0f, vz = 0.0f;
            if (kz < d.cz) {
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) {
                    const int ky = ky0 + sy + bb;
                    float hx = 0.0f, hy = 0.0f, hz = 0.0f;
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const int kx = kx0 + sx + a;
                        if (kx < d.cx && ky < d.cy) {
                            const float3 q = coef3(P, ((long long)kz * d.cy + ky) * d.cx + kx);
                            fin = fin && __builtin_isfinite(q.x) && __builtin_isfinite(q.y) &&
                                  __builtin_isfinite(q.z);
                            hx = fmaf(wxs[a], q.x, hx);
                            hy = fmaf(wxs[a], q.y, hy);
                            hz = fmaf(wxs[a], q.z, hz);
                        }
                    }
                    vx = fmaf(wys[bb], hx, vx);
                    vy = fmaf(wys[bb], hy, vy);
                    vz = fmaf(wys[bb], hz, vz);
                }
            }
#pragma unroll
            for (int zl = 0; zl < TL::TZ; ++zl) {
                const int c = kzl - sz_s[zl];
                if (c >= 0 && c < 4) {
                    const float w = ((const float *)&wz_s[zl])[c];
                    ux[zl] = fmaf(w, vx, ux[zl]);
                    uy[zl] = fmaf(w, vy, uy[zl]);
                    uz[zl] = fmaf(w, vz, uz[zl]);
                }
            }
        }
    }
    // Sample positions x - u(x), kept clamped to +-1e6 for the interpolation.
    // The tap box comes from the extremes of the unclamped positions (results
    // of arithmetic, so v_min / v_max need not quiet them first; clamping the
    // extremes afterwards gives the extremes of the clamped positions).  A
    // non-finite coefficient makes the box test fail below, and the tile then
    // takes the global-memory path, which is safe for any position.
    // Lanes beyond the image in x or y and slices beyond it in z take the
    // coordinates of the last column, row or slice (their table entries are
    // those of the last ones already), so they repeat a real voxel's position
    // and leave the box as it is; they are not stored.
    float px[TL::TZ], py[TL::TZ], pz[TL::TZ], qx[TL::TZ], qy[TL::TZ], qz[TL::TZ];
    const float xc = (float)min(x, d.nx - 1), yc = (float)min(y, d.ny - 1);
#pragma unroll
    for (int zl = 0; zl < TL::TZ; ++zl) {
        qx[zl] = xc - ux[zl];
        qy[zl] = yc - uy[zl];
        qz[zl] = (float)min(z0 + zl, d.nz - 1) - uz[zl];
        // NaN positions (non-finite coefficients) only reach the global
        // path below, which clamps them again
        px[zl] = clamp_pos_nn(qx[zl]);
        py[zl] = clamp_pos_nn(qy[zl]);
        pz[zl] = clamp_pos_nn(qz[zl]);
    }
    float mnx = min_n<TL::TZ>(qx), mny = min_n<TL::TZ>(qy), mnz = min_n<TL::TZ>(qz);
    float mxx = max_n<TL::TZ>(qx), mxy = max_n<TL::TZ>(qy), mxz = max_n<TL::TZ>(qz);
    if (!fin)
        mnx = -FLT_MAX;
    // bounding box of the taps of all valid voxels of the tile; the barrier
    // also ends every read of the phase-2 intermediates, which the box reuses
    wave_min3_max3(mnx, mny, mnz, mxx, mxy, mxz);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wid] = mnx;
        red[1][wid] = mny;
        red[2][wid] = mnz;
        red[3][wid] = mxx;
        red[4][wid] = mxy;
        red[5][wid] = mxz;
    }
    __syncthreads();
    static_assert(kBlock / 64 == 4, "one float4 of wave results per quantity");
    mnx = min_n<4>(red[0]);
    mny = min_n<4>(red[1]);
    mnz = min_n<4>(red[2]);
    mxx = max_n<4>(red[3]);
    mxy = max_n<4>(red[4]);
    mxz = max_n<4>(red[5]);
    // every lane holds the same box; readfirstlane makes that visible to the
    // compiler so that the fill below runs its row arithmetic on the scalar unit
    const int rx0 = __builtin_amdgcn_readfirstlane(fastfloor16(clamp_pos(mnx)) - LO);
    const int ry0 = __builtin_amdgcn_readfirstlane(fastfloor16(clamp_pos(mny)) - LO);
    const int rz0 = __builtin_amdgcn_readfirstlane(fastfloor16(clamp_pos(mnz)) - LO);
    const int RX = __builtin_amdgcn_readfirstlane(fastfloor16(clamp_pos(mxx)) - LO + NS) - rx0;
    const int RY = __builtin_amdgcn_readfirstlane(fastfloor16(clamp_pos(mxy)) - LO + NS) - ry0;
    const int RZ = __builtin_amdgcn_readfirstlane(fastfloor16(clamp_pos(mxz)) - LO + NS) - rz0;
    const bool use_lds = RX <= TL::PX && RY <= TL::PY && RZ <= TL::PZ; // block-uniform
    const float *Mi = M + d.img_off;
    if (use_lds) {
        // phase 3: the box rows straight into LDS
        fill_box_rows<TL::PX, TL::PY>(region, Mi, d.nx, d.ny, d.nz, rx0, ry0, rz0, RX, RY, RZ);
        // the loads count on vmcnt only: wait for them before the barrier
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
    if (!valid)
        return;
    const float2 cl = clamp[i];
    // W stores as buffer stores: the lane's column offset plus the slice's
    // scalar offset (images are below 2^31 bytes)
    const __amdgpu_buffer_rsrc_t rwo = raw_rsrc(W + d.img_off);
    const unsigned wvo = (unsigned)(y * d.nx + x) * 4u;
    const unsigned planeb = (unsigned)(d.nx * d.ny) * 4u;
    auto store = [&](int zl, float v) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rwo, wvo, (unsigned)(z0 + zl) * planeb, 0);
    };
    if (use_lds) {
        // phase 4: the taps at one base address plus compile-time offsets;
        // the first tap's box coordinates (a, b, c) are the floors less the
        // box origin (plus LO): small integers, exact in float, and so is
        // the element offset c PX PY + b PX + a (below the box size, < 2^24)
        const float ox = (float)(rx0 + LO), oy = (float)(ry0 + LO), oz = (float)(rz0 + LO);
        // the clamp bounds in vector registers once (not copied per voxel)
        float vlo = cl.x, vhi = cl.y;
        asm("" : "+v"(vlo), "+v"(vhi));
#pragma unroll
        for (int zl = 0; zl < TL::TZ; ++zl) {
            if (zl < nzl) {
                float wa[NS], wb[NS], wc[NS];
                const float a = kweights_sf<DEG>(px[zl], wa) - ox;
                const float b = kweights_sf<DEG>(py[zl], wb) - oy;
                const float c = kweights_sf<DEG>(pz[zl], wc) - oz;
                const unsigned o0 =
                    (unsigned)fmaf(c, (float)(TL::PY * TL::PX), fmaf(b, (float)TL::PX, a)) << 2;
                const float v = box_contract<DEG, TL::PX, TL::PY>(region, o0, wa, wb, wc) * kWScale3<DEG>;
                store(zl, clamp_val(v, vlo, vhi));
            }
        }
    } else {
#pragma unroll
        for (int zl = 0; zl < TL::TZ; ++zl)
            if (zl < nzl)
                store(zl, interp3<DEG, false>(Mi, d.nx, d.ny, d.nz, clamp_pos(px[zl]), clamp_pos(py[zl]),
                                              clamp_pos(pz[zl]), cl.x, cl.y, nullptr));
    }
}

// 2D warp.  Block = TX x TY pixels, thread = TY * TX / kBlock pixels of one
// x column (rows y0 + ly + (kBlock / TX) k).  One coefficient window per tile.
// BOX: the tile's taps in an LDS box sized from the extremes of its sample
// positions; otherwise each pixel gathers its taps from the image (small
// images stay in L2), with no box extent, fill or barrier.  Both contract in
// the same order, so every variant computes the same image.
template <int DEG, int TX, int TY, bool BOX>
__global__ void __launch_bounds__(kBlock)
    k_warp2d(const InstDesc *__restrict__ inst, const int *__restrict__ ids, int n_inst,
             const int *__restrict__ prefix, const int4 *__restrict__ tiles, int boff, int nblocks,
             const double *__restrict__ params, const float4 *__restrict__ tw, const int *__restrict__ ts,
             const float *__restrict__ M, const float2 *__restrict__ clamp, float *__restrict__ W)
{
    constexpr int NS = DEG == 3 ? 4 : 2, LO = DEG == 3 ? 1 : 0;
    static_assert(TX <= kBlock && kBlock % TX == 0 && TY % (kBlock / TX) == 0, "tile shape");
    static_assert((TX + 4) * (TY + 4) <= kW2Win && TY * (TX + 4) <= kW2Region / 2, "tile LDS budget");
    constexpr int RPT = TY / (kBlock / TX);   // rows per thread
    // dynamic LDS: the image box (kW2Region floats; without a box only the
    // y-contracted displacement, TY x (TX + 4) float2), then the coefficient
    // window, sized by the host for the launch's largest window (about 1 KB
    // at rate 8, so that six blocks fit a CU)
    constexpr int REG = BOX ? kW2Region : TY * (TX + 4) * 2;
    extern __shared__ float smem[];
    float *region = smem;
    float2 *win = (float2 *)(smem + REG);
    __shared__ float red[4][kBlock / 64];
    __shared__ float4 wy_s[TY];
    __shared__ int sy_s[TY];
    const int b = boff + xcd_remap(blockIdx.x, nblocks);
    // the tile's instance, origin and coefficient window: from the tile
    // table in one load, or derived from the tile prefix and axis tables
    // (a chain of dependent loads)
    int i, x0, y0, kx0, ky0, wxn, wyn;
    InstDesc d;
    if (tiles) {
        const int4 t0 = tiles[2 * b], t1 = tiles[2 * b + 1];
        i = t0.x;
        x0 = t0.y;
        y0 = t0.z;
        kx0 = t0.w;
        ky0 = t1.x;
        wxn = t1.y;
        wyn = t1.z;
        d = inst[i];
    } else {
        const int gi = find_inst(prefix, n_inst, b);
        i = ids[gi];
        d = inst[i];
        const int lb = b - prefix[gi];
        const int tiles_x = (d.nx + TX - 1) / TX;
        x0 = (lb % tiles_x) * TX;
        y0 = (lb / tiles_x) * TY;
        kx0 = ts[d.tab_x + x0];
        ky0 = ts[d.tab_y + y0];
        wxn = ts[d.tab_x + min(x0 + TX, d.nx) - 1] + 4 - kx0;
        wyn = ts[d.tab_y + min(y0 + TY, d.ny) - 1] + 4 - ky0;
    }
    const int x = x0 + (threadIdx.x & (TX - 1));
    const int ly = threadIdx.x / TX;
    const bool xvalid = x < d.nx;
    const int xc = min(x, d.nx - 1);
    const double *P = params + d.par_off;
    // this thread's x table entry and the clamp range, loaded now so that
    // their latency overlaps the window's (they are used after two barriers)
    const int tsx = ts[d.tab_x + xc];
    const float4 wxv = tw[d.tab_x + xc];
    const float2 cl = clamp[i];
    // the tile's y tables (rows clamped to the image) once into LDS
    if (threadIdx.x < TY) {
        const int yy = min(y0 + (int)threadIdx.x, d.ny - 1);
        wy_s[threadIdx.x] = tw[d.tab_y + yy];
        sy_s[threadIdx.x] = ts[d.tab_y + yy] - ky0;
    }
    for (int e = threadIdx.x; e < wxn * wyn; e += kBlock) {
        const int ky = ky0 + e / wxn, kx = kx0 + e % wxn;
        win[e] = (kx < d.cx && ky < d.cy) ? coef2(P, (long long)ky * d.cx + kx) : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    // displacement contracted along y once per tile row and knot column
    // (into the box memory, which is free until the box fill)
    float2 *vy = (float2 *)region;
    {
        const float iwx = 1.0f / (float)wxn;
        for (int e = threadIdx.x; e < TY * wxn; e += kBlock) {
            const int yl = small_div(e, iwx), kx = e - yl * wxn;
            const float4 wyv = wy_s[yl];
            const float wys[4] = {wyv.x, wyv.y, wyv.z, wyv.w};
            const float2 *c = win + sy_s[yl] * wxn + kx;
            float vx = 0.0f, vyy = 0.0f;
#pragma unroll
            for (int bb = 0; bb < 4; ++bb) {
                const float2 q = c[bb * wxn];
                vx = fmaf(wys[bb], q.x, vx);
                vyy = fmaf(wys[bb], q.y, vyy);
            }
            vy[e] = make_float2(vx, vyy);
        }
    }
    __syncthreads();
    const int sx = tsx - kx0;
    const float wxs[4] = {wxv.x, wxv.y, wxv.z, wxv.w};
    float px[RPT], py[RPT];
    float mnx = FLT_MAX, mny = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int yl = ly + k * (kBlock / TX);
        const int y = y0 + yl;
        const float2 *wr = vy + yl * wxn + sx;
        float ux = 0.0f, uy = 0.0f;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const float2 q = wr[a];
            ux = fmaf(wxs[a], q.x, ux);
            uy = fmaf(wxs[a], q.y, uy);
        }
        px[k] = clamp_pos((float)x - ux);
        py[k] = clamp_pos((float)y - uy);
        if (xvalid && y < d.ny) {
            mnx = fminf(mnx, px[k]);
            mxx = fmaxf(mxx, px[k]);
            mny = fminf(mny, py[k]);
            mxy = fmaxf(mxy, py[k]);
        }
    }
    int rx0 = 0, ry0 = 0, RX = 0;
    bool use_lds = false;
    const float *Mi = M + d.img_off;
    if constexpr (BOX) {
        // wave extremes by the interleaved DPP chain (the y pair twice: it
        // takes three of each); positions are clamped, never NaN
        float mny2 = mny, mxy2 = mxy;
        wave_min3_max3(mnx, mny, mny2, mxx, mxy, mxy2);
        const int wid = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            red[0][wid] = mnx;
            red[1][wid] = mny;
            red[2][wid] = mxx;
            red[3][wid] = mxy;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kBlock / 64; ++k) {
            mnx = fminf(mnx, red[0][k]);
            mny = fminf(mny, red[1][k]);
            mxx = fmaxf(mxx, red[2][k]);
            mxy = fmaxf(mxy, red[3][k]);
        }
        rx0 = fastfloor16(mnx) - LO;
        ry0 = fastfloor16(mny) - LO;
        RX = fastfloor16(mxx) - LO + NS - rx0;
        const int RY = fastfloor16(mxy) - LO + NS - ry0;
        use_lds = (long long)RX * RY <= kW2Region;
        if (use_lds) {
            fill_box2(region, Mi, d.nx, d.ny, rx0, ry0, RX, RY);
            __syncthreads();
        }
    }
    if (!xvalid)
        return;
    float *Wi = W + d.img_off;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int y = y0 + ly + k * (kBlock / TX);
        if (y < d.ny) {
            float v;
            if (use_lds) {
                float wa[NS], wb[NS];
                const int s0 = kweights<DEG, false>(px[k], wa, nullptr);
                const int s1 = kweights<DEG, false>(py[k], wb, nullptr);
                v = clamp_val(contract_lds2<DEG>(region, RX, s0 - rx0, s1 - ry0, wa, wb), cl.x, cl.y);
            } else {
                v = interp2<DEG, false>(Mi, d.nx, d.ny, px[k], py[k], cl.x, cl.y, nullptr);
            }
            Wi[(long long)y * d.nx + x] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// stage 1: residual, force, SSD partial, first contraction of the adjoint
// ---------------------------------------------------------------------------

// Dynamic LDS layout of the stage-1 kernels (offsets in floats):
//   acc  [nx][4 * DIM + 1]  per-column band sums (t, component); the odd row
//                           stride keeps the column-parallel stores and the
//                           cell-parallel loads free of bank conflicts
//   xw   [nx] float4        x table weights
//   xs   [nx] int           x table start indices
//   cells[cx + 1] int       x band table
//   S    [ncx][4][4 * DIM]  per x band partial sums for each of the 4 weights
//   win  [4][cx + 4]        exact mode: coefficient window of the band (float4 / float2)
struct Stage1Layout {
    int acc_stride;
    size_t acc, xw, xs, cl, S, win, bytes;
};

__host__ __device__ inline Stage1Layout stage1_layout(int dim, int nx, int cx, int ncx, bool exact)
{
    Stage1Layout L;
    const int T = 4 * dim;
    L.acc_stride = T + 1;
    size_t f = 0;
    L.acc = f;
    f += (size_t)nx * L.acc_stride;
    f = (f + 3) & ~(size_t)3;
    L.xw = f;
    f += (size_t)4 * nx;
    L.xs = f;
    f += (size_t)nx;
    L.cl = f;
    f += (size_t)(cx + 1);
    L.S = f;
    f += (size_t)ncx * 4 * T;
    f = (f + 3) & ~(size_t)3;
    L.win = f;
    if (exact)
        f += (size_t)4 * (cx + 4) * (dim == 3 ? 4 : 2);
    L.bytes = f * 4;
    return L;
}

// load at a block-uniform row pointer plus a 32-bit byte offset: maps to a
// scalar base and one vector offset register, no 64-bit address arithmetic
__device__ __forceinline__ float ld_row(const float *__restrict__ row, unsigned byte_off)
{
    return *(const float *)((const char *)row + byte_off);
}

// rows of a band walked per iteration of the stage-1 loop: unrolled in
// reference mode (independent loads of consecutive rows in flight), not in
// exact mode (the interpolation body is large)
template <int MODE>
constexpr int kRowUnroll = MODE == 0 ? 4 : 1;

// x contraction of the per-column band sums into the block's partial-sum row:
// P[(t * cx + kx) * DIM + comp] = sum_x wx_{kx - start(x)}(x) * acc[x][t*DIM + comp].
// Two fixed-order steps: for each x band c and weight index a,
// S[c][a][tc] = sum over the band's columns in ascending x (one thread per
// band, weight index and half of the tc: one band lookup and one weight read
// per column for T / 2 sums); then P = sum_{a = 0..3} S[kx - a][a][tc].
template <int DIM>
__device__ __forceinline__ void contract_x(const float *acc_s, int AS, const float4 *xw_s, const int *cl_s,
                                                  float *S_s, int ncx, int cx, float *__restrict__ Pout)
{
    constexpr int T = 4 * DIM, TH = T / 2;
    static_assert(T % 2 == 0, "two halves");
    for (int it = threadIdx.x; it < ncx * 8; it += kBlock) {
        const int cell = it >> 3, a = (it >> 1) & 3, h = it & 1;
        const int xb = cl_s[cell], xe = cl_s[cell + 1];
        float sum[TH];
#pragma unroll
        for (int m = 0; m < TH; ++m)
            sum[m] = 0.0f;
        for (int xx = xb; xx < xe; ++xx) {
            const float w = ((const float *)&xw_s[xx])[a];
            const float *row = acc_s + AS * xx + h * TH;
#pragma unroll
            for (int m = 0; m < TH; ++m)
                sum[m] = fmaf(w, row[m], sum[m]);
        }
        float *dst = S_s + (cell * 4 + a) * T + h * TH;
#pragma unroll
        for (int m = 0; m < TH; ++m)
            dst[m] = sum[m];
    }
    __syncthreads();
    for (int o = threadIdx.x; o < T * cx; o += kBlock) {
        const int comp = o % DIM;
        const int q = o / DIM;
        const int kx = q % cx;
        const int tc = (q / cx) * DIM + comp;
        float sum = 0.0f;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int cell = kx - a;
            if (cell >= 0 && cell < ncx)
                sum += S_s[(cell * 4 + a) * T + tc];
        }
        Pout[o] = sum;
    }
}

// Stage 1, 3D: block = (z slice, y knot band).  Each thread owns x columns;
// rows of the band are walked with the y neighbours rotated through registers
// (reference mode).  Force f = grad * d with the reference's central
// differences of W, zero border components (mia/3d/image.cc:347-477) or, in
// exact mode, the analytic gradient of the interpolated moving image at
// x - u(x).  f * wy_t(j) accumulates in registers, the x contraction runs
// through LDS with the x table staged there.
template <int DEG, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_stage1_3d(const InstDesc *__restrict__ inst, int n_inst, const int *__restrict__ prefix, int nblocks,
                const double *__restrict__ params, const float4 *__restrict__ tw, const int *__restrict__ ts,
                const int *__restrict__ cells, const float *__restrict__ M, const float2 *__restrict__ clamp,
                const float *__restrict__ W, const float *__restrict__ R, float thr, float *__restrict__ Pout_all,
                double *__restrict__ ssd_part, unsigned *__restrict__ cnt_part)
{
    extern __shared__ float smem[];
    __shared__ double red[kBlock / 64];
    __shared__ unsigned redu[kBlock / 64];
    const int b = xcd_remap(blockIdx.x, nblocks);
    const int i = find_inst(prefix, n_inst, b);
    const InstDesc d = inst[i];
    if (MODE == 0 && d.zcell) // k_stage1z_3d's instance
        return;
    const int lb = b - prefix[i];
    const int band = lb % d.ncy;
    const int z = lb / d.ncy;
    const int j0 = cells[d.cell_y + band], j1 = cells[d.cell_y + band + 1];
    const int nx = d.nx, ny = d.ny, nz = d.nz, cx = d.cx;
    const Stage1Layout L = stage1_layout(3, nx, cx, d.ncx, MODE == 1);
    const int AS = L.acc_stride;
    float *acc_s = smem + L.acc;
    float4 *xw_s = (float4 *)(smem + L.xw);
    int *xs_s = (int *)(smem + L.xs);
    int *cl_s = (int *)(smem + L.cl);
    float *S_s = smem + L.S;
    float4 *win = (float4 *)(smem + L.win);
    const int wxn = cx + 4;
    for (int x = threadIdx.x; x < nx; x += kBlock) {
        xw_s[x] = tw[d.tab_x + x];
        xs_s[x] = ts[d.tab_x + x];
    }
    for (int k = threadIdx.x; k <= cx; k += kBlock)
        cl_s[k] = cells[d.cell_x + k];
    if constexpr (MODE == 1) {
        const int sz = ts[d.tab_z + z];
        const float4 wz = tw[d.tab_z + z];
        const float wzs[4] = {wz.x, wz.y, wz.z, wz.w};
        const double *P = params + d.par_off;
        for (int e = threadIdx.x; e < 4 * wxn; e += kBlock) {
            const int ky = band + e / wxn, kx = e % wxn;
            float ax = 0.0f, ay = 0.0f, az = 0.0f;
            if (kx < cx && ky < d.cy) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int kz = sz + c;
                    if (kz < d.cz) {
                        const float3 q = coef3(P, ((long long)kz * d.cy + ky) * cx + kx);
                        ax = fmaf(wzs[c], q.x, ax);
                        ay = fmaf(wzs[c], q.y, ay);
                        az = fmaf(wzs[c], q.z, az);
                    }
                }
            }
            win[e] = make_float4(ax, ay, az, 0.0f);
        }
    }
    __syncthreads();
    const long long plane = (long long)nx * ny;
    const float *Wi = W + d.img_off;
    const float *Ri = R + d.img_off;
    const float *Mi = M + d.img_off;
    const float2 cl = clamp[i];
    const bool iz = z > 0 && z < nz - 1;
    double ssd = 0.0;
    unsigned cnt = 0;
    for (int x = threadIdx.x; x < nx; x += kBlock) {
        float acc[12];
#pragma unroll
        for (int k = 0; k < 12; ++k)
            acc[k] = 0.0f;
        const int sx = xs_s[x];
        const float4 wxv = xw_s[x];
        const float wxs[4] = {wxv.x, wxv.y, wxv.z, wxv.w};
        const bool ix = x > 0 && x < nx - 1;
        const int xm = ix ? x - 1 : x, xp = ix ? x + 1 : x;
        const long long zoff = (long long)z * plane;
        const long long zm = iz ? zoff - plane : zoff, zp = iz ? zoff + plane : zoff;
        // Row pointers are block-uniform and the column indices are
        // non-negative 32-bit values, so every load below is a scalar base
        // plus one constant per-lane offset (no per-load address arithmetic).
        const unsigned ux = 4u * (unsigned)x, uxm = 4u * (unsigned)xm, uxp = 4u * (unsigned)xp; // byte offsets
        const float *Wz = Wi + zoff, *Wzm = Wi + zm, *Wzp = Wi + zp, *Rz = Ri + zoff;
        // rows j-1, j, j+1 of column x (reference mode)
        float wm = 0.0f, wc = 0.0f, wp = 0.0f;
        if constexpr (MODE == 0) {
            wc = ld_row(Wz + (long long)j0 * nx, ux);
            wm = ld_row(Wz + (long long)max(j0 - 1, 0) * nx, ux);
        }
#pragma unroll kRowUnroll<MODE>
        for (int j = j0; j < j1; ++j) {
            const long long ro = (long long)j * nx;
            const float4 wyv = tw[d.tab_y + j];
            const float wys[4] = {wyv.x, wyv.y, wyv.z, wyv.w};
            float w, gx, gy, gz;
            if constexpr (MODE == 0) {
                const bool iy = j > 0 && j < ny - 1;
                wp = ld_row(Wz + (long long)min(j + 1, ny - 1) * nx, ux);
                w = wc;
                // all four neighbours are loaded (clamped indices keep them in
                // the image) and the border rule is a select
                const float wxp = ld_row(Wz + ro, uxp), wxm = ld_row(Wz + ro, uxm);
                const float wzp = ld_row(Wzp + ro, ux), wzm = ld_row(Wzm + ro, ux);
                gx = ix ? 0.5f * (wxp - wxm) : 0.0f;
                gy = iy ? 0.5f * (wp - wm) : 0.0f;
                gz = iz ? 0.5f * (wzp - wzm) : 0.0f;
                wm = wc;
                wc = wp;
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
                w = interp3<DEG, true>(Mi, nx, ny, nz, clamp_pos((float)x - ux), clamp_pos((float)j - uy),
                                       clamp_pos((float)z - uz), cl.x, cl.y, g);
                gx = g[0];
                gy = g[1];
                gz = g[2];
            }
            const float r = ld_row(Rz + ro, ux);
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
            acc_s[AS * x + k] = acc[k];
    }
    // the block's SSD (and, with autothresh, mask count) as block_sum adds
    // them: wave sums, then the waves in order; their partials share the
    // barrier that completes the column sums (without autothresh the count
    // is not used and its partial is 0)
    const double sw = wave_sum(ssd);
    const unsigned cw = thr != 0.0f ? wave_sum_u(cnt) : 0u;
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = sw;
        redu[threadIdx.x >> 6] = cw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        unsigned c = 0;
        for (int k = 0; k < kBlock / 64; ++k) {
            s += red[k];
            c += redu[k];
        }
        ssd_part[b] = s;
        cnt_part[b] = c;
    }
    contract_x<3>(acc_s, AS, xw_s, cl_s, S_s, d.ncx, cx,
                  Pout_all + d.part_off + ((long long)z * d.ncy + band) * 12 * cx);
}

// register window of k_stage1z_3d (band rows + 2): y bands of up to 9 rows,
// i.e. spline rates up to about 9
constexpr int kS1zRows = 11;

// One float per lane from a buffer at a per-lane byte offset plus a scalar
// byte offset (the builtin returns the raw 32 bits)
__device__ __forceinline__ float ld_buf(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// x neighbours within a wave: lane l receives lane l + 1 (wave_shl:1) or
// lane l - 1 (wave_shr:1); the lane at the wave's edge receives `edge`
__device__ __forceinline__ float lane_next(float v, float edge)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x130, 0xf, 0xf, false));
}

__device__ __forceinline__ float lane_prev(float v, float edge)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x138, 0xf, 0xf, false));
}

// Stage 1, 3D, reference gradient, one block per (z knot cell, y knot band):
// the slices z of the cell (those whose first z knot is the cell, so that
// their four z knots are cell .. cell + 3) and the rows of the band; one
// thread per x column (nx <= kBlock).  Each thread keeps its column of W for
// the band's rows and the rows either side in registers for three slices
// (z - 1, z, z + 1) and walks z, loading at each step the next slice of W and
// the slice's rows of R (buffer loads at a scalar slice offset plus a scalar
// row offset fixed for the block).  Nothing is loaded ahead: the four waves
// per SIMD cover each other's loads, and the registers a prefetch would hold
// are what keeps the kernel free of spills.  The x neighbours come from the
// neighbouring lanes (DPP); the wave's edge lanes load the column beyond the
// wave themselves, so the waves of a block never wait for each other during
// the walk.  The gradient rule, residual and force are those of k_stage1_3d.
// f * wy_t(j) accumulates over the band's rows (y weights read from LDS per
// row); at the end of the slice the 12 sums (t, component) are folded with
// the slice's four z weights into 48 per column (k = kz - cell, t,
// component), kept in registers across the cell.  The x contraction then runs
// through LDS once per k with contract_x's two fixed-order steps.  Output of
// the block: P[cell][band][k][t][kx][comp] (48 * cx floats), so the reduce
// reads every partial sum once.  MAXR >= band rows + 2.  THR: autothresh on
// (the mask and its count; without it every voxel counts and no count is kept).
template <int DEG, int MAXR, bool THR>
__global__ void __launch_bounds__(kBlock, 4)
    k_stage1z_3d(const InstDesc *__restrict__ inst, int n_inst, const int *__restrict__ prefix, int nblocks,
                 const float4 *__restrict__ tw, const int *__restrict__ ts, const int *__restrict__ cells,
                 const float *__restrict__ W, const float *__restrict__ R, float thr, float *__restrict__ Pout_all,
                 double *__restrict__ ssd_part, unsigned *__restrict__ cnt_part)
{
    constexpr int NW = kBlock / 64;
    extern __shared__ float smem[];
    __shared__ double red[NW];
    __shared__ unsigned redu[NW];
    __shared__ float4 wy_s[MAXR];            // y weights of the window rows (read per row, not held)
    const int b = xcd_remap(blockIdx.x, nblocks);
    const int i = find_inst(prefix, n_inst, b);
    const InstDesc d = inst[i];
    if (!d.zcell) // k_stage1_3d's instance
        return;
    const int lb = b - prefix[i];
    const int band = lb % d.ncy;
    const int cz = lb / d.ncy;
    const int j0 = cells[d.cell_y + band], j1 = cells[d.cell_y + band + 1];
    const int zb = cells[d.cell_z + cz], ze = cells[d.cell_z + cz + 1];
    const int nx = d.nx, ny = d.ny, nz = d.nz, cx = d.cx;
    const Stage1Layout L = stage1_layout(3, nx, cx, d.ncx, false);
    const int AS = L.acc_stride;
    float *acc_s = smem + L.acc;
    float4 *xw_s = (float4 *)(smem + L.xw);
    int *cl_s = (int *)(smem + L.cl);
    float *S_s = smem + L.S;
    for (int x = threadIdx.x; x < nx; x += kBlock)
        xw_s[x] = tw[d.tab_x + x];
    for (int k = threadIdx.x; k <= cx; k += kBlock)
        cl_s[k] = cells[d.cell_x + k];
    const int nrw = j1 - j0 + 2; // window rows j0 - 1 .. j1 (clamped to the image)
    if (threadIdx.x >= 1 && threadIdx.x + 1 < nrw)
        wy_s[threadIdx.x] = tw[d.tab_y + j0 + threadIdx.x - 1];
    const int x = threadIdx.x;
    const int lane = x & 63;
    const bool ix = x > 0 && x < nx - 1;
    const __amdgpu_buffer_rsrc_t rwi = raw_rsrc(W + d.img_off), rri = raw_rsrc(R + d.img_off);
    const unsigned vo = (unsigned)min(x, nx - 1) * 4u;
    // the column beyond the wave for its edge lanes (x - 1 for lane 0, x + 1
    // for lane 63; the others read their own column), clamped to the row on
    // both sides: lane 0 of a wave that starts beyond the row (nx <= 64 w)
    // would otherwise read past the row, and on the image's last row past
    // the image
    const unsigned veo = lane == 0 ? (unsigned)min(max(x - 1, 0), nx - 1) * 4u
                                   : lane == 63 ? (unsigned)min(x + 1, nx - 1) * 4u : vo;
    // byte offset of window row r within a slice (rows clamped; the same for
    // every slice of the block, so these stay in scalar registers) and of
    // slice z (clamped): each load adds the two
    const unsigned planeb = (unsigned)(nx * ny) * 4u;
    unsigned roff[MAXR];
#pragma unroll
    for (int r = 0; r < MAXR; ++r)
        roff[r] = (unsigned)(min(max(j0 - 1 + r, 0), ny - 1) * nx) * 4u;
    auto zoff = [&](int z) { return (unsigned)min(max(z, 0), nz - 1) * planeb; };
    auto soff = [&](int z, int r) { return zoff(z) + roff[r]; };
    float wm[MAXR], wc[MAXR], wp[MAXR], rc[MAXR];
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
        wm[r] = wc[r] = wp[r] = rc[r] = 0.0f;
        if (r < nrw) {
            wm[r] = ld_buf(rwi, vo, soff(zb - 1, r));
            wc[r] = ld_buf(rwi, vo, soff(zb, r));
            wp[r] = ld_buf(rwi, vo, soff(zb + 1, r));
        }
    }
    double ssd = 0.0;
    unsigned cnt_w = 0; // the wave's count, a scalar
    float a48[48];
#pragma unroll
    for (int k = 0; k < 48; ++k)
        a48[k] = 0.0f;
    // lanes beyond the image count nothing and their force is never stored
    const bool xin = x < nx;
    for (int z = zb; z < ze; ++z) {
        // this step's loads: W of slice z + 1 (the cell's first step has
        // it from the prologue), R of slice z and the edge lanes' columns
        if (z > zb) {
#pragma unroll
            for (int r = 0; r < MAXR; ++r)
                if (r < nrw)
                    wp[r] = ld_buf(rwi, vo, soff(z + 1, r));
        }
#pragma unroll
        for (int r = 1; r + 1 < MAXR; ++r)
            if (r + 1 < nrw)
                rc[r] = ld_buf(rri, vo, soff(z, r));
        // lane 0: W at x - 1, lane 63: W at x + 1 (the DPP shifts below use
        // it only there; at the image's edges the x difference is zeroed)
        float ev[MAXR];
#pragma unroll
        for (int r = 1; r + 1 < MAXR; ++r)
            ev[r] = r + 1 < nrw ? ld_buf(rwi, veo, soff(z, r)) : 0.0f;
        if (z == zb)
            __syncthreads(); // the y weights in LDS
        const float4 wzv = tw[d.tab_z + z];
        const float wzs[4] = {wzv.x, wzv.y, wzv.z, wzv.w};
        const bool iz = z > 0 && z < nz - 1;

        float acc[12];
#pragma unroll
        for (int k = 0; k < 12; ++k)
            acc[k] = 0.0f;
#pragma unroll
        for (int r = 1; r + 1 < MAXR; ++r) {
            if (r + 1 < nrw) {
                const int j = j0 + r - 1;
                const float w = wc[r];
                const float e = ev[r];
                const float wxm = lane_prev(w, e), wxp = lane_next(w, e);
                const float4 wyv = wy_s[r];
                const float wys[4] = {wyv.x, wyv.y, wyv.z, wyv.w};
                const bool iy = j > 0 && j < ny - 1;
                // without autothresh every voxel counts and the count is unused
                const bool on = THR ? xin && w > thr : xin;
                const float dl = on ? w - rc[r] : 0.0f;
                ssd += (double)dl * (double)dl;
                if constexpr (THR)
                    cnt_w += (unsigned)__popcll(__ballot(on));
                // force = gradient * difference with the reference's 0.5 on
                // the difference: the same products as (0.5 * (w+ - w-)) * dl,
                // 0.5 being exact; the y and z components are zero on the
                // image border, their sums are skipped there (uniform branches)
                const float hdl = 0.5f * dl;
                const float fx = (wxp - wxm) * (ix ? hdl : 0.0f);
                const float fy = (wc[r + 1] - wc[r - 1]) * (iy ? hdl : 0.0f);
                const float fz = (wp[r] - wm[r]) * (iz ? hdl : 0.0f);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    acc[3 * t + 0] = fmaf(wys[t], fx, acc[3 * t + 0]);
                    acc[3 * t + 1] = fmaf(wys[t], fy, acc[3 * t + 1]);
                    acc[3 * t + 2] = fmaf(wys[t], fz, acc[3 * t + 2]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int m = 0; m < 12; ++m)
                a48[12 * k + m] = fmaf(wzs[k], acc[m], a48[12 * k + m]);
        // (rows 0 and MAXR - 1 of the slice below and of R are never read)
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            if (r >= 1 && r + 1 < MAXR)
                wm[r] = wc[r];
            wc[r] = wp[r];
        }
    }
    // the block's SSD and the waves' counts, added as block_sum adds them
    // (wave sums, then the waves in order) by thread 0 after the first pass's
    // barrier
    const double sw = wave_sum(ssd);
    if (lane == 0) {
        red[threadIdx.x >> 6] = sw;
        redu[threadIdx.x >> 6] = cnt_w;
    }
    float *Pb = Pout_all + d.part_off + ((long long)cz * d.ncy + band) * 48 * cx;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k > 0)
            __syncthreads(); // the previous pass is done with acc_s and S_s
        if (x < nx)
#pragma unroll
            for (int m = 0; m < 12; ++m)
                acc_s[AS * x + m] = a48[12 * k + m];
        __syncthreads();
        if (k == 0 && threadIdx.x == 0) {
            double s = 0.0;
            unsigned c = 0;
            for (int w = 0; w < NW; ++w) {
                s += red[w];
                c += redu[w];
            }
            ssd_part[b] = s;
            cnt_part[b] = c;
        }
        contract_x<3>(acc_s, AS, xw_s, cl_s, S_s, d.ncx, cx, Pb + 12 * k * cx);
    }
}

// Stage 1, 2D: block = y knot band.
template <int DEG, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_stage1_2d(const InstDesc *__restrict__ inst, int n_inst, const int *__restrict__ prefix,
                const int4 *__restrict__ blocks, int nblocks,
                const double *__restrict__ params, const float4 *__restrict__ tw, const int *__restrict__ ts,
                const int *__restrict__ cells, const float *__restrict__ M, const float2 *__restrict__ clamp,
                const float *__restrict__ W, const float *__restrict__ R, float thr, float *__restrict__ Pout_all,
                double *__restrict__ ssd_part, unsigned *__restrict__ cnt_part)
{
    extern __shared__ float smem[];
    __shared__ double red[kBlock / 64];
    __shared__ unsigned redu[kBlock / 64];
    const int b = xcd_remap(blockIdx.x, nblocks);
    // the block's instance, band and rows: from the block table in one load,
    // or derived from the block prefix and band table (dependent loads)
    int i, band, j0, j1;
    if (blocks) {
        const int4 e = blocks[b];
        i = e.x;
        band = e.y;
        j0 = e.z;
        j1 = e.w;
    } else {
        i = find_inst(prefix, n_inst, b);
        band = b - prefix[i];
    }
    const InstDesc d = inst[i];
    if (!blocks) {
        j0 = cells[d.cell_y + band];
        j1 = cells[d.cell_y + band + 1];
    }
    const int nx = d.nx, ny = d.ny, cx = d.cx;
    const Stage1Layout L = stage1_layout(2, nx, cx, d.ncx, MODE == 1);
    const int AS = L.acc_stride;
    float *acc_s = smem + L.acc;
    float4 *xw_s = (float4 *)(smem + L.xw);
    int *xs_s = (int *)(smem + L.xs);
    int *cl_s = (int *)(smem + L.cl);
    float *S_s = smem + L.S;
    float2 *win = (float2 *)(smem + L.win);
    const int wxn = cx + 4;
    for (int x = threadIdx.x; x < nx; x += kBlock) {
        xw_s[x] = tw[d.tab_x + x];
        xs_s[x] = ts[d.tab_x + x];
    }
    for (int k = threadIdx.x; k <= cx; k += kBlock)
        cl_s[k] = cells[d.cell_x + k];
    if constexpr (MODE == 1) {
        const double *P = params + d.par_off;
        for (int e = threadIdx.x; e < 4 * wxn; e += kBlock) {
            const int ky = band + e / wxn, kx = e % wxn;
            win[e] = (kx < cx && ky < d.cy) ? coef2(P, (long long)ky * cx + kx) : make_float2(0.0f, 0.0f);
        }
    }
    __syncthreads();
    const float *Wi = W + d.img_off;
    const float *Ri = R + d.img_off;
    const float *Mi = M + d.img_off;
    const float2 cl = clamp[i];
    double ssd = 0.0;
    unsigned cnt = 0;
    const int lane = threadIdx.x & 63;
    for (int x = threadIdx.x; x < nx; x += kBlock) {
        float acc[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            acc[k] = 0.0f;
        const int sx = xs_s[x];
        const float4 wxv = xw_s[x];
        const float wxs[4] = {wxv.x, wxv.y, wxv.z, wxv.w};
        const bool ix = x > 0 && x < nx - 1;
        // the column beyond the wave for its edge lanes (x - 1 for lane 0,
        // x + 1 for lane 63, clamped to the row; the others read their own):
        // the x neighbours come from the neighbouring lanes (DPP)
        const long long xe = lane == 0 ? max(x - 1, 0) : lane == 63 ? min(x + 1, nx - 1) : x;
        // one row's residual, force and its y-weighted sums (both modes)
        auto add_row = [&](float w, float gx, float gy, float r, const float *wys) {
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
        };
        if constexpr (MODE == 0) {
            // rows four at a time: their loads (clamped rows) first, then the
            // rows in order; W of the column rolls through wm, wc
            float wc = Wi[(long long)j0 * nx + x];
            float wm = Wi[(long long)max(j0 - 1, 0) * nx + x];
            for (int j = j0; j < j1; j += 4) {
                float wpv[4], ev[4], rv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int jj = min(j + u, j1 - 1);
                    wpv[u] = Wi[(long long)min(jj + 1, ny - 1) * nx + x];
                    ev[u] = Wi[(long long)jj * nx + xe];
                    rv[u] = Ri[(long long)jj * nx + x];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (j + u < j1) {
                        const int jr = j + u;
                        const float4 wyv = tw[d.tab_y + jr];
                        const float wys[4] = {wyv.x, wyv.y, wyv.z, wyv.w};
                        const bool iy = jr > 0 && jr < ny - 1;
                        // every lane takes part in the shifts (a lane left out
                        // would hand its neighbours the edge value instead of
                        // its column): the opaque use keeps them out of the
                        // border lanes' branch below
                        float wxp = lane_next(wc, ev[u]), wxm = lane_prev(wc, ev[u]);
                        asm volatile("" : "+v"(wxp), "+v"(wxm));
                        const float gx = ix ? 0.5f * (wxp - wxm) : 0.0f;
                        const float gy = iy ? 0.5f * (wpv[u] - wm) : 0.0f;
                        add_row(wc, gx, gy, rv[u], wys);
                        wm = wc;
                        wc = wpv[u];
                    }
            }
        } else {
            for (int j = j0; j < j1; ++j) {
                const float4 wyv = tw[d.tab_y + j];
                const float wys[4] = {wyv.x, wyv.y, wyv.z, wyv.w};
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
                const float w = interp2<DEG, true>(Mi, nx, ny, clamp_pos((float)x - ux), clamp_pos((float)j - uy),
                                                   cl.x, cl.y, g);
                add_row(w, g[0], g[1], Ri[(long long)j * nx + x], wys);
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            acc_s[AS * x + k] = acc[k];
    }
    // the block's SSD (and, with autothresh, mask count) as block_sum adds
    // them: wave sums, then the waves in order; their partials share the
    // barrier that completes the column sums (without autothresh the count
    // is not used and its partial is 0)
    const double sw = wave_sum(ssd);
    const unsigned cw = thr != 0.0f ? wave_sum_u(cnt) : 0u;
    if (lane == 0) {
        red[threadIdx.x >> 6] = sw;
        redu[threadIdx.x >> 6] = cw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        unsigned c = 0;
        for (int k = 0; k < kBlock / 64; ++k) {
            s += red[k];
            c += redu[k];
        }
        ssd_part[b] = s;
        cnt_part[b] = c;
    }
    contract_x<2>(acc_s, AS, xw_s, cl_s, S_s, d.ncx, cx, Pout_all + d.part_off + (long long)band * 8 * cx);
}

// ---------------------------------------------------------------------------
// finalize and reduce
// ---------------------------------------------------------------------------

// Per instance: SSD sum and mask count in a fixed order, cost and gradient
// scale (mia/template/ssd.hh:54-99,125-181; weight:
// mia/template/fullcost.cxx:36-45).  Run by thread block `i` of either kernel.
// This thread's share of an instance's SSD (and mask count) partials
// [beg, end) as the finalize adds them: thread t < kBlock takes partials
// t, t + kBlock, ..., four in flight before they are added (same order; a
// missing one adds 0).
template <bool COUNT>
__device__ __forceinline__ void ssd_thread_share(int beg, int end, const double *__restrict__ ssd_part,
                                                 const unsigned *__restrict__ cnt_part, double &s, unsigned &c)
{
    s = 0.0;
    c = 0;
    for (int k0 = beg + (int)threadIdx.x; k0 < end; k0 += 4 * kBlock) {
        double a[4];
        unsigned n[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            // unconditional loads from a clamped index (a load under a
            // branch is waited for at the branch's end)
            const int k = k0 + u * kBlock;
            const int kc = min(k, end - 1);
            const double av = ssd_part[kc];
            a[u] = k < end ? av : 0.0;
            if constexpr (COUNT) {
                const unsigned nv = cnt_part[kc];
                n[u] = k < end ? nv : 0u;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            s += a[u];
            if constexpr (COUNT)
                c += n[u];
        }
    }
}

// Gradient component e3 of a 2D instance from the stage-1 partial sums: the
// four y bands that touch its knot added in FP64 in this fixed order (the
// adjoint of spline.cc:731-791 completed along y), times sc, plus the
// instance's penalty gradient.  The four partial sums are loaded first
// (unconditional loads from a clamped band; a missing one then reads as 0).
__device__ __forceinline__ double reduce2_elem(const InstDesc &d, const float *__restrict__ P, long long e3,
                                               double sc, const PenaltyIn &pen, int i)
{
    int comp, kx, ky;
    if (2LL * d.cx * d.cy < (1LL << 31)) { // 32-bit index arithmetic (instance-uniform)
        const unsigned ue = (unsigned)e3 >> 1;
        comp = (int)(e3 & 1);
        ky = (int)(ue / (unsigned)d.cx);
        kx = (int)(ue - (unsigned)ky * (unsigned)d.cx);
    } else {
        comp = (int)(e3 % 2);
        const long long e = e3 / 2;
        kx = (int)(e % d.cx);
        ky = (int)(e / d.cx);
    }
    float v[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int band = ky - t;
        const int bc = min(max(band, 0), d.ncy - 1);
        const float q = P[d.part_off + (long long)bc * 8 * d.cx + (t * d.cx + kx) * 2 + comp];
        v[t] = band >= 0 && band < d.ncy ? q : 0.0f;
    }
    double g = 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t)
        g += v[t];
    return grad_with_penalty(sc, g, pen, i, d.par_off + e3);
}

__device__ __forceinline__ void finalize_instance(const InstDesc &d, int i, SsdIndex si,
                                                  const double *__restrict__ ssd_part,
                                                  const unsigned *__restrict__ cnt_part, int norm, float thr,
                                                  double weight, double *__restrict__ costs,
                                                  double *__restrict__ gscale, const PenaltyIn &pen)
{
    __shared__ double red[kBlock / 64];
    __shared__ unsigned redu[kBlock / 64];
    const bool z = si.zmode && d.zcell;
    const int beg = z ? si.zoff + si.s1z[i] : si.s1[i], end = z ? si.zoff + si.s1z[i + 1] : si.s1[i + 1];
    double s;
    unsigned c = 0;
    // the mask count only matters with autothresh
    if (thr != 0.0f) {
        ssd_thread_share<true>(beg, end, ssd_part, cnt_part, s, c);
        c = block_sum_u(c, redu);
    } else {
        ssd_thread_share<false>(beg, end, ssd_part, cnt_part, s, c);
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
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
        {
#pragma clang fp contract(off)
            double f = weight * cost;
            if (pen.on && pen.on[i])
                f = f + pen.w[i] * pen.value[i];
            costs[i] = f;
        }
        if (gscale)
            gscale[i] = -sc * weight;
    }
}

__global__ void __launch_bounds__(kBlock)
    k_finalize(const InstDesc *__restrict__ inst, SsdIndex si, const double *__restrict__ ssd_part,
               const unsigned *__restrict__ cnt_part, int norm, float thr, double weight, double *__restrict__ costs,
               double *__restrict__ gscale, PenaltyIn pen)
{
    const int i = blockIdx.x;
    finalize_instance(inst[i], i, si, ssd_part, cnt_part, norm, thr, weight, costs, gscale, pen);
}

// Per coefficient component: gradient = scale * sum_z wz(z) sum_t
// P[z][ky - t][t][kx][comp] in FP64, fixed order (the adjoint of
// spline.cc:731-791 completed along y and z).  FUSED: the scale is
// -weight * (norm ? 1/N : 1), known without the mask count (no autothresh);
// otherwise it comes from k_finalize.  The first nfin blocks (0 or n_inst) do
// the finalize work.  Instances that k_reducez_3d handles are skipped.
template <int DIM, bool FUSED>
__global__ void __launch_bounds__(kBlock)
    k_reduce(const InstDesc *__restrict__ inst, int n_inst, const int *__restrict__ prefix,
             const int2 *__restrict__ blocks, const float *__restrict__ P, const float4 *__restrict__ tw, const int *__restrict__ ts,
             const int *__restrict__ cells, SsdIndex si, int nfin, const double *__restrict__ ssd_part,
             const unsigned *__restrict__ cnt_part, int norm, float thr, double weight, double *__restrict__ costs,
             const double *__restrict__ gscale, double *__restrict__ grad, PenaltyIn pen)
{
    int b = blockIdx.x;
    if (b < nfin) {
        finalize_instance(inst[b], b, si, ssd_part, cnt_part, norm, thr, weight, costs, nullptr, pen);
        return;
    }
    b -= nfin;
    // the block's instance and its block within the instance: from the block
    // table in one load (2D), or from the block prefix
    int i, lb;
    if (blocks) {
        const int2 e = blocks[b];
        i = e.x;
        lb = e.y;
    } else {
        i = find_inst(prefix, n_inst, b);
        lb = b - prefix[i];
    }
    const InstDesc d = inst[i];
    if (si.zmode && d.zcell)
        return;
    const long long ncoef = (long long)d.cx * d.cy * d.cz;
    double sc;
    if constexpr (FUSED)
        sc = -weight * (norm ? 1.0 / ((double)d.nx * d.ny * d.nz) : 1.0);
    else
        sc = gscale[i];
    if constexpr (DIM == 3) {
        // 64 outputs per block; wave g sums the g-th quarter of the z support,
        // the quarters are added in order through LDS
        __shared__ double part[kBlock / 64][64];
        const int lane = threadIdx.x & 63, g4 = threadIdx.x >> 6;
        const long long e3 = (long long)lb * 64 + lane;
        double g = 0.0;
        if (e3 < 3 * ncoef) {
            const int comp = (int)(e3 % 3);
            const long long e = e3 / 3;
            const int kx = (int)(e % d.cx);
            const int ky = (int)((e / d.cx) % d.cy);
            const int kz = (int)(e / ((long long)d.cx * d.cy));
            const int zb = cells[d.cell_z + max(kz - 3, 0)], ze = cells[d.cell_z + kz + 1];
            const int len = ze - zb;
            const int za = zb + (len * g4) / 4, zc = zb + (len * (g4 + 1)) / 4;
            const float *twf = (const float *)tw;
            const long long row = 12LL * d.cx;
            // kU slices at a time: their weights and partial sums are loaded
            // before any is added (a slice at a time waits out the memory
            // latency once per slice), then added in the same z and t order
            constexpr int kU = 8;
            for (int z0 = za; z0 < zc; z0 += kU) {
                float v[kU][4], wzf[kU];
                int tsz[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u)
                    tsz[u] = ts[d.tab_z + min(z0 + u, zc - 1)];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int z = min(z0 + u, zc - 1);
                    wzf[u] = twf[4 * (d.tab_z + z) + (kz - tsz[u])];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        // unconditional loads from a clamped band (a load
                        // under a branch is waited for at the branch's end)
                        const int band = min(max(ky - t, 0), d.ncy - 1);
                        v[u][t] = P[d.part_off + ((long long)z * d.ncy + band) * row + (t * d.cx + kx) * 3 + comp];
                    }
                }
#pragma unroll
                for (int u = 0; u < kU; ++u)
                    if (z0 + u < zc) {
                        double s = 0.0;
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const int band = ky - t;
                            if (band >= 0 && band < d.ncy)
                                s += v[u][t];
                        }
                        g += (double)wzf[u] * s;
                    }
            }
        }
        part[g4][lane] = g;
        __syncthreads();
        if (g4 == 0 && e3 < 3 * ncoef)
            grad[d.par_off + e3] = grad_with_penalty(
                sc, ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane], pen, i, d.par_off + e3);
        return;
    } else {
        const long long e3 = (long long)lb * kBlock + threadIdx.x;
        if (e3 >= 2 * ncoef)
            return;
        grad[d.par_off + e3] = reduce2_elem(d, P, e3, sc, pen, i);
    }
}

// Reduce of k_stage1z_3d's partial sums, one gradient component per thread:
// gradient(kz, ky, kx, comp) = scale * sum_{k < 4} sum_{t < 4}
// P[kz - k][ky - t][k][t][kx][comp], in FP64 and this fixed order (the
// adjoint of spline.cc:731-791 completed along y and z).  Every partial sum
// is read once.  FUSED and nfin as in k_reduce; instances without zcell are
// skipped.
template <bool FUSED>
__global__ void __launch_bounds__(kBlock)
    k_reducez_3d(const InstDesc *__restrict__ inst, int n_inst, const int *__restrict__ prefix,
                 const float *__restrict__ P, SsdIndex si, int nfin, const double *__restrict__ ssd_part,
                 const unsigned *__restrict__ cnt_part, int norm, float thr, double weight,
                 double *__restrict__ costs, const double *__restrict__ gscale, double *__restrict__ grad,
                 PenaltyIn pen)
{
    int b = blockIdx.x;
    if (b < nfin) {
        finalize_instance(inst[b], b, si, ssd_part, cnt_part, norm, thr, weight, costs, nullptr, pen);
        return;
    }
    b -= nfin;
    const int i = find_inst(prefix, n_inst, b);
    const InstDesc d = inst[i];
    if (!d.zcell)
        return;
    // 32-bit index arithmetic: the set-up keeps images below 2^29 voxels, and
    // the coefficient grid is smaller than the image
    const int ncoef3 = 3 * d.cx * d.cy * d.cz;
    const int e3 = (b - prefix[i]) * kBlock + threadIdx.x;
    if (e3 >= ncoef3)
        return;
    double sc;
    if constexpr (FUSED)
        sc = -weight * (norm ? 1.0 / ((double)d.nx * d.ny * d.nz) : 1.0);
    else
        sc = gscale[i];
    const unsigned ue = (unsigned)e3 / 3u;
    const int comp = e3 - 3 * (int)ue;
    const unsigned row = ue / (unsigned)d.cx;
    const int kx = (int)(ue - row * (unsigned)d.cx);
    const int kz = (int)(row / (unsigned)d.cy);
    const int ky = (int)(row - (unsigned)kz * (unsigned)d.cy);
    const float *Pi = P + d.part_off + 3 * kx + comp;
    const long long cell_stride = 48LL * d.cx;
    // all 16 partial sums in flight first (unconditional loads from a
    // clamped cell and band, a missing one then reads as 0), then the sum in
    // the fixed (k, t) order
    float v[16];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int cz = kz - k;
        const int czc = min(max(cz, 0), d.ncz - 1);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int band = ky - t;
            const int bc = min(max(band, 0), d.ncy - 1);
            const bool in = cz >= 0 && cz < d.ncz && band >= 0 && band < d.ncy;
            const float q = Pi[((long long)czc * d.ncy + bc) * cell_stride + (4 * k + t) * 3LL * d.cx];
            v[4 * k + t] = in ? q : 0.0f;
        }
    }
    double g = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q)
        g += (double)v[q];
    grad[d.par_off + e3] = grad_with_penalty(sc, g, pen, i, d.par_off + e3);
}

// ---------------------------------------------------------------------------
// set-up of the allocation-free evaluation
// ---------------------------------------------------------------------------

// Block counts of a one-instance batch (the totals of the per-instance block
// prefixes that the host set-up builds).
struct OneInstCounts {
    int s1, red, s1z, redz, tiles;
};

// Layout of the small integer arrays of a one-instance batch: the four block
// prefixes {0, total}, then the warp launch's instance list {0} and tile
// prefix {0, tiles}.
constexpr int kOneInstInts = 11;

// One block writes what the host set-up builds and copies for a batch of one
// instance: the instance description, the block prefixes, the warp launch's
// lists and the spline axis tables (spline_axis_entry, the host's arithmetic)
// with their band tables.  Kernel arguments only, so the sequence can be
// captured in a graph.
__global__ void __launch_bounds__(kBlock)
    k_one_inst_setup(InstDesc d, int dim, OneInstCounts nc, InstDesc *__restrict__ inst, int *__restrict__ ints,
                     float4 *__restrict__ tw, int *__restrict__ ts, int *__restrict__ cells)
{
    if (threadIdx.x == 0) {
        *inst = d;
        const int v[kOneInstInts] = {0, nc.s1, 0, nc.red, 0, nc.s1z, 0, nc.redz, 0, 0, nc.tiles};
        for (int k = 0; k < kOneInstInts; ++k)
            ints[k] = v[k];
    }
    const int ns[3] = {d.nx, d.ny, d.nz}, cs[3] = {d.cx, d.cy, d.cz};
    const int toff[3] = {d.tab_x, d.tab_y, d.tab_z}, coff[3] = {d.cell_x, d.cell_y, d.cell_z};
    for (int a = 0; a < 3; ++a) {
        if (a < dim) {
            const int n = ns[a], c = cs[a];
            for (int i = threadIdx.x; i < n; i += kBlock) {
                float w[4];
                ts[toff[a] + i] = spline_axis_entry(i, n, c, w);
                tw[toff[a] + i] = make_float4(w[0], w[1], w[2], w[3]);
            }
            // band table: first sample whose start index is >= k (n if none);
            // the start indices ascend, so a binary search over recomputed
            // entries finds it without reading another thread's stores
            for (int k = threadIdx.x; k <= c; k += kBlock) {
                int lo = 0, hi = n;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    float w[4];
                    if (spline_axis_entry(mid, n, c, w) < k)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                cells[coff[a] + k] = lo;
            }
        } else if (threadIdx.x == 0) {
            // the z axis of a 2D instance: one sample with weight 1, one band
            ts[toff[a]] = 0;
            tw[toff[a]] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
            cells[coff[a]] = 0;
            cells[coff[a] + 1] = 1;
        }
    }
}

} // namespace mia_hip

// On-device batched L-BFGS: after each batched cost + gradient evaluation,
// one step of the minimizer runs for every instance, so a whole
// optimisation iteration (warp, stage 1, reduce, step) stays on the GPU and can
// be captured in a graph.  The algorithm is the one of the host
// LBFGSMinimizer (mia_amd/host/minimizer.hh): the L-BFGS direction (written
// in the compact form of Byrd, Nocedal & Schnabel 1994, equal to the host's
// two-loop recursion of Nocedal & Wright, Alg. 7.4/7.5), backtracking Armijo
// line search with safeguarded quadratic interpolation, first step scaled to
// move no parameter by more than `first_step`, `maxeval` counting every
// evaluation.
//
// Instances of up to kLbRegN parameters take the whole step in one block
// (k_lbfgs_step_small, at the end of this file): for them a step is a chain
// of short dependent phases that one block runs faster than several launches.
//
// Larger instances: the parameter vector is cut into segments of kSeg
// elements, one block each, over three launches, so that a large instance (44,616 parameters at
// the full level of the 3D registration) streams its pair history through
// many CUs instead of one.  No block waits for or signals another: each pass
// writes per-segment partial sums, and in the next pass every block of the
// instance adds them up in ascending segment order and takes the same scalar
// decisions (the launch boundary makes the sums visible).  An instance's
// segments depend only on its own size, so its trajectory does not depend on
// the batch it runs in.
//
//   pass 1: the evaluated point's element updates (best point, X/G, the
//           candidate pair s, y in the ring's spare slot, or the backtracking
//           point) and partial sums of every product the step needs (s.y,
//           y.y, s.s, s.g, y.g, max |g|, and s_c.y, y_c.y, s_c.g, y_c.g per
//           stored pair c);
//   pass 2: pair acceptance, finish, the compact coefficients; elements of
//           the direction D = -H g and the trial point (or the best point
//           when finished); partial sums of g.D;
//   pass 3: the descent check (restart from steepest descent when it fails)
//           and the line-search state for the next step.
#pragma once
#include "kernels.hh"

namespace mia_hip {

constexpr int kMaxMemory = 16;
constexpr int kSegT = 256;           // threads of a segment block
constexpr int kSegK = 2;             // elements per thread
constexpr int kSeg = kSegT * kSegK;  // elements per segment
constexpr int kSegW = kSegT / 64;    // waves per segment block
constexpr int kRing = kMaxMemory + 1; // ring slots: memory pairs and a spare for the candidate
constexpr int kOptBlock = 512;       // threads of the one-block step of small instances

struct OptState {
    double f, gd, alpha, best_f;
    double gamma, next_alpha; // newest pair's s.y / y.y; the next step length (pass 1 -> pass 2)
    int phase;                // 0: first evaluation pending, 1: line search, 2: done
    int evals;
    int k, head;              // pairs stored, ring slot of the oldest
    int mode;                 // pass 2: 0 nothing, 1 new direction, 2 finished (best point back)
    int first;                // the direction is the first (steepest descent) one
};

// Buffers (doubles) per instance i with n = dim * ncoef parameters at offset
// par_off: X, G, D, BX laid out like the parameters; S and Y hold `memory + 1`
// vectors (ring slots) per instance at offset (memory + 1) * par_off; SY and
// YY (memory + 1)^2 pair products per instance by ring slot; ST the state
// between steps, ST2 the state decided by pass 2; PART and PART2 kPart
// partial sums per segment (segment prefix) of passes 1 and 2.
struct OptBuffers {
    double *X, *G, *D, *BX, *S, *Y, *best_cost;
    int *done; // pinned host memory, read by the host between chunks
    OptState *st, *ST2;
    double *SY, *YY, *PART, *PART2;
    const int *seg_prefix;
};

constexpr int kPart = 6 + 4 * kMaxMemory;

// partial-sum indices: the new pair's s.y, y.y, s.s, s.g, y.g, max |g|, then per
// stored pair c (oldest first) s_c.y, y_c.y, s_c.g, y_c.g; pass 2 uses kPSY for g.D
enum { kPSY = 0, kPYY, kPSS, kPSG, kPYG, kPGMAX, kPPair };

// A double moved across lanes by DPP (both halves; lanes that the row mask
// leaves out, or whose source is out of range, receive `fill`)
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp64(double v, double fill)
{
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Wave reduction in a fixed order, result in lane 63: within each row of 16
// lanes (quad swaps, half-row and row mirrors), then row 0 into row 1 and row
// 2 into row 3 (row_bcast:15), then lane 31 into rows 2 and 3 (row_bcast:31).
template <bool MAX>
__device__ __forceinline__ double wave_reduce63(double v)
{
    const double id = MAX ? -INFINITY : 0.0;
    auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
    v = op(v, dpp64<0xB1, 0xf>(v, id));
    v = op(v, dpp64<0x4E, 0xf>(v, id));
    v = op(v, dpp64<0x141, 0xf>(v, id));
    v = op(v, dpp64<0x140, 0xf>(v, id));
    v = op(v, dpp64<0x142, 0xa>(v, id));
    v = op(v, dpp64<0x143, 0xc>(v, id));
    return v;
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Wave sums of 16 values at once (fixed order): each lane stores its 16 values into its wave's tb
// (16 rows of 65 doubles), then lane 4v + q adds row v's 16 entries of lanes 16q .. 16q + 15 (starting
// at a lane-dependent rotation, to spread the banks) and the quad combines its four partial sums.
// Returns the wave sum of value lane / 4 (in all four lanes of the quad).  Contains two block barriers,
// so every thread of the block calls it the same number of times.
constexpr int kTbRow = 65;

__device__ __forceinline__ double transpose_sum16(const double *v16, double *tb)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int v = 0; v < 16; ++v)
        tb[v * kTbRow + lane] = v16[v];
    __syncthreads();
    const int v = lane >> 2, q = lane & 3;
    const double *row = tb + v * kTbRow + q * 16;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 16; ++j)
        acc += row[(j + 2 * q) & 15];
    acc += dpp64<0xB1, 0xf>(acc, 0.0);
    acc += dpp64<0x4E, 0xf>(acc, 0.0);
    __syncthreads(); // tb free for the next call
    return acc;
}

// this thread's elements of a segment (element base + thread + k * kSegT)
__device__ __forceinline__ void seg_load(double *v, const double *__restrict__ src, long long base, long long n)
{
#pragma unroll
    for (int k = 0; k < kSegK; ++k) {
        const long long e = base + threadIdx.x + (long long)k * kSegT;
        v[k] = e < n ? src[e] : 0.0;
    }
}

__device__ __forceinline__ void seg_store(double *dst, const double *v, long long base, long long n)
{
#pragma unroll
    for (int k = 0; k < kSegK; ++k) {
        const long long e = base + threadIdx.x + (long long)k * kSegT;
        if (e < n)
            dst[e] = v[k];
    }
}

// Pairs are read kLbP at a time, all loads of a group issued before any use.
constexpr int kLbP = 8;

// stored pairs j0 .. j0 + kLbP - 1 (those below kk, ring slots from head) into sc, yc
__device__ __forceinline__ void seg_load_group(double (*sc)[kSegK], double (*yc)[kSegK], int j0, int kk, int head,
                                               int R, const double *S, const double *Y, long long base, long long n)
{
#pragma unroll
    for (int t = 0; t < kLbP; ++t)
        if (j0 + t < kk) {
            const int slot = (head + j0 + t) % R;
            seg_load(sc[t], S + (long long)slot * n, base, n);
            seg_load(yc[t], Y + (long long)slot * n, base, n);
        }
}

// The segment's partial sums: the block's wave totals of each value (in wp,
// complete after a barrier) added over the waves in order by thread v.
__device__ __forceinline__ void seg_partials_out(const double (*wp)[kPart], int nvals, double *__restrict__ out)
{
    const int v = threadIdx.x;
    if (v < nvals) {
        double t = wp[0][v];
#pragma unroll
        for (int w = 1; w < kSegW; ++w)
            t = v == kPGMAX ? fmax(t, wp[w][v]) : t + wp[w][v];
        out[v] = t;
    }
}

// Totals over an instance's segments in ascending order (thread v: value v);
// every block of the instance computes the same totals.
// The loads go out 16 segments at a time (unconditional, from clamped
// segment indices) before they are added in ascending order: a load per
// added term would wait out the memory latency once per segment.
__device__ __forceinline__ void instance_totals(const double *__restrict__ part, int nseg, int nvals, double *tot)
{
    constexpr int U = 16;
    const int v = threadIdx.x;
    if (v < nvals) {
        const bool mx = v == kPGMAX;
        double t = part[v];
        for (int s0 = 1; s0 < nseg; s0 += U) {
            double x[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                x[u] = part[(long long)min(s0 + u, nseg - 1) * kPart + v];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (s0 + u < nseg)
                    t = mx ? fmax(t, x[u]) : t + x[u];
        }
        tot[v] = t;
    }
    __syncthreads();
}

// Coefficients of the compact direction for k >= 1 pairs from ring slot head (wave 0, lane i for pair i,
// oldest first): cu[i] = p_i, cw[i] = gamma w_i, so that H g = gamma g + sum_i (cu_i s_i - cw_i y_i):
//   w = R^-1 a,  p = R^-T((Dg + gamma Y^T Y) w - gamma b),
// with a = S^T g, b = Y^T g, R the upper triangle of S^T Y (R_ij = s_i.y_j, i older or equal), Dg its
// diagonal and gamma = s.y / y.y of the newest pair (the two-loop's scaling).  Pair i's products with g
// come from the totals at stored index i + drop, or from the new pair's s.g, y.g when `fresh` and
// i == k - 1.  mSY, mYY: R x R tables by ring slot.
__device__ __forceinline__ void lbfgs_compact_coeffs(int R, int k, int head, int drop, bool fresh,
                                                     const double *tot, const double *mSY, const double *mYY,
                                                     double gamma, double *cu, double *cw)
{
    const int i = threadIdx.x;
    const int il = i < k ? i : k - 1;
    const int si = (head + il) % R;
    const bool isnew = fresh && il == k - 1;
    const int po = kPPair + 4 * (il + drop);
    const double a = tot[isnew ? kPSG : po + 2];
    const double b = tot[isnew ? kPYG : po + 3];
    const double rinv = 1.0 / mSY[si * R + si];
    // w = R^-1 a (back substitution, column by column)
    double r = a, wi = 0.0;
#pragma unroll
    for (int j = kMaxMemory - 1; j >= 0; --j)
        if (j < k) {
            const int sj = (head + j) % R;
            const double wj = readlane_d(r * rinv, j);
            if (il == j)
                wi = wj;
            if (il < j)
                r -= mSY[si * R + sj] * wj;
        }
    // t = (Dg + gamma Y^T Y) w - gamma b
    double yw = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxMemory; ++j)
        if (j < k) {
            const int sj = (head + j) % R;
            yw += mYY[si * R + sj] * readlane_d(wi, j);
        }
    // p = R^-T t (forward substitution)
    r = mSY[si * R + si] * wi + gamma * yw - gamma * b;
    double pi = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxMemory; ++j)
        if (j < k) {
            const int sj = (head + j) % R;
            const double pj = readlane_d(r * rinv, j);
            if (il == j)
                pi = pj;
            if (il > j)
                r -= mSY[sj * R + si] * pj;
        }
    if (i < k) {
        cu[i] = pi;
        cw[i] = gamma * wi;
    }
}

// Block of segment b: its instance, segment index and count.
struct SegPos {
    int i, sg, nseg;
    long long n, base;
};

__device__ __forceinline__ SegPos seg_pos(const InstDesc *__restrict__ inst, const int *__restrict__ prefix,
                                          int n_inst, int dim)
{
    SegPos p;
    const int b = blockIdx.x;
    p.i = find_inst(prefix, n_inst, b);
    p.sg = b - prefix[p.i];
    p.nseg = prefix[p.i + 1] - prefix[p.i];
    const InstDesc d = inst[p.i];
    p.n = (long long)dim * d.cx * d.cy * d.cz;
    p.base = (long long)p.sg * kSeg;
    return p;
}

// What the evaluated point means for the line search, the same in every
// block: 0 first evaluation, 1 accepted (Armijo), 2 backtrack to step *an,
// 3 finished (evaluations spent, or the step became negligible).
__device__ __forceinline__ int step_kind(const OptState &st, double ft, int maxeval, double *an, bool *better)
{
    *an = 0.0;
    *better = false;
    if (st.phase == 0)
        return 0;
    const bool finite = isfinite(ft);
    *better = finite && ft < st.best_f;
    if (finite && ft <= st.f + 1e-4 * st.alpha * st.gd)
        return 1;
    if (st.evals + 1 >= maxeval)
        return 3;
    const double a = st.alpha;
    double b = 0.5 * a;
    if (finite) {
        const double den = 2.0 * (ft - st.f - a * st.gd);
        if (den > 0.0)
            b = -st.gd * a * a / den;
    }
    b = fmin(fmax(b, 0.1 * a), 0.5 * a);
    *an = b;
    return b < 1e-20 ? 3 : 2;
}

// number of pass-1 partial sums of a step of this kind with kk stored pairs
__device__ __forceinline__ int pass1_nvals(int kind, int kk)
{
    return kind == 0 ? kPGMAX + 1 : kind == 1 ? kPPair + 4 * kk : 0;
}

// Pass 1 (segments): the evaluated point's element updates and the partial
// sums of every product the step needs.
__global__ void __launch_bounds__(kSegT)
    k_lbfgs_pass1(const InstDesc *__restrict__ inst, int n_inst, int dim, OptBuffers ob, double *__restrict__ xt,
                  const double *__restrict__ gt, const double *__restrict__ costs, int memory, int maxeval)
{
    __shared__ double wp[kSegW][kPart];
    __shared__ double tbuf[kSegW * 16 * kTbRow];
    const SegPos q = seg_pos(inst, ob.seg_prefix, n_inst, dim);
    const int i = q.i;
    const long long n = q.n, base = q.base;
    const OptState st = ob.st[i];
    if (st.phase == 2)
        return;
    const int R = memory + 1;
    const long long off = inst[i].par_off;
    double *X = ob.X + off, *G = ob.G + off, *D = ob.D + off, *BX = ob.BX + off;
    double *S = ob.S + (long long)R * off, *Y = ob.Y + (long long)R * off;
    double *XT = xt + off;
    const double *GT = gt + off;
    double an;
    bool better;
    const int kind = step_kind(st, costs[i], maxeval, &an, &better);
    double rxt[kSegK], rgt[kSegK];
    seg_load(rxt, XT, base, n);
    seg_load(rgt, GT, base, n);
    if (better)
        seg_store(BX, rxt, base, n);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *tb = tbuf + w * 16 * kTbRow;
    const int kk = st.k, head = st.head;
    if (kind == 0) {
        double gm = 0.0;
#pragma unroll
        for (int k = 0; k < kSegK; ++k)
            gm = fmax(gm, fabs(rgt[k]));
        seg_store(X, rxt, base, n);
        seg_store(G, rgt, base, n);
        seg_store(BX, rxt, base, n);
        gm = wave_reduce63<true>(gm);
        if (lane < kPGMAX)
            wp[w][lane] = 0.0; // the pair sums are unused here: zero, not stale
        if (lane == 63)
            wp[w][kPGMAX] = gm;
    } else if (kind == 1) {
        double rx[kSegK], rg[kSegK], sv[kSegK], yv[kSegK];
        seg_load(rx, X, base, n);
        seg_load(rg, G, base, n);
        double sy = 0.0, yy = 0.0, ss = 0.0, sg = 0.0, yg = 0.0, gm = 0.0;
#pragma unroll
        for (int k = 0; k < kSegK; ++k) {
            sv[k] = rxt[k] - rx[k];
            yv[k] = rgt[k] - rg[k];
            sy += sv[k] * yv[k];
            yy += yv[k] * yv[k];
            ss += sv[k] * sv[k];
            sg += sv[k] * rgt[k];
            yg += yv[k] * rgt[k];
            gm = fmax(gm, fabs(rgt[k]));
        }
        // the candidate pair into the spare slot (part of the ring only if
        // pass 2 accepts it)
        const int ns = (head + kk) % R;
        seg_store(S + (long long)ns * n, sv, base, n);
        seg_store(Y + (long long)ns * n, yv, base, n);
        seg_store(X, rxt, base, n);
        seg_store(G, rgt, base, n);
        {
            const double h[16] = {sy, yy, ss, sg, yg, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            const double t = transpose_sum16(h, tb);
            gm = wave_reduce63<true>(gm);
            if ((lane & 3) == 3 && lane < 4 * kPGMAX)
                wp[w][lane >> 2] = t;
            if (lane == 63)
                wp[w][kPGMAX] = gm;
        }
        for (int j0 = 0; j0 < kk; j0 += kLbP) {
            double sc[kLbP][kSegK], yc[kLbP][kSegK];
            seg_load_group(sc, yc, j0, kk, head, R, S, Y, base, n);
            double dd[4 * kLbP];
#pragma unroll
            for (int t = 0; t < kLbP; ++t) {
                double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
#pragma unroll
                for (int k = 0; k < kSegK; ++k) {
                    d0 += sc[t][k] * yv[k];
                    d1 += yc[t][k] * yv[k];
                    d2 += sc[t][k] * rgt[k];
                    d3 += yc[t][k] * rgt[k];
                }
                dd[4 * t] = d0;
                dd[4 * t + 1] = d1;
                dd[4 * t + 2] = d2;
                dd[4 * t + 3] = d3;
            }
#pragma unroll
            for (int h = 0; h < kLbP / 4; ++h) {
                const double t = transpose_sum16(dd + 16 * h, tb);
                const int j = j0 + 4 * h + (lane >> 4);
                if ((lane & 3) == 3 && j < kk)
                    wp[w][kPPair + 4 * (j0 + 4 * h) + (lane >> 2)] = t;
            }
        }
    } else if (kind == 2) {
        double rx[kSegK], dv[kSegK], xv[kSegK];
        seg_load(rx, X, base, n);
        seg_load(dv, D, base, n);
#pragma unroll
        for (int k = 0; k < kSegK; ++k)
            xv[k] = rx[k] + an * dv[k];
        seg_store(XT, xv, base, n);
    }
    __syncthreads();
    seg_partials_out(wp, pass1_nvals(kind, kk), ob.PART + ((long long)ob.seg_prefix[i] + q.sg) * kPart);
}

// Pass 2 (segments): every block adds up pass 1's partial sums and takes the
// step's decisions (pair acceptance, finish, the compact coefficients), the
// same in each; then its elements of the direction and the trial point (or
// the best point when finished) and the partial sums of g.D.  Block 0 of
// the instance stores the decided state in ST2 and the pair-product entries
// of the new pair's slot (the spare slot, which no block reads in this pass).
__global__ void __launch_bounds__(kSegT)
    k_lbfgs_pass2(const InstDesc *__restrict__ inst, int n_inst, int dim, OptBuffers ob, double *__restrict__ xt,
                  const double *__restrict__ costs, int memory, int maxeval, double first_step)
{
    __shared__ double wp[kSegW][kPart];
    __shared__ double tot[kPart];
    __shared__ double mSY[kRing * kRing], mYY[kRing * kRing];
    __shared__ double cu[kMaxMemory], cw[kMaxMemory];
    __shared__ OptState s;
    __shared__ int flags[3]; // fresh (pair stored), drop (oldest pair replaced), new slot
    const SegPos q = seg_pos(inst, ob.seg_prefix, n_inst, dim);
    const int i = q.i;
    const long long n = q.n, base = q.base;
    const OptState st = ob.st[i];
    if (st.phase == 2) {
        if (q.sg == 0 && threadIdx.x == 0) {
            OptState s2 = st;
            s2.mode = 0;
            ob.ST2[i] = s2;
        }
        return;
    }
    const int R = memory + 1;
    const long long off = inst[i].par_off;
    double *X = ob.X + off, *G = ob.G + off, *D = ob.D + off, *BX = ob.BX + off;
    const double *S = ob.S + (long long)R * off, *Y = ob.Y + (long long)R * off;
    double *XT = xt + off;
    const double ft = costs[i];
    const bool finite = isfinite(ft);
    double an;
    bool better;
    const int kind = step_kind(st, ft, maxeval, &an, &better);
    const int kk = st.k, head = st.head;
    instance_totals(ob.PART + (long long)ob.seg_prefix[i] * kPart, q.nseg, pass1_nvals(kind, kk), tot);
    for (int t = threadIdx.x; t < R * R; t += kSegT) {
        mSY[t] = ob.SY[(long long)R * R * i + t];
        mYY[t] = ob.YY[(long long)R * R * i + t];
    }
    if (threadIdx.x == 0) {
        s = st;
        flags[0] = flags[1] = 0;
        flags[2] = -1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bool finish = false, search = false;
        if (kind == 0) {
            s.f = ft;
            s.evals = 1;
            s.best_f = finite ? ft : INFINITY;
            s.k = 0;
            s.head = 0;
            const double gmax = tot[kPGMAX];
            if (!finite || gmax <= 0.0 || s.evals >= maxeval) {
                finish = true;
            } else {
                search = true;
                s.first = 1;
                s.next_alpha = fmin(1.0, first_step / gmax);
            }
        } else {
            const double f0 = s.f;
            s.evals += 1;
            if (better)
                s.best_f = ft;
            if (kind == 1) {
                const double sy = tot[kPSY], yy = tot[kPYY], ss = tot[kPSS], gmax = tot[kPGMAX];
                if (sy > 1e-12 * sqrt(ss * yy)) {
                    const int drop = kk < memory ? 0 : 1;
                    const int ns = (head + kk) % R;
                    // the new pair's products with the pairs that stay
                    for (int jo = drop; jo < kk; ++jo) {
                        const int sc = (head + jo) % R;
                        mSY[sc * R + ns] = tot[kPPair + 4 * jo];
                        mYY[sc * R + ns] = mYY[ns * R + sc] = tot[kPPair + 4 * jo + 1];
                    }
                    mSY[ns * R + ns] = sy;
                    mYY[ns * R + ns] = yy;
                    if (kk < memory)
                        s.k = kk + 1;
                    else
                        s.head = (head + 1) % R;
                    flags[0] = 1;
                    flags[1] = drop;
                    flags[2] = ns;
                }
                s.f = ft;
                if (s.evals >= maxeval || gmax <= 0.0 || f0 == ft) {
                    finish = true;
                } else {
                    search = true;
                    s.first = 0;
                    s.next_alpha = 1.0;
                }
            } else if (kind == 3) {
                finish = true;
            } else {
                s.alpha = an;
            }
        }
        s.mode = finish ? 2 : search ? 1 : 0;
        s.gamma = 1.0;
        if (search && s.k > 0) {
            const int sl = (s.head + s.k - 1) % R;
            s.gamma = mSY[sl * R + sl] / mYY[sl * R + sl];
        }
    }
    __syncthreads();
    if (s.mode == 1 && s.k > 0 && threadIdx.x < 64)
        lbfgs_compact_coeffs(R, s.k, s.head, flags[1], flags[0] != 0, tot, mSY, mYY, s.gamma, cu, cw);
    if (q.sg == 0) {
        const int ns = flags[2];
        if (ns >= 0)
            for (int t = threadIdx.x; t < R; t += kSegT) {
                ob.SY[(long long)R * R * i + t * R + ns] = mSY[t * R + ns];
                ob.YY[(long long)R * R * i + t * R + ns] = mYY[t * R + ns];
                ob.YY[(long long)R * R * i + ns * R + t] = mYY[ns * R + t];
            }
        if (threadIdx.x == 0)
            ob.ST2[i] = s;
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int nvals = 0;
    if (s.mode == 2) {
        double bx[kSegK];
        seg_load(bx, BX, base, n);
        seg_store(XT, bx, base, n);
    } else if (s.mode == 1) {
        // D = -(gamma g + sum_j (cu_j s_j - cw_j y_j)), the trial point X + alpha D
        const int k = s.k, hd = s.head;
        double g[kSegK], x[kSegK], dv[kSegK];
        seg_load(g, G, base, n);
        seg_load(x, X, base, n);
#pragma unroll
        for (int k2 = 0; k2 < kSegK; ++k2)
            dv[k2] = s.gamma * g[k2];
        for (int j0 = 0; j0 < k; j0 += kLbP) {
            double sc[kLbP][kSegK], yc[kLbP][kSegK];
            seg_load_group(sc, yc, j0, k, hd, R, S, Y, base, n);
#pragma unroll
            for (int t = 0; t < kLbP; ++t)
                if (j0 + t < k) {
                    const double u = cu[j0 + t], v = cw[j0 + t];
#pragma unroll
                    for (int k2 = 0; k2 < kSegK; ++k2) {
                        dv[k2] += u * sc[t][k2];
                        dv[k2] -= v * yc[t][k2];
                    }
                }
        }
        double gd = 0.0, xv[kSegK];
#pragma unroll
        for (int k2 = 0; k2 < kSegK; ++k2) {
            dv[k2] = -dv[k2];
            const long long e = base + threadIdx.x + (long long)k2 * kSegT;
            if (e < n)
                gd += g[k2] * dv[k2];
            xv[k2] = x[k2] + s.next_alpha * dv[k2];
        }
        seg_store(D, dv, base, n);
        seg_store(XT, xv, base, n);
        gd = wave_reduce63<false>(gd);
        if (lane == 63)
            wp[w][0] = gd;
        nvals = 1;
    }
    __syncthreads();
    seg_partials_out(wp, nvals, ob.PART2 + ((long long)ob.seg_prefix[i] + q.sg) * kPart);
}

// Pass 3 (segments): every block adds up g.D; a direction that does not
// descend is replaced by steepest descent (D = -G, the first step's length
// from max |g| of pass 1, g.D = -g.g), each block rewriting its elements.
// Block 0 stores the line-search state for the next step.
__global__ void __launch_bounds__(kSegT)
    k_lbfgs_pass3(const InstDesc *__restrict__ inst, int n_inst, int dim, OptBuffers ob, double *__restrict__ xt,
                  double first_step)
{
    __shared__ double tot[kPart];
    __shared__ double red[kSegW];
    const SegPos q = seg_pos(inst, ob.seg_prefix, n_inst, dim);
    const int i = q.i;
    const long long n = q.n, base = q.base;
    OptState s = ob.ST2[i];
    if (s.phase == 2)
        return;
    const long long off = inst[i].par_off;
    double *X = ob.X + off, *G = ob.G + off, *D = ob.D + off;
    double *XT = xt + off;
    if (s.mode == 1) {
        instance_totals(ob.PART2 + (long long)ob.seg_prefix[i] * kPart, q.nseg, 1, tot);
        if (tot[0] < 0.0) {
            s.gd = tot[0];
            s.alpha = s.next_alpha;
        } else {
            // restart: g.g and max |g| over the instance (the instance's G,
            // one pass in every block), then this segment's elements
            double gg = 0.0, gm = 0.0;
            for (long long e = threadIdx.x; e < n; e += kSegT) {
                gg += G[e] * G[e];
                gm = fmax(gm, fabs(G[e]));
            }
            const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
            gg = wave_reduce63<false>(gg);
            gm = wave_reduce63<true>(gm);
            if (lane == 63) {
                red[w] = gg;
                tot[1 + w] = gm;
            }
            __syncthreads();
            double tg = red[0], tm = tot[1];
#pragma unroll
            for (int v = 1; v < kSegW; ++v) {
                tg += red[v];
                tm = fmax(tm, tot[1 + v]);
            }
            const double alpha = tm > 0.0 ? fmin(1.0, first_step / tm) : 1.0;
            double g[kSegK], x[kSegK], dv[kSegK], xv[kSegK];
            seg_load(g, G, base, n);
            seg_load(x, X, base, n);
#pragma unroll
            for (int k2 = 0; k2 < kSegK; ++k2) {
                dv[k2] = -g[k2];
                xv[k2] = x[k2] + alpha * dv[k2];
            }
            seg_store(D, dv, base, n);
            seg_store(XT, xv, base, n);
            s.k = 0;
            s.head = 0;
            s.first = 1;
            s.gd = -tg;
            s.alpha = alpha;
        }
        s.phase = 1;
    } else if (s.mode == 2) {
        s.phase = 2;
    }
    if (q.sg == 0 && threadIdx.x == 0) {
        s.mode = 0;
        ob.st[i] = s;
        ob.best_cost[i] = s.best_f;
        if (s.phase == 2)
            ob.done[i] = 1;
    }
}

// ---------------------------------------------------------------------------
// small instances: the whole step in one block
// ---------------------------------------------------------------------------

// Block sum / max with the result in every thread (fixed order).  red holds
// two sets of per-wave results used alternately (par flips per call, the same
// in every thread), so one barrier per reduction suffices: a set is written
// again only after every thread has passed the next reduction's barrier.
template <bool MAX>
__device__ __forceinline__ double block_reduce_all(double v, double *red, int &par)
{
    v = wave_reduce63<MAX>(v);
    double *r = red + par * (kOptBlock / 64);
    par ^= 1;
    if ((threadIdx.x & 63) == 63)
        r[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = r[0];
#pragma unroll
    for (int k = 1; k < kOptBlock / 64; ++k)
        s = MAX ? fmax(s, r[k]) : s + r[k];
    return s;
}

__device__ __forceinline__ double block_sum_all(double v, double *red, int &par)
{
    return block_reduce_all<false>(v, red, par);
}

__device__ __forceinline__ double block_max_all(double v, double *red, int &par)
{
    return block_reduce_all<true>(v, red, par);
}

// Instances of up to kLbRegN parameters run the whole step in one block of
// kOptBlock threads with the vectors in registers (element e = thread +
// k * kOptBlock, k < kLbK): for them the step is a chain of short dependent
// phases, faster within one block than split over launches.
constexpr int kLbK = 4;
constexpr long long kLbRegN = (long long)kLbK * kOptBlock;

// This thread's k-th element of a vector: e = thread + k kOptBlock, or with
// P (every vector of the instance starts 16-byte aligned and has an even
// length) adjacent pairs, e = 2 thread + (k & 1) + (k >> 1) 2 kOptBlock,
// loaded and stored two at a time (half the memory instructions, so that a
// step's loads fit the wave's count of loads in flight)
template <bool P>
__device__ __forceinline__ long long lb_e(int k)
{
    return P ? 2LL * threadIdx.x + (k & 1) + (long long)(k >> 1) * 2 * kOptBlock
             : (long long)threadIdx.x + (long long)k * kOptBlock;
}

template <bool P>
__device__ __forceinline__ void lb_load(double *v, const double *__restrict__ src, long long n)
{
    if constexpr (P) {
#pragma unroll
        for (int k = 0; k < kLbK; k += 2) {
            const long long e = lb_e<P>(k);
            // an unconditional load from a clamped pair (n is even)
            const double2 q = *(const double2 *)(src + min(e, n - 2));
            v[k] = e < n ? q.x : 0.0;
            v[k + 1] = e < n ? q.y : 0.0;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kLbK; ++k) {
            const long long e = lb_e<P>(k);
            v[k] = e < n ? src[e] : 0.0;
        }
    }
}

// Compact form of the same direction for instances of up to kLbRegN
// parameters (Byrd, Nocedal & Schnabel 1994, eq. 2.6-2.7, with H0 = gamma I):
//   H g = gamma g + [S  gamma Y] [p; q],  w = R^-1 a,  p = R^-T((Dg + gamma Y^T Y) w - gamma b),  q = -w,
// with a = S^T g, b = Y^T g, R the upper triangle of S^T Y (R_ij = s_i.y_j, i older or equal), Dg its
// diagonal and gamma = s.y / y.y of the newest pair (the two-loop's scaling).  All dot products of an
// evaluation come from one pass (lbfgs_accept_pass) and one batched reduction, and the small triangular
// systems are solved by the lanes of wave 0, instead of 2k + 1 dependent block reductions.  The pair
// products s_i.y_j and y_i.y_j are kept per instance in memory x memory tables indexed by ring slot.
__device__ __forceinline__ double part_total(const double (*part)[kPart], int v)
{
    double t = part[0][v];
#pragma unroll
    for (int w = 1; w < kOptBlock / 64; ++w)
        t += part[w][v];
    return t;
}

__device__ __forceinline__ double part_max(const double (*part)[kPart], int v)
{
    double t = part[0][v];
#pragma unroll
    for (int w = 1; w < kOptBlock / 64; ++w)
        t = fmax(t, part[w][v]);
    return t;
}


template <bool P>
__device__ __forceinline__ void lb_store(double *dst, const double *v, long long n)
{
    if constexpr (P) {
#pragma unroll
        for (int k = 0; k < kLbK; k += 2) {
            const long long e = lb_e<P>(k);
            if (e < n)
                *(double2 *)(dst + e) = make_double2(v[k], v[k + 1]);
        }
    } else {
#pragma unroll
        for (int k = 0; k < kLbK; ++k) {
            const long long e = lb_e<P>(k);
            if (e < n)
                dst[e] = v[k];
        }
    }
}

// Pairs j0 .. j0 + kLbP - 1 (those below kk) from ring slot head into sc, yc
template <bool P>
__device__ __forceinline__ void lb_load_group(double (*sc)[kLbK], double (*yc)[kLbK], int j0, int kk, int head,
                                              int memory, const double *S, const double *Y, long long n)
{
#pragma unroll
    for (int t = 0; t < kLbP; ++t)
        if (j0 + t < kk) {
            const int slot = (head + j0 + t) % memory;
            lb_load<P>(sc[t], S + (long long)slot * n, n);
            lb_load<P>(yc[t], Y + (long long)slot * n, n);
        }
}

// With memory <= kLbRegMem the one-block step keeps every stored pair in
// registers for the whole step: read once, oldest first, they serve both the
// product pass and the direction, which otherwise reads them from memory a
// second time.
constexpr int kLbRegMem = 10;

// Stored pairs 0 .. kk - 1 (oldest first, ring slots from head) into ps, py.
// Every load is issued unconditionally (slots beyond kk repeat the newest
// stored pair, or slot head when none is stored: in bounds, cached, unused),
// so that none waits at the end of a branch.
template <bool P>
__device__ __forceinline__ void lb_load_pairs(double (*ps)[kLbK], double (*py)[kLbK], int kk, int head, int memory,
                                              const double *S, const double *Y, long long n)
{
#pragma unroll
    for (int j = 0; j < kLbRegMem; ++j) {
        const int jc = j < kk ? j : (kk > 0 ? kk - 1 : 0);
        const int slot = (head + jc) % memory;
        lb_load<P>(ps[j], S + (long long)slot * n, n);
        lb_load<P>(py[j], Y + (long long)slot * n, n);
    }
}

// Accepted step, vectors in registers (xt, x, gv, g: this thread's elements of XT, X, GT, G):
// X <- XT, G <- GT, the new pair s = XT - X, y = GT - G left in sv, yv, and the wave partials of every product of kPart written to part[wave][.] (complete after the
// caller's barrier).  kk pairs are stored from ring slot head.
template <bool P>
__device__ __forceinline__ void lbfgs_accept_pass(long long n, int memory, int kk, int head, const double *S,
                                                  const double *Y, double *X, double *G, const double *xt,
                                                  const double *x, const double *gv, const double *g, double *sv,
                                                  double *yv, double (*part)[kPart], double *tbuf)
{
    double sy = 0.0, yy = 0.0, ss = 0.0, sg = 0.0, yg = 0.0, gm = 0.0;
#pragma unroll
    for (int k = 0; k < kLbK; ++k) {
        sv[k] = xt[k] - x[k];
        yv[k] = gv[k] - g[k];
        sy += sv[k] * yv[k];
        yy += yv[k] * yv[k];
        ss += sv[k] * sv[k];
        sg += sv[k] * gv[k];
        yg += yv[k] * gv[k];
        gm = fmax(gm, fabs(gv[k]));
    }
    lb_store<P>(X, xt, n);
    lb_store<P>(G, gv, n);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *tb = tbuf + w * 16 * kTbRow;
    {
        const double h[16] = {sy, yy, ss, sg, yg, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        const double t = transpose_sum16(h, tb);
        gm = wave_reduce63<true>(gm);
        if ((lane & 3) == 3 && lane < 4 * kPGMAX)
            part[w][lane >> 2] = t;
        if (lane == 63)
            part[w][kPGMAX] = gm;
    }
    for (int j0 = 0; j0 < kk; j0 += kLbP) {
        double sc[kLbP][kLbK], yc[kLbP][kLbK];
        lb_load_group<P>(sc, yc, j0, kk, head, memory, S, Y, n);
        double dd[4 * kLbP];
#pragma unroll
        for (int t = 0; t < kLbP; ++t) {
            double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
#pragma unroll
            for (int k = 0; k < kLbK; ++k) {
                d0 += sc[t][k] * yv[k];
                d1 += yc[t][k] * yv[k];
                d2 += sc[t][k] * gv[k];
                d3 += yc[t][k] * gv[k];
            }
            dd[4 * t] = d0;
            dd[4 * t + 1] = d1;
            dd[4 * t + 2] = d2;
            dd[4 * t + 3] = d3;
        }
#pragma unroll
        for (int h = 0; h < kLbP / 4; ++h) {
            const double t = transpose_sum16(dd + 16 * h, tb);
            const int j = j0 + 4 * h + (lane >> 4);
            if ((lane & 3) == 3 && j < kk)
                part[w][kPPair + 4 * (j0 + 4 * h) + (lane >> 2)] = t;
        }
    }
}

// Block sums of many values per thread for the register-resident step, in a
// fixed order: each quad of lanes adds its four values (DPP), the quad's
// first lane stores the sum into st[value][quad] (kQs = 128 quads per
// block); after a barrier, collect_quad_sums adds each value's quad sums in
// eight chunks of 16 into part[chunk][value], which part_total / part_max
// then add in chunk order.  One barrier for any number of values, where a
// transposed sum of 16 values takes two.
constexpr int kQs = kOptBlock / 4;
constexpr int kQsVals = kPPair + 4 * kLbRegMem; // values of a step with kLbRegMem pairs

template <bool MAX>
__device__ __forceinline__ void quad_stage(double v, double *st, int val)
{
    const double id = MAX ? -INFINITY : 0.0;
    auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
    v = op(v, dpp64<0xB1, 0xf>(v, id));
    v = op(v, dpp64<0x4E, 0xf>(v, id));
    if ((threadIdx.x & 3) == 0)
        st[val * kQs + (threadIdx.x >> 2)] = v;
}

__device__ __forceinline__ void collect_quad_sums(const double *st, int nvals, double (*part)[kPart])
{
    static_assert(kOptBlock / 64 == 8 && kQs == 8 * 16, "eight chunks of 16 quad sums");
    for (int t = threadIdx.x; t < nvals * 8; t += kOptBlock) {
        const int v = t >> 3, c = t & 7;
        const double *q = st + v * kQs + c * 16;
        double x[16];
#pragma unroll
        for (int j = 0; j < 16; ++j)
            x[j] = q[j];
        double s = x[0];
#pragma unroll
        for (int j = 1; j < 16; ++j)
            s = v == kPGMAX ? fmax(s, x[j]) : s + x[j];
        part[c][v] = s;
    }
}

// lbfgs_accept_pass with the kk stored pairs already in registers (ps, py,
// oldest first): the same products, each pair's four computed and staged as
// soon as its loads arrive, summed over the block by quads (quad_stage,
// collect_quad_sums); part[chunk][.] is complete after the caller's barrier.
template <bool P>
__device__ __forceinline__ void lbfgs_accept_pass_reg(long long n, int kk, double *X, double *G, const double *xt,
                                                      const double *x, const double *gv, const double *g,
                                                      double *sv, double *yv, const double (*ps)[kLbK],
                                                      const double (*py)[kLbK], double (*part)[kPart], double *st)
{
    double sy = 0.0, yy = 0.0, ss = 0.0, sg = 0.0, yg = 0.0, gm = 0.0;
#pragma unroll
    for (int k = 0; k < kLbK; ++k) {
        sv[k] = xt[k] - x[k];
        yv[k] = gv[k] - g[k];
        sy += sv[k] * yv[k];
        yy += yv[k] * yv[k];
        ss += sv[k] * sv[k];
        sg += sv[k] * gv[k];
        yg += yv[k] * gv[k];
        gm = fmax(gm, fabs(gv[k]));
    }
    lb_store<P>(X, xt, n);
    lb_store<P>(G, gv, n);
    quad_stage<false>(sy, st, kPSY);
    quad_stage<false>(yy, st, kPYY);
    quad_stage<false>(ss, st, kPSS);
    quad_stage<false>(sg, st, kPSG);
    quad_stage<false>(yg, st, kPYG);
    quad_stage<true>(gm, st, kPGMAX);
#pragma unroll
    for (int j = 0; j < kLbRegMem; ++j)
        if (j < kk) {
            double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
#pragma unroll
            for (int k = 0; k < kLbK; ++k) {
                d0 += ps[j][k] * yv[k];
                d1 += py[j][k] * yv[k];
                d2 += ps[j][k] * gv[k];
                d3 += py[j][k] * gv[k];
            }
            quad_stage<false>(d0, st, kPPair + 4 * j);
            quad_stage<false>(d1, st, kPPair + 4 * j + 1);
            quad_stage<false>(d2, st, kPPair + 4 * j + 2);
            quad_stage<false>(d3, st, kPPair + 4 * j + 3);
        }
    __syncthreads();
    collect_quad_sums(st, kPPair + 4 * kk, part);
}

// Coefficients of the compact direction for k >= 1 pairs from ring slot head (wave 0, lane i for pair i,
// oldest first): cu[i] = p_i, cw[i] = gamma w_i, so that H g = gamma g + sum_i (cu_i s_i - cw_i y_i).
// Pair i's products with g come from part[] at stored index i + drop, or from the new pair's s.g, y.g
// when `fresh` and i == k - 1.
__device__ __forceinline__ void block_compact_coeffs(int memory, int k, int head, int drop, bool fresh,
                                                     const double (*part)[kPart], const double *mSY,
                                                     const double *mYY, double gamma, double *cu, double *cw)
{
    const int i = threadIdx.x;
    const int il = i < k ? i : k - 1;
    const int si = head + il < memory ? head + il : head + il - memory;
    const bool isnew = fresh && il == k - 1;
    const int po = kPPair + 4 * (il + drop);
    const double a = part_total(part, isnew ? kPSG : po + 2);
    const double b = part_total(part, isnew ? kPYG : po + 3);
    const double rinv = 1.0 / mSY[si * memory + si];
    // ring slot of stored pair j (j clamped to the stored ones): the table
    // value each step needs is read one step ahead, so that its LDS latency
    // overlaps the previous step instead of stalling the dependent chain
    auto slot = [&](int j) {
        const int jc = min(max(j, 0), k - 1);
        return head + jc < memory ? head + jc : head + jc - memory;
    };
    // w = R^-1 a (back substitution, column by column: j = k - 1 .. 0)
    double r = a, wi = 0.0;
    double mcur = mSY[si * memory + slot(k - 1)];
#pragma unroll
    for (int t = 0; t < kMaxMemory; ++t)
        if (t < k) {
            const int j = k - 1 - t;
            const double mnext = mSY[si * memory + slot(j - 1)];
            const double wj = readlane_d(r * rinv, j);
            if (il == j)
                wi = wj;
            if (il < j)
                r -= mcur * wj;
            mcur = mnext;
        }
    // t = (Dg + gamma Y^T Y) w - gamma b
    double yw = 0.0;
    mcur = mYY[si * memory + slot(0)];
#pragma unroll
    for (int j = 0; j < kMaxMemory; ++j)
        if (j < k) {
            const double mnext = mYY[si * memory + slot(j + 1)];
            yw += mcur * readlane_d(wi, j);
            mcur = mnext;
        }
    // p = R^-T t (forward substitution)
    r = mSY[si * memory + si] * wi + gamma * yw - gamma * b;
    double pi = 0.0;
    mcur = mSY[slot(0) * memory + si];
#pragma unroll
    for (int j = 0; j < kMaxMemory; ++j)
        if (j < k) {
            const double mnext = mSY[slot(j + 1) * memory + si];
            const double pj = readlane_d(r * rinv, j);
            if (il == j)
                pi = pj;
            if (il > j)
                r -= mcur * pj;
            mcur = mnext;
        }
    if (i < k) {
        cu[i] = pi;
        cw[i] = gamma * wi;
    }
}

// D <- -H g from the coefficients (all threads; g: this thread's gradient elements, sc / yc: the first
// group of pairs, already loaded), also left in dv; returns g.D
template <bool P>
__device__ __forceinline__ double lbfgs_compact_direction(long long n, int memory, int k, int head,
                                                          const double *S, const double *Y, const double *g,
                                                          double (*sc)[kLbK], double (*yc)[kLbK], double *D,
                                                          double *dv, double gamma, const double *cu,
                                                          const double *cw, double *red, int &rpar)
{
#pragma unroll
    for (int k2 = 0; k2 < kLbK; ++k2)
        dv[k2] = gamma * g[k2];
    for (int j0 = 0; j0 < k; j0 += kLbP) {
        if (j0 > 0)
            lb_load_group<P>(sc, yc, j0, k, head, memory, S, Y, n);
#pragma unroll
        for (int t = 0; t < kLbP; ++t)
            if (j0 + t < k) {
                const double u = cu[j0 + t], v = cw[j0 + t];
#pragma unroll
                for (int k2 = 0; k2 < kLbK; ++k2) {
                    dv[k2] += u * sc[t][k2];
                    dv[k2] -= v * yc[t][k2];
                }
            }
    }
    double gd = 0.0;
#pragma unroll
    for (int k2 = 0; k2 < kLbK; ++k2) {
        dv[k2] = -dv[k2];
        const long long e = lb_e<P>(k2);
        if (e < n)
            gd += g[k2] * dv[k2];
    }
    lb_store<P>(D, dv, n);
    return block_sum_all(gd, red, rpar);
}


// lbfgs_compact_direction with the pairs in registers: the stored pairs
// drop .. kk - 1 (ps, py) and, when a pair was stored by this step, the new
// one (sv, yv), oldest first, added in the same order.
template <bool P>
__device__ __forceinline__ double lbfgs_compact_direction_reg(long long n, int kk, int drop, bool fresh, int knew,
                                                              const double *g, const double (*ps)[kLbK],
                                                              const double (*py)[kLbK], const double *sv,
                                                              const double *yv, double *D, double *dv, double gamma,
                                                              const double *cu, const double *cw, double *red,
                                                              int &rpar)
{
#pragma unroll
    for (int k2 = 0; k2 < kLbK; ++k2)
        dv[k2] = gamma * g[k2];
#pragma unroll
    for (int j = 0; j < kLbRegMem; ++j)
        if (j >= drop && j < kk) {
            const double u = cu[j - drop], v = cw[j - drop];
#pragma unroll
            for (int k2 = 0; k2 < kLbK; ++k2) {
                dv[k2] += u * ps[j][k2];
                dv[k2] -= v * py[j][k2];
            }
        }
    if (fresh) {
        const double u = cu[knew - 1], v = cw[knew - 1];
#pragma unroll
        for (int k2 = 0; k2 < kLbK; ++k2) {
            dv[k2] += u * sv[k2];
            dv[k2] -= v * yv[k2];
        }
    }
    double gd = 0.0;
#pragma unroll
    for (int k2 = 0; k2 < kLbK; ++k2) {
        dv[k2] = -dv[k2];
        const long long e = lb_e<P>(k2);
        if (e < n)
            gd += g[k2] * dv[k2];
    }
    lb_store<P>(D, dv, n);
    return block_sum_all(gd, red, rpar);
}

// One step for an instance of up to kLbRegN parameters, one block each (the
// other instances return at once: they take k_lbfgs_pass1..3).  Same step as
// the segmented passes, with the ring of `memory` slots, products summed
// over the block's waves.  REG (memory <= kLbRegMem): the stored pairs are read
// once into registers and serve both passes.
template <bool REG, bool P>
__global__ void __launch_bounds__(kOptBlock)
    k_lbfgs_step_small(const InstDesc *__restrict__ inst, int dim, OptBuffers ob, double *__restrict__ xt,
                       const double *__restrict__ gt, const double *__restrict__ costs, int memory, int maxeval,
                       double first_step)
{
    __shared__ double red[2 * (kOptBlock / 64)];
    __shared__ double part[kOptBlock / 64][kPart];
    __shared__ double mSY[kMaxMemory * kMaxMemory], mYY[kMaxMemory * kMaxMemory];
    __shared__ double cu[kMaxMemory], cw[kMaxMemory];
    __shared__ double tbuf[REG ? 1 : (kOptBlock / 64) * 16 * kTbRow]; // transposed sums (pairs from memory)
    __shared__ double st[REG ? kQsVals * kQs : 1];                      // quad sums (pairs in registers)
    int rpar = 0; // which half of red the next block reduction uses
    __shared__ OptState s;
    const int i = blockIdx.x;
    const InstDesc d = inst[i];
    const long long n = (long long)dim * d.cx * d.cy * d.cz;
    if (n > kLbRegN)
        return;
    const int R = memory + 1; // allocation stride of the ring and the tables
    double *SYi = ob.SY + (long long)R * R * i, *YYi = ob.YY + (long long)R * R * i;
    bool fresh = false; // a pair was stored by this step
    int drop = 0;       // ... and replaced the oldest one
    int kold = 0;       // pairs stored before this step (accepted step)
    const long long off = d.par_off;
    double *X = ob.X + off, *G = ob.G + off, *D = ob.D + off, *BX = ob.BX + off;
    double *S = ob.S + (long long)R * off, *Y = ob.Y + (long long)R * off;
    double *XT = xt + off;
    const double *GT = gt + off;
    // this thread's elements of XT, X, GT, G, loaded while the state is read
    double rxt[kLbK], rx[kLbK], rgt[kLbK], rg[kLbK];
    double sv[kLbK], yv[kLbK];                              // the new pair (accepted step)
    double ps[REG ? kLbRegMem : 1][kLbK], py[REG ? kLbRegMem : 1][kLbK]; // stored pairs (REG)
    lb_load<P>(rxt, XT, n);
    lb_load<P>(rgt, GT, n);
    lb_load<P>(rx, X, n);
    lb_load<P>(rg, G, n);
    const double ft = costs[i];
    if (threadIdx.x == 0)
        s = ob.st[i];
    __syncthreads();
    if (s.phase == 2)
        return;
    const bool finite = isfinite(ft);
    bool search = false; // compute a new direction
    bool first = false;
    bool finish = false;
    bool backtrack = false;
    if (s.phase == 0) {
        lb_store<P>(X, rxt, n);
        lb_store<P>(G, rgt, n);
        lb_store<P>(BX, rxt, n);
        double gmax = 0.0;
#pragma unroll
        for (int k = 0; k < kLbK; ++k)
            gmax = fmax(gmax, fabs(rgt[k]));
        if (threadIdx.x == 0) {
            s.f = ft;
            s.evals = 1;
            s.best_f = finite ? ft : INFINITY;
            s.k = 0;
            s.head = 0;
        }
        gmax = block_max_all(gmax, red, rpar);
        if (!finite || gmax <= 0.0 || 1 >= maxeval)
            finish = true;
        else {
            search = true;
            first = true;
        }
    } else {
        const double f0 = s.f, a0 = s.alpha, gd0 = s.gd;
        const bool better = finite && ft < s.best_f;
        if (better)
            lb_store<P>(BX, rxt, n);
        __syncthreads();
        if (threadIdx.x == 0) {
            s.evals += 1;
            if (better)
                s.best_f = ft;
        }
        __syncthreads();
        const bool armijo = finite && ft <= f0 + 1e-4 * a0 * gd0;
        if (armijo) {
            const int kk = s.k, head = s.head;
            if constexpr (REG)
                lb_load_pairs<P>(ps, py, kk, head, memory, S, Y, n);
            for (int t = threadIdx.x; t < memory * memory; t += kOptBlock) {
                mSY[t] = SYi[t];
                mYY[t] = YYi[t];
            }
            if constexpr (REG)
                lbfgs_accept_pass_reg<P>(n, kk, X, G, rxt, rx, rgt, rg, sv, yv, ps, py, part, st);
            else
                lbfgs_accept_pass<P>(n, memory, kk, head, S, Y, X, G, rxt, rx, rgt, rg, sv, yv, part, tbuf);
            kold = kk;
            __syncthreads();
            const double sy = part_total(part, kPSY), yy = part_total(part, kPYY), ss = part_total(part, kPSS);
            const double gmax = part_max(part, kPGMAX);
            if (sy > 1e-12 * sqrt(ss * yy)) {
                fresh = true;
                drop = kk < memory ? 0 : 1;
                const int ns = kk < memory ? (head + kk) % memory : head;
                lb_store<P>(S + (long long)ns * n, sv, n);
                lb_store<P>(Y + (long long)ns * n, yv, n);
                // the new pair's products with the pairs that stay
                if (threadIdx.x < kk - drop) {
                    const int jo = threadIdx.x + drop;
                    const int sc = (head + jo) % memory;
                    const double v0 = part_total(part, kPPair + 4 * jo), v1 = part_total(part, kPPair + 4 * jo + 1);
                    mSY[sc * memory + ns] = SYi[sc * memory + ns] = v0;
                    mYY[sc * memory + ns] = YYi[sc * memory + ns] = v1;
                    mYY[ns * memory + sc] = YYi[ns * memory + sc] = v1;
                }
                if (threadIdx.x == 0) {
                    mSY[ns * memory + ns] = SYi[ns * memory + ns] = sy;
                    mYY[ns * memory + ns] = YYi[ns * memory + ns] = yy;
                    if (s.k < memory)
                        s.k += 1;
                    else
                        s.head = (s.head + 1) % memory;
                }
            }
            if (threadIdx.x == 0)
                s.f = ft;
            __syncthreads();
            if (s.evals >= maxeval || gmax <= 0.0 || f0 == ft)
                finish = true;
            else
                search = true;
        } else if (s.evals >= maxeval) {
            finish = true;
        } else {
            backtrack = true;
        }
    }
    if (backtrack) {
        const double a = s.alpha;
        double an = 0.5 * a;
        if (finite) {
            const double den = 2.0 * (ft - s.f - a * s.gd);
            if (den > 0.0)
                an = -s.gd * a * a / den;
        }
        an = fmin(fmax(an, 0.1 * a), 0.5 * a);
        if (an < 1e-20) {
            finish = true;
        } else {
            double xv[kLbK], dv[kLbK];
            lb_load<P>(dv, D, n);
#pragma unroll
            for (int k = 0; k < kLbK; ++k)
                xv[k] = rx[k] + an * dv[k];
            lb_store<P>(XT, xv, n);
            __syncthreads();
            if (threadIdx.x == 0)
                s.alpha = an;
        }
    }
    if (search) {
        const int k = s.k, head = s.head;
        double dv[kLbK]; // the direction's elements of this thread
        double sc[REG ? 1 : kLbP][kLbK], yc[REG ? 1 : kLbP][kLbK];
        if constexpr (!REG)
            lb_load_group<P>(sc, yc, 0, k, head, memory, S, Y, n);
        double gamma = 1.0;
        if (k > 0) {
            const int sl = (head + k - 1) % memory;
            gamma = mSY[sl * memory + sl] / mYY[sl * memory + sl];
            if (threadIdx.x < 64)
                block_compact_coeffs(memory, k, head, drop, fresh, part, mSY, mYY, gamma, cu, cw);
            __syncthreads();
        }
        double gd;
        if constexpr (REG)
            gd = lbfgs_compact_direction_reg<P>(n, kold, drop, fresh, k, rgt, ps, py, sv, yv, D, dv, gamma, cu, cw, red,
                                             rpar);
        else
            gd = lbfgs_compact_direction<P>(n, memory, k, head, S, Y, rgt, sc, yc, D, dv, gamma, cu, cw, red, rpar);
        if (!(gd < 0.0)) { // not a descent direction: restart from steepest descent
            double g2 = 0.0;
#pragma unroll
            for (int k2 = 0; k2 < kLbK; ++k2) {
                dv[k2] = -rgt[k2];
                const long long e = lb_e<P>(k2);
                if (e < n)
                    g2 += rgt[k2] * dv[k2];
            }
            lb_store<P>(D, dv, n);
            gd = block_sum_all(g2, red, rpar);
            if (threadIdx.x == 0) {
                s.k = 0;
                s.head = 0;
            }
            first = true;
        }
        double alpha = 1.0;
        if (first) {
            double dmax = 0.0;
#pragma unroll
            for (int k2 = 0; k2 < kLbK; ++k2)
                dmax = fmax(dmax, fabs(dv[k2]));
            dmax = block_max_all(dmax, red, rpar);
            if (dmax > 0.0)
                alpha = fmin(1.0, first_step / dmax);
        }
        // X is XT of this evaluation here (accepted step or first evaluation)
        double xv[kLbK];
#pragma unroll
        for (int k2 = 0; k2 < kLbK; ++k2)
            xv[k2] = rxt[k2] + alpha * dv[k2];
        lb_store<P>(XT, xv, n);
        __syncthreads();
        if (threadIdx.x == 0) {
            s.alpha = alpha;
            s.gd = gd;
            s.phase = 1;
        }
    }
    if (finish) {
        double bx[kLbK];
        lb_load<P>(bx, BX, n);
        __syncthreads();
        lb_store<P>(XT, bx, n);
        if (threadIdx.x == 0) {
            s.phase = 2;
            ob.done[i] = 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ob.st[i] = s;
        ob.best_cost[i] = s.best_f;
    }
}

} // namespace mia_hip

// On-device batched L-BFGS: one thread block per registration instance runs
// one step of the minimizer after each batched cost + gradient evaluation, so
// a whole optimisation iteration (warp, stage 1, reduce, step) stays on the
// GPU and can be captured in a graph.  The algorithm is the one of the host
// LBFGSMinimizer (mia_amd/host/minimizer.hh): the L-BFGS direction (two-loop
// recursion of Nocedal & Wright, Alg. 7.4/7.5, or for instances of up to
// kLbRegN parameters the equal compact form), backtracking Armijo line search with safeguarded
// quadratic interpolation, first step scaled to move no parameter by more
// than `first_step`, `maxeval` counting every evaluation.  Dot products are
// block reductions in a fixed order, so an instance's trajectory does not
// depend on the batch it runs in.
#pragma once
#include "kernels.hh"

namespace mia_hip {

constexpr int kMaxMemory = 16;
// threads per instance: eight waves, so that a CU keeps twice the loads of the
// pair history in flight that four waves would
constexpr int kOptBlock = 512;

struct OptState {
    double f, gd, alpha, best_f;
    int phase;  // 0: first evaluation pending, 1: line search, 2: done
    int evals;
    int k, head; // memory pairs stored, index of the oldest
};

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

// Buffers (doubles) per instance i with n = dim * ncoef parameters at offset
// par_off: X, G, D, BX laid out like the parameters; S and Y hold `memory`
// vectors per instance at offset memory * par_off; RHO memory scalars at
// offset memory * i.  xt (the evaluation's parameter buffer), gt (its gradient
// output) and costs are the evaluation's own buffers.
struct OptBuffers {
    double *X, *G, *D, *BX, *S, *Y, *RHO, *best_cost;
    int *done;
    OptState *st;
    double *SY, *YY; // memory x memory pair products per instance (offset memory^2 * i), by ring slot
};

// Instances of up to kLbRegN parameters keep the block's vectors in registers
// (element e = thread + k * kOptBlock, k < kLbK) and compute the direction in the
// compact form below; larger ones run the two-loop recursion over global
// memory in k_lbfgs_step.
constexpr int kLbK = 4;
constexpr long long kLbRegN = (long long)kLbK * kOptBlock;

__device__ __forceinline__ void lb_load(double *v, const double *__restrict__ src, long long n)
{
#pragma unroll
    for (int k = 0; k < kLbK; ++k) {
        const long long e = threadIdx.x + (long long)k * kOptBlock;
        v[k] = e < n ? src[e] : 0.0;
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
constexpr int kPart = 6 + 4 * kMaxMemory;

// part[] value indices: the new pair's s.y, y.y, s.s, s.g, y.g, max |g|, then per stored pair c (oldest
// first) s_c.y, y_c.y, s_c.g, y_c.g
enum { kPSY = 0, kPYY, kPSS, kPSG, kPYG, kPGMAX, kPPair };

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

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Pairs are read kLbP at a time, all loads of a group issued before any use,
// so that a pass over k pairs waits about k / kLbP memory latencies.
constexpr int kLbP = 8;

__device__ __forceinline__ void lb_store(double *dst, const double *v, long long n)
{
#pragma unroll
    for (int k = 0; k < kLbK; ++k) {
        const long long e = threadIdx.x + (long long)k * kOptBlock;
        if (e < n)
            dst[e] = v[k];
    }
}

// Pairs j0 .. j0 + kLbP - 1 (those below kk) from ring slot head into sc, yc
__device__ __forceinline__ void lb_load_group(double (*sc)[kLbK], double (*yc)[kLbK], int j0, int kk, int head,
                                              int memory, const double *S, const double *Y, long long n)
{
#pragma unroll
    for (int t = 0; t < kLbP; ++t)
        if (j0 + t < kk) {
            const int slot = (head + j0 + t) % memory;
            lb_load(sc[t], S + (long long)slot * n, n);
            lb_load(yc[t], Y + (long long)slot * n, n);
        }
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

// Accepted step, vectors in registers (xt, x, gv, g: this thread's elements of XT, X, GT, G):
// X <- XT, G <- GT, the new pair s = XT - X, y = GT - G left in sv, yv, and the wave partials of every product of kPart written to part[wave][.] (complete after the
// caller's barrier).  kk pairs are stored from ring slot head.
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
    lb_store(X, xt, n);
    lb_store(G, gv, n);
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
        lb_load_group(sc, yc, j0, kk, head, memory, S, Y, n);
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

// Coefficients of the compact direction for k >= 1 pairs from ring slot head (wave 0, lane i for pair i,
// oldest first): cu[i] = p_i, cw[i] = gamma w_i, so that H g = gamma g + sum_i (cu_i s_i - cw_i y_i).
// Pair i's products with g come from part[] at stored index i + drop, or from the new pair's s.g, y.g
// when `fresh` and i == k - 1.
__device__ __forceinline__ void lbfgs_compact_coeffs(int memory, int k, int head, int drop, bool fresh,
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
    // w = R^-1 a (back substitution, column by column)
    double r = a, wi = 0.0;
#pragma unroll
    for (int j = kMaxMemory - 1; j >= 0; --j)
        if (j < k) {
            const int sj = head + j < memory ? head + j : head + j - memory;
            const double wj = readlane_d(r * rinv, j);
            if (il == j)
                wi = wj;
            if (il < j)
                r -= mSY[si * memory + sj] * wj;
        }
    // t = (Dg + gamma Y^T Y) w - gamma b
    double yw = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxMemory; ++j)
        if (j < k) {
            const int sj = head + j < memory ? head + j : head + j - memory;
            yw += mYY[si * memory + sj] * readlane_d(wi, j);
        }
    // p = R^-T t (forward substitution)
    r = mSY[si * memory + si] * wi + gamma * yw - gamma * b;
    double pi = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxMemory; ++j)
        if (j < k) {
            const int sj = head + j < memory ? head + j : head + j - memory;
            const double pj = readlane_d(r * rinv, j);
            if (il == j)
                pi = pj;
            if (il > j)
                r -= mSY[sj * memory + si] * pj;
        }
    if (i < k) {
        cu[i] = pi;
        cw[i] = gamma * wi;
    }
}

// D <- -H g from the coefficients (all threads; g: this thread's gradient elements, sc / yc: the first
// group of pairs, already loaded), also left in dv; returns g.D
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
            lb_load_group(sc, yc, j0, k, head, memory, S, Y, n);
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
        const long long e = threadIdx.x + (long long)k2 * kOptBlock;
        if (e < n)
            gd += g[k2] * dv[k2];
    }
    lb_store(D, dv, n);
    return block_sum_all(gd, red, rpar);
}

__global__ void __launch_bounds__(kOptBlock)
    k_lbfgs_step(const InstDesc *__restrict__ inst, int dim, OptBuffers ob, double *__restrict__ xt,
                 const double *__restrict__ gt, const double *__restrict__ costs, int memory, int maxeval,
                 double first_step)
{
    __shared__ double red[2 * (kOptBlock / 64)];
    __shared__ double alph[kMaxMemory];
    __shared__ double part[kOptBlock / 64][kPart];
    __shared__ double mSY[kMaxMemory * kMaxMemory], mYY[kMaxMemory * kMaxMemory];
    __shared__ double cu[kMaxMemory], cw[kMaxMemory];
    __shared__ double tbuf[(kOptBlock / 64) * 16 * kTbRow];
    int rpar = 0; // which half of red the next block reduction uses
    __shared__ OptState s;
    const int i = blockIdx.x;
    const InstDesc d = inst[i];
    const long long n = (long long)dim * d.cx * d.cy * d.cz;
    const bool regs = n <= kLbRegN; // compact direction with the vectors in registers
    double *SYi = ob.SY + (long long)memory * memory * i, *YYi = ob.YY + (long long)memory * memory * i;
    bool fresh = false; // a pair was stored by this step (regs path)
    int drop = 0;       // ... and replaced the oldest one
    const long long off = d.par_off;
    double *X = ob.X + off, *G = ob.G + off, *D = ob.D + off, *BX = ob.BX + off;
    double *S = ob.S + (long long)memory * off, *Y = ob.Y + (long long)memory * off;
    double *RHO = ob.RHO + (long long)memory * i;
    double *XT = xt + off;
    const double *GT = gt + off;
    // this thread's elements of XT, X, GT, G (regs path), loaded while the state is read
    double rxt[kLbK], rx[kLbK], rgt[kLbK], rg[kLbK];
    if (regs) {
        lb_load(rxt, XT, n);
        lb_load(rgt, GT, n);
        lb_load(rx, X, n);
        lb_load(rg, G, n);
    }
    const double ft = costs[i];
    if (threadIdx.x == 0)
        s = ob.st[i];
    __syncthreads();
    if (s.phase == 2)
        return;
    const bool finite = isfinite(ft);
    bool search = false;      // compute a new direction
    bool first = false;
    bool finish = false;
    bool backtrack = false;
    if (s.phase == 0) {
        for (long long e = threadIdx.x; e < n; e += kOptBlock) {
            X[e] = XT[e];
            G[e] = GT[e];
            BX[e] = XT[e];
        }
        if (threadIdx.x == 0) {
            s.f = ft;
            s.evals = 1;
            s.best_f = finite ? ft : INFINITY;
            s.k = 0;
            s.head = 0;
        }
        __syncthreads();
        double gmax = 0.0;
        for (long long e = threadIdx.x; e < n; e += kOptBlock)
            gmax = fmax(gmax, fabs(GT[e]));
        gmax = block_max_all(gmax, red, rpar);
        if (!finite || gmax <= 0.0 || s.evals >= maxeval)
            finish = true;
        else {
            search = true;
            first = true;
        }
    } else {
        const double f0 = s.f, a0 = s.alpha, gd0 = s.gd;
        const bool better = finite && ft < s.best_f;
        if (better && regs) {
            lb_store(BX, rxt, n);
        } else if (better) {
            for (long long e = threadIdx.x; e < n; e += kOptBlock)
                BX[e] = XT[e];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s.evals += 1;
            if (better)
                s.best_f = ft;
        }
        __syncthreads();
        const bool armijo = finite && ft <= f0 + 1e-4 * a0 * gd0;
        if (armijo && regs) {
            for (int t = threadIdx.x; t < memory * memory; t += kOptBlock) {
                mSY[t] = SYi[t];
                mYY[t] = YYi[t];
            }
            const int kk = s.k, head = s.head;
            double sv[kLbK], yv[kLbK];
            lbfgs_accept_pass(n, memory, kk, head, S, Y, X, G, rxt, rx, rgt, rg, sv, yv, part, tbuf);
            __syncthreads();
            const double sy = part_total(part, kPSY), yy = part_total(part, kPYY), ss = part_total(part, kPSS);
            const double gmax = part_max(part, kPGMAX);
            if (sy > 1e-12 * sqrt(ss * yy)) {
                fresh = true;
                drop = kk < memory ? 0 : 1;
                const int ns = kk < memory ? (head + kk) % memory : head;
                double *Sj = S + (long long)ns * n, *Yj = Y + (long long)ns * n;
#pragma unroll
                for (int k = 0; k < kLbK; ++k) {
                    const long long e = threadIdx.x + (long long)k * kOptBlock;
                    if (e < n) {
                        Sj[e] = sv[k];
                        Yj[e] = yv[k];
                    }
                }
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
                    RHO[ns] = 1.0 / sy;
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
        } else if (armijo) {
            double sy = 0.0, yy = 0.0, ss = 0.0;
            for (long long e = threadIdx.x; e < n; e += kOptBlock) {
                const double se = XT[e] - X[e], ye = GT[e] - G[e];
                sy += se * ye;
                yy += ye * ye;
                ss += se * se;
            }
            sy = block_sum_all(sy, red, rpar);
            yy = block_sum_all(yy, red, rpar);
            ss = block_sum_all(ss, red, rpar);
            if (sy > 1e-12 * sqrt(ss * yy)) {
                const int slot = s.k < memory ? (s.head + s.k) % memory : s.head;
                double *Sj = S + (long long)slot * n, *Yj = Y + (long long)slot * n;
                for (long long e = threadIdx.x; e < n; e += kOptBlock) {
                    Sj[e] = XT[e] - X[e];
                    Yj[e] = GT[e] - G[e];
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    RHO[slot] = 1.0 / sy;
                    if (s.k < memory)
                        s.k += 1;
                    else
                        s.head = (s.head + 1) % memory;
                }
            }
            double gmax = 0.0;
            for (long long e = threadIdx.x; e < n; e += kOptBlock) {
                X[e] = XT[e];
                G[e] = GT[e];
                gmax = fmax(gmax, fabs(GT[e]));
            }
            gmax = block_max_all(gmax, red, rpar);
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
        double a = s.alpha;
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
            if (regs) {
                double xv[kLbK], dv[kLbK];
                lb_load(dv, D, n);
#pragma unroll
                for (int k = 0; k < kLbK; ++k)
                    xv[k] = rx[k] + an * dv[k];
                lb_store(XT, xv, n);
            } else {
                for (long long e = threadIdx.x; e < n; e += kOptBlock)
                    XT[e] = X[e] + an * D[e];
            }
            __syncthreads();
            if (threadIdx.x == 0)
                s.alpha = an;
        }
    }
    if (search) {
        const int k = s.k, head = s.head;
        double gd;
        double dv[kLbK]; // the direction's elements of this thread (regs path)
        if (regs) {
            double sc[kLbP][kLbK], yc[kLbP][kLbK];
            lb_load_group(sc, yc, 0, k, head, memory, S, Y, n);
            double gamma = 1.0;
            if (k > 0) {
                const int sl = (head + k - 1) % memory;
                gamma = mSY[sl * memory + sl] / mYY[sl * memory + sl];
                if (threadIdx.x < 64)
                    lbfgs_compact_coeffs(memory, k, head, drop, fresh, part, mSY, mYY, gamma, cu, cw);
                __syncthreads();
            }
            gd = lbfgs_compact_direction(n, memory, k, head, S, Y, rgt, sc, yc, D, dv, gamma, cu, cw, red, rpar);
        } else {
        // two-loop recursion on q (kept in D), newest pair first
        for (long long e = threadIdx.x; e < n; e += kOptBlock)
            D[e] = G[e];
        __syncthreads();
        for (int j = k - 1; j >= 0; --j) {
            const int slot = (head + j) % memory;
            const double *Sj = S + (long long)slot * n, *Yj = Y + (long long)slot * n;
            double a = 0.0;
            for (long long e = threadIdx.x; e < n; e += kOptBlock)
                a += Sj[e] * D[e];
            a = block_sum_all(a, red, rpar) * RHO[slot];
            if (threadIdx.x == 0)
                alph[j] = a;
            for (long long e = threadIdx.x; e < n; e += kOptBlock)
                D[e] -= a * Yj[e];
            __syncthreads();
        }
        if (k > 0) {
            const int slot = (head + k - 1) % memory;
            const double *Yj = Y + (long long)slot * n;
            double yy = 0.0;
            for (long long e = threadIdx.x; e < n; e += kOptBlock)
                yy += Yj[e] * Yj[e];
            yy = block_sum_all(yy, red, rpar);
            const double gamma = 1.0 / (RHO[slot] * yy);
            for (long long e = threadIdx.x; e < n; e += kOptBlock)
                D[e] *= gamma;
            __syncthreads();
        }
        for (int j = 0; j < k; ++j) {
            const int slot = (head + j) % memory;
            const double *Sj = S + (long long)slot * n, *Yj = Y + (long long)slot * n;
            double bsum = 0.0;
            for (long long e = threadIdx.x; e < n; e += kOptBlock)
                bsum += Yj[e] * D[e];
            const double b = block_sum_all(bsum, red, rpar) * RHO[slot];
            const double c = alph[j] - b;
            for (long long e = threadIdx.x; e < n; e += kOptBlock)
                D[e] += c * Sj[e];
            __syncthreads();
        }
        gd = 0.0;
        for (long long e = threadIdx.x; e < n; e += kOptBlock) {
            D[e] = -D[e];
            gd += G[e] * D[e];
        }
        gd = block_sum_all(gd, red, rpar);
        }
        if (!(gd < 0.0)) { // not a descent direction: restart from steepest descent
            gd = 0.0;
            for (long long e = threadIdx.x; e < n; e += kOptBlock) {
                D[e] = -G[e];
                gd += G[e] * D[e];
            }
            gd = block_sum_all(gd, red, rpar);
            if (threadIdx.x == 0) {
                s.k = 0;
                s.head = 0;
            }
            first = true;
            if (regs)
                lb_load(dv, D, n);
        }
        double alpha = 1.0;
        if (first) {
            double dmax = 0.0;
            for (long long e = threadIdx.x; e < n; e += kOptBlock)
                dmax = fmax(dmax, fabs(D[e]));
            dmax = block_max_all(dmax, red, rpar);
            if (dmax > 0.0)
                alpha = fmin(1.0, first_step / dmax);
        }
        if (regs) { // X is XT of this evaluation here (accepted step or first evaluation)
            double xv[kLbK];
#pragma unroll
            for (int k2 = 0; k2 < kLbK; ++k2)
                xv[k2] = rxt[k2] + alpha * dv[k2];
            lb_store(XT, xv, n);
        } else {
            for (long long e = threadIdx.x; e < n; e += kOptBlock)
                XT[e] = X[e] + alpha * D[e];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s.alpha = alpha;
            s.gd = gd;
            s.phase = 1;
        }
    }
    if (finish) {
        for (long long e = threadIdx.x; e < n; e += kOptBlock)
            XT[e] = BX[e];
        __syncthreads();
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

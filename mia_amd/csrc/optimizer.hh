// On-device batched L-BFGS: one thread block per registration instance runs
// one step of the minimizer after each batched cost + gradient evaluation, so
// a whole optimisation iteration (warp, stage 1, reduce, step) stays on the
// GPU and can be captured in a graph.  The algorithm is the one of the host
// LBFGSMinimizer (mia_amd/host/minimizer.hh): two-loop recursion (Nocedal &
// Wright, Alg. 7.4/7.5), backtracking Armijo line search with safeguarded
// quadratic interpolation, first step scaled to move no parameter by more
// than `first_step`, `maxeval` counting every evaluation.  Dot products are
// block reductions in a fixed order, so an instance's trajectory does not
// depend on the batch it runs in.
#pragma once
#include "kernels.hh"

namespace mia_hip {

constexpr int kMaxMemory = 16;

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
    double *r = red + par * (kBlock / 64);
    par ^= 1;
    if ((threadIdx.x & 63) == 63)
        r[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = r[0];
#pragma unroll
    for (int k = 1; k < kBlock / 64; ++k)
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
};

// Instances of up to kLbRegN parameters run the direction search with the
// block's vectors in registers (element e = thread + k * kBlock, k < kLbK):
// the same operations in the same per-thread order as the global-memory
// loop below (bitwise-equal results), with each pass's S and Y loaded one
// pass ahead, so that only the history is read from memory.
constexpr int kLbK = 8;
constexpr long long kLbRegN = (long long)kLbK * kBlock;

__device__ __forceinline__ void lb_load(double *v, const double *__restrict__ src, long long n)
{
#pragma unroll
    for (int k = 0; k < kLbK; ++k) {
        const long long e = threadIdx.x + (long long)k * kBlock;
        v[k] = e < n ? src[e] : 0.0;
    }
}

// q <- H g (two-loop recursion), D <- -q, returns g.D; pairs `kk` from `head`
__device__ __forceinline__ double lbfgs_search_regs(long long n, int memory, int kk, int head, const double *S,
                                                    const double *Y, const double *RHO, const double *G,
                                                    double *D, double *red, int &rpar, double *alph)
{
    double q[kLbK], g[kLbK], sv[kLbK], yv[kLbK], sn[kLbK], yn[kLbK];
    lb_load(g, G, n);
#pragma unroll
    for (int k = 0; k < kLbK; ++k)
        q[k] = g[k];
    if (kk > 0) {
        const int slot = (head + kk - 1) % memory;
        lb_load(sv, S + (long long)slot * n, n);
        lb_load(yv, Y + (long long)slot * n, n);
    }
    // newest pair first
    for (int j = kk - 1; j >= 0; --j) {
        const int slot = (head + j) % memory;
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < kLbK; ++k)
            a += sv[k] * q[k];
        if (j > 0) {
            const int nslot = (head + j - 1) % memory;
            lb_load(sn, S + (long long)nslot * n, n);
            lb_load(yn, Y + (long long)nslot * n, n);
        }
        a = block_sum_all(a, red, rpar) * RHO[slot];
        if (threadIdx.x == 0)
            alph[j] = a;
#pragma unroll
        for (int k = 0; k < kLbK; ++k) {
            q[k] -= a * yv[k];
            sv[k] = sn[k];
            yv[k] = yn[k];
        }
    }
    if (kk > 0) {
        const int slot = (head + kk - 1) % memory;
        lb_load(yv, Y + (long long)slot * n, n);
        double yy = 0.0;
#pragma unroll
        for (int k = 0; k < kLbK; ++k)
            yy += yv[k] * yv[k];
        yy = block_sum_all(yy, red, rpar);
        const double gamma = 1.0 / (RHO[slot] * yy);
#pragma unroll
        for (int k = 0; k < kLbK; ++k)
            q[k] *= gamma;
        lb_load(sv, S + (long long)(head % memory) * n, n);
        lb_load(yv, Y + (long long)(head % memory) * n, n);
    }
    __syncthreads(); // alph[] complete
    for (int j = 0; j < kk; ++j) {
        const int slot = (head + j) % memory;
        double bsum = 0.0;
#pragma unroll
        for (int k = 0; k < kLbK; ++k)
            bsum += yv[k] * q[k];
        if (j + 1 < kk) {
            const int nslot = (head + j + 1) % memory;
            lb_load(sn, S + (long long)nslot * n, n);
            lb_load(yn, Y + (long long)nslot * n, n);
        }
        const double b = block_sum_all(bsum, red, rpar) * RHO[slot];
        const double c = alph[j] - b;
#pragma unroll
        for (int k = 0; k < kLbK; ++k) {
            q[k] += c * sv[k];
            sv[k] = sn[k];
            yv[k] = yn[k];
        }
    }
    double gd = 0.0;
#pragma unroll
    for (int k = 0; k < kLbK; ++k) {
        const long long e = threadIdx.x + (long long)k * kBlock;
        if (e < n) {
            const double dk = -q[k];
            D[e] = dk;
            gd += g[k] * dk;
        }
    }
    return block_sum_all(gd, red, rpar);
}

__global__ void __launch_bounds__(kBlock)
    k_lbfgs_step(const InstDesc *__restrict__ inst, int dim, OptBuffers ob, double *__restrict__ xt,
                 const double *__restrict__ gt, const double *__restrict__ costs, int memory, int maxeval,
                 double first_step)
{
    __shared__ double red[2 * (kBlock / 64)];
    __shared__ double alph[kMaxMemory];
    int rpar = 0; // which half of red the next block reduction uses
    __shared__ OptState s;
    const int i = blockIdx.x;
    const InstDesc d = inst[i];
    const long long n = (long long)dim * d.cx * d.cy * d.cz;
    const long long off = d.par_off;
    double *X = ob.X + off, *G = ob.G + off, *D = ob.D + off, *BX = ob.BX + off;
    double *S = ob.S + (long long)memory * off, *Y = ob.Y + (long long)memory * off;
    double *RHO = ob.RHO + (long long)memory * i;
    double *XT = xt + off;
    const double *GT = gt + off;
    if (threadIdx.x == 0)
        s = ob.st[i];
    __syncthreads();
    if (s.phase == 2)
        return;
    const double ft = costs[i];
    const bool finite = isfinite(ft);
    bool search = false;      // compute a new direction
    bool first = false;
    bool finish = false;
    bool backtrack = false;
    if (s.phase == 0) {
        for (long long e = threadIdx.x; e < n; e += kBlock) {
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
        for (long long e = threadIdx.x; e < n; e += kBlock)
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
        if (better)
            for (long long e = threadIdx.x; e < n; e += kBlock)
                BX[e] = XT[e];
        __syncthreads();
        if (threadIdx.x == 0) {
            s.evals += 1;
            if (better)
                s.best_f = ft;
        }
        __syncthreads();
        const bool armijo = finite && ft <= f0 + 1e-4 * a0 * gd0;
        if (armijo) {
            double sy = 0.0, yy = 0.0, ss = 0.0;
            for (long long e = threadIdx.x; e < n; e += kBlock) {
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
                for (long long e = threadIdx.x; e < n; e += kBlock) {
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
            for (long long e = threadIdx.x; e < n; e += kBlock) {
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
            for (long long e = threadIdx.x; e < n; e += kBlock)
                XT[e] = X[e] + an * D[e];
            __syncthreads();
            if (threadIdx.x == 0)
                s.alpha = an;
        }
    }
    if (search) {
        const int k = s.k, head = s.head;
        double gd;
        if (n <= kLbRegN) {
            gd = lbfgs_search_regs(n, memory, k, head, S, Y, RHO, G, D, red, rpar, alph);
        } else {
        // two-loop recursion on q (kept in D), newest pair first
        for (long long e = threadIdx.x; e < n; e += kBlock)
            D[e] = G[e];
        __syncthreads();
        for (int j = k - 1; j >= 0; --j) {
            const int slot = (head + j) % memory;
            const double *Sj = S + (long long)slot * n, *Yj = Y + (long long)slot * n;
            double a = 0.0;
            for (long long e = threadIdx.x; e < n; e += kBlock)
                a += Sj[e] * D[e];
            a = block_sum_all(a, red, rpar) * RHO[slot];
            if (threadIdx.x == 0)
                alph[j] = a;
            for (long long e = threadIdx.x; e < n; e += kBlock)
                D[e] -= a * Yj[e];
            __syncthreads();
        }
        if (k > 0) {
            const int slot = (head + k - 1) % memory;
            const double *Yj = Y + (long long)slot * n;
            double yy = 0.0;
            for (long long e = threadIdx.x; e < n; e += kBlock)
                yy += Yj[e] * Yj[e];
            yy = block_sum_all(yy, red, rpar);
            const double gamma = 1.0 / (RHO[slot] * yy);
            for (long long e = threadIdx.x; e < n; e += kBlock)
                D[e] *= gamma;
            __syncthreads();
        }
        for (int j = 0; j < k; ++j) {
            const int slot = (head + j) % memory;
            const double *Sj = S + (long long)slot * n, *Yj = Y + (long long)slot * n;
            double bsum = 0.0;
            for (long long e = threadIdx.x; e < n; e += kBlock)
                bsum += Yj[e] * D[e];
            const double b = block_sum_all(bsum, red, rpar) * RHO[slot];
            const double c = alph[j] - b;
            for (long long e = threadIdx.x; e < n; e += kBlock)
                D[e] += c * Sj[e];
            __syncthreads();
        }
        gd = 0.0;
        for (long long e = threadIdx.x; e < n; e += kBlock) {
            D[e] = -D[e];
            gd += G[e] * D[e];
        }
        gd = block_sum_all(gd, red, rpar);
        }
        if (!(gd < 0.0)) { // not a descent direction: restart from steepest descent
            gd = 0.0;
            for (long long e = threadIdx.x; e < n; e += kBlock) {
                D[e] = -G[e];
                gd += G[e] * D[e];
            }
            gd = block_sum_all(gd, red, rpar);
            if (threadIdx.x == 0) {
                s.k = 0;
                s.head = 0;
            }
            first = true;
        }
        double alpha = 1.0;
        if (first) {
            double dmax = 0.0;
            for (long long e = threadIdx.x; e < n; e += kBlock)
                dmax = fmax(dmax, fabs(D[e]));
            dmax = block_max_all(dmax, red, rpar);
            if (dmax > 0.0)
                alpha = fmin(1.0, first_step / dmax);
        }
        for (long long e = threadIdx.x; e < n; e += kBlock)
            XT[e] = X[e] + alpha * D[e];
        __syncthreads();
        if (threadIdx.x == 0) {
            s.alpha = alpha;
            s.gd = gd;
            s.phase = 1;
        }
    }
    if (finish) {
        for (long long e = threadIdx.x; e < n; e += kBlock)
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

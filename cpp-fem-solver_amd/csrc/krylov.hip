// krylov.hip — PCG, BiCGSTAB and restarted GMRES on the device.
//
// Algorithms and stopping rules follow the reference exactly:
//   pcg       ode-solver/src/math/array.cpp:3445-3601
//   bicgstab  ode-solver/src/math/array.cpp:3602-3842 (true-residual replacement)
//   gmres     ode-solver/src/math/array.cpp:4190-4401 (right preconditioning,
//             modified Gram-Schmidt, Givens rotations)
// The scalars live in a device KState. A reduction writes per-block partial
// sums; its tail (a one-block kernel, or every block of the next kernel
// re-adding them in the same order) forms the total, and one thread evaluates
// the scalar recurrences, the breakdown tests and the stopping test. All
// kernels are gated on the state,
// so a block of iterations can be captured once as a hipGraph and replayed
// until the device reports completion; the host reads back 4 bytes per
// replay. Reductions are deterministic (fixed grid and order).
//
// FP32_IR: the chosen Krylov method runs in float32 on the scaled residual
// inside float64 iterative refinement x += ||r|| d, r = b - A x (float64).
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "reduce.hpp"
#include "solver.hpp"

namespace fem {

enum PcMode : int { PCM_GEN = 0, PCM_NONE = 1, PCM_DIAG = 2 };

// FP32_IR: scaled inner residual below which a float32 breakdown counts as
// round-off near convergence rather than a failure of the method
constexpr double kSoftExit = 1e-3;

// PCG with M = G^T G: most partial sums re-read in all by the blocks of one
// of its fused kernels (blocks x partial sums per block)
constexpr uint64_t kFacFold = 8ull << 20;

__device__ __forceinline__ void kfail(KState* st, int code) {
    if (st->soft) { st->done = 1; st->err = -code; }
    else { st->done = 2; st->err = code; }
}

#define GRID_LOOP(i, n) for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)
#define DONE_RET(st) if (flag_load(&(st)->done)) return

// ---------------------------------------------------------------------------
// PCG kernels. Each reduction kernel writes per-block partials; a Fin*
// functor, run by the one-block k_finalize kernel, evaluates the scalar
// recurrence on the totals. Functors repeat the gate of their kernel.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_pcg_init(const T* b, T* x, T* r, size_t n, KState* st, double* part) {
    DONE_RET(st);
    double s = 0.0;
    GRID_LOOP(i, n) {
        const T bi = b[i];
        x[i] = (T)0;
        r[i] = bi;
        s += (double)bi * (double)bi;
    }
    double v[1] = {s};
    block_partials<1>(v, part);
}
struct FinPcgInit {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done) return;
        st->normb = sqrt(t[0]);
        st->normr = st->normb;
        st->it = 0;
        st->it_end = 0;
        st->iters = 0;
        if (st->normb < 1e-300) { st->done = 1; return; }
        if (!st->minicg && st->normr / st->normb < st->tol) { st->done = 1; return; }
    }
};

template <typename T, int PCM>
__device__ __forceinline__ T zval(const T* r, const T* z, const T* d, size_t i) {
    if constexpr (PCM == PCM_NONE) return r[i];
    else if constexpr (PCM == PCM_DIAG) return r[i] / d[i];
    else return z[i];
}

// z (given or implied) -> p = z, rho = r^T z
template <typename T, int PCM>
__global__ void k_pcg_init_rz(const T* r, const T* z, const T* d, T* p, size_t n, KState* st, double* part) {
    DONE_RET(st);
    double s = 0.0;
    GRID_LOOP(i, n) {
        const T zi = zval<T, PCM>(r, z, d, i);
        p[i] = zi;
        s += (double)r[i] * (double)zi;
    }
    double v[1] = {s};
    block_partials<1>(v, part);
}
struct FinPcgInitRz {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done) return;
        if (t[0] <= 0.0 && !st->minicg) { kfail(st, SE_PCG_RZ); return; }
        st->rz = t[0];
    }
};

template <typename T>
__global__ void k_dot2(const T* a, const T* b, size_t n, KState* st, double* part) {
    DONE_RET(st);
    double s = 0.0;
    GRID_LOOP(i, n) s += (double)a[i] * (double)b[i];
    double v[1] = {s};
    block_partials<1>(v, part);
}
struct FinPcgPq {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done) return;
        const double pap = t[0];
        st->pap = pap;
        st->iters += 1;
        if (!st->minicg) {
            if (fabs(pap) < 1e-30) { kfail(st, SE_PCG_PAP_SMALL); return; }
            if (pap <= 0.0) { kfail(st, SE_PCG_PAP_NEG); return; }
        }
        st->alpha = st->rz / pap;
    }
};

__device__ __forceinline__ void pcg_finish(KState* st, double rz_new) {
    if (rz_new <= 0.0 && !st->minicg) { kfail(st, SE_PCG_RZ); return; }
    st->beta = rz_new / st->rz;
    st->rz = rz_new;
    st->it_end = st->it;
    st->it += 1;
    if (st->it >= st->max_iter) st->done = 3;
}

// x += alpha p ; r -= alpha q ; ||r|| stopping test (and rho for none/diag)
template <typename T, int PCM>
__global__ void k_pcg_xr(T* x, T* r, const T* p, const T* q, const T* d, size_t n, KState* st, double* part) {
    DONE_RET(st);
    const T a = (T)st->alpha;
    double rr = 0.0, rz = 0.0;
    GRID_LOOP(i, n) {
        x[i] += a * p[i];
        const T ri = r[i] - a * q[i];
        r[i] = ri;
        rr += (double)ri * (double)ri;
        if constexpr (PCM == PCM_DIAG) rz += (double)ri * (double)(ri / d[i]);
    }
    if constexpr (PCM == PCM_NONE) rz = rr;
    double v[2] = {rr, rz};
    block_partials<2>(v, part);
}
template <int PCM>
struct FinPcgXr {
    KState* st;
    __device__ void operator()(double (&t)[2]) const {
        if (st->done) return;
        st->normr = sqrt(t[0]);
        if (st->normr / st->normb < st->tol) { st->done = 1; return; }
        if constexpr (PCM != PCM_GEN) pcg_finish(st, t[1]);
        else st->rz_prev = st->rz;
    }
};

// k_pcg_xr with the p.q reduction tail folded in: every block adds the
// product's npq partial sums exactly as the one-block tail k_finalize<1>
// does (same order, same block size), block 0 runs that tail's state update
// (FinPcgPq), and every block uses alpha = rz / p.q, the value FinPcgPq
// stores. st->rz is not written in this kernel. Block 0 may set st->done
// here (breakdown); the other blocks do not rely on seeing it: each runs the
// same breakdown test on its own copy of p.q and leaves x and r untouched,
// as the unfused path (whose update kernel is gated off) does.
template <typename T, int PCM>
__global__ __launch_bounds__(kBlock) void k_pcg_xr_f(T* x, T* r, const T* p, const T* q, const T* d, size_t n,
                                                     KState* st, double* part, const double* __restrict__ pq_part,
                                                     unsigned npq) {
    DONE_RET(st);
    double t[1] = {0.0};
    add_partials<1>(pq_part, npq, t);
    block_sum<1>(t);
    if (blockIdx.x == 0 && threadIdx.x == 0) FinPcgPq{st}(t);
    if (!st->minicg && (fabs(t[0]) < 1e-30 || t[0] <= 0.0)) return;
    const T a = (T)(st->rz / t[0]);
    double rr = 0.0, rz = 0.0;
    GRID_LOOP(i, n) {
        x[i] += a * p[i];
        const T ri = r[i] - a * q[i];
        r[i] = ri;
        rr += (double)ri * (double)ri;
        if constexpr (PCM == PCM_DIAG) rz += (double)ri * (double)(ri / d[i]);
    }
    if constexpr (PCM == PCM_NONE) rz = rr;
    double v[2] = {rr, rz};
    block_partials<2>(v, part);
}

struct FinPcgRz {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done) return;
        pcg_finish(st, t[0]);
    }
};

template <typename T, int PCM>
__global__ void k_pcg_p(const T* r, const T* z, const T* d, T* p, size_t n, KState* st) {
    DONE_RET(st);
    const T b = (T)st->beta;
    GRID_LOOP(i, n) p[i] = zval<T, PCM>(r, z, d, i) + b * p[i];
}

// k_pcg_p with the r.z reduction tail folded in (as k_pcg_xr_f): block 0
// runs FinPcgRz, every block uses beta = r.z / rz_prev (FinPcgXr saved the
// previous rz there; st->rz itself changes in this kernel)
template <typename T, int PCM>
__global__ __launch_bounds__(kBlock) void k_pcg_p_f(const T* r, const T* z, const T* d, T* p, size_t n, KState* st,
                                                    const double* __restrict__ rz_part, unsigned nrz) {
    DONE_RET(st);
    double t[1] = {0.0};
    add_partials<1>(rz_part, nrz, t);
    block_sum<1>(t);
    if (blockIdx.x == 0 && threadIdx.x == 0) FinPcgRz{st}(t);
    const T b = (T)(t[0] / st->rz_prev);
    GRID_LOOP(i, n) p[i] = zval<T, PCM>(r, z, d, i) + b * p[i];
}

// ---------------------------------------------------------------------------
// PCG with a factorized sparse approximate inverse M = G^T G in three
// launches per iteration instead of six (q = A p; x/r update; ||r|| tail;
// t = G r; z = G^T t; p update):
//   1. q = A p with the partial sums of p.q (the operator's fused product);
//   2. k_pcg_fac_t, over G's CSR-stream blocking: the p.q tail (every block
//      re-adds the partial sums in the one-block tail's order, block 0 stores
//      the state), then t = G (r - alpha q) with r - alpha q formed in the
//      gather (r and q are only read here), and per block the partial sums of
//      ||r - alpha q||^2 and ||t||^2 = r^T M r over its rows;
//   3. k_pcg_fac_p, over G^T's blocking: per row x += alpha p and
//      r -= alpha q; then both tails, and unless ||r|| met the tolerance,
//      p = G^T t + beta p.
// x, r, t, p and r.z equal those of the six-launch sequence bit for bit
// (same expressions, same summation order); ||r||^2 is summed over G's row
// blocks instead of the vector kernels' grid, so ||r|| may differ in its
// last bits.
// ---------------------------------------------------------------------------
template <typename T, bool PK>
__global__ __launch_bounds__(kBlock) void k_pcg_fac_t(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                                      const uint16_t* __restrict__ cd,
                                                      const uint32_t* __restrict__ cbase, const T* __restrict__ v,
                                                      const uint2* __restrict__ bnd, const T* __restrict__ r,
                                                      const T* __restrict__ q, T* __restrict__ t, KState* st,
                                                      const double* __restrict__ pq_part, unsigned npq,
                                                      double* __restrict__ part) {
    constexpr int STAGE = stage_nnz<T>();
    __shared__ T prod[STAGE];
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
    const uint2 lo = bnd[bid], hi = bnd[bid + 1];
    if (flag_load(&st->done)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->fac_go = 0;
        return;
    }
    double pq[1] = {0.0};
    add_partials<1>(pq_part, npq, pq);
    block_sum<1>(pq);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        FinPcgPq{st}(pq);
        st->rz_prev = st->rz;  // beta's denominator: block 0 of k_pcg_fac_p rewrites rz
        st->fac_go = st->done ? 0 : 1;
    }
    if (!st->minicg && (fabs(pq[0]) < 1e-30 || pq[0] <= 0.0)) return;
    const T a = (T)(st->rz / pq[0]);
    const auto gx = [r, q, a](uint32_t c) { return r[c] - a * q[c]; };
    const uint32_t r0 = lo.x, r1 = hi.x, k0 = lo.y, k1 = hi.y;
    const uint32_t cnt = k1 - k0;
    const uint32_t tid = threadIdx.x;
    double s2[2] = {0.0, 0.0};  // ||r - alpha q||^2, ||t||^2 over the block's rows
    if (stage_fits<STAGE>(k0, cnt)) {
        constexpr int RPT = (stage_rows<T>() + kBlock - 1) / kBlock;
        uint32_t rb[RPT], re[RPT];
        T rn[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const uint32_t row = r0 + tid + (uint32_t)j * kBlock;
            rb[j] = re[j] = 0;
            rn[j] = (T)0;
            if (row < r1) {
                rb[j] = rp[row] - k0;
                re[j] = rp[row + 1] - k0;
                rn[j] = gx(row);
            }
        }
        if (cnt > 0) stage_products_g<T, STAGE, PK>(ci, cd, cbase, v, gx, k0, cnt, prod);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const uint32_t row = r0 + tid + (uint32_t)j * kBlock;
            if (row < r1) {
                T s = (T)0;
                for (uint32_t k = rb[j]; k < re[j]; ++k) s += prod[k];
                t[row] = s;
                s2[0] += (double)rn[j] * (double)rn[j];
                s2[1] += (double)s * (double)s;
            }
        }
    } else {
        // one row longer than the stage: strided products, block tree
        double s = 0.0;
        for (uint32_t k = k0 + tid; k < k1; k += kBlock) s += (double)(v[k] * gx(col_of<PK>(ci, cd, cbase, k)));
        double tot[1] = {s};
        block_sum<1>(tot);
        if (tid == 0) {
            const T tv = (T)tot[0], rv = gx(r0);
            t[r0] = tv;
            s2[0] = (double)rv * (double)rv;
            s2[1] = (double)tv * (double)tv;
        }
    }
    block_sum<2>(s2);
    if (tid == 0) {
        part[2 * bid] = s2[0];
        part[2 * bid + 1] = s2[1];
    }
}

// block 0 of k_pcg_fac_p: the ||r|| stopping test (FinPcgXr) and, if the
// solve goes on, the r.z recurrence (FinPcgRz); rz_prev is left as
// k_pcg_fac_t set it
struct FinPcgFac {
    KState* st;
    __device__ void operator()(double (&t)[2]) const {
        if (st->done) return;
        st->normr = sqrt(t[0]);
        if (st->normr / st->normb < st->tol) { st->done = 1; return; }
        pcg_finish(st, t[1]);
    }
};

template <typename T, bool PK>
__global__ __launch_bounds__(kBlock) void k_pcg_fac_p(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                                      const uint16_t* __restrict__ cd,
                                                      const uint32_t* __restrict__ cbase, const T* __restrict__ v,
                                                      const uint2* __restrict__ bnd, const T* __restrict__ t, T* x,
                                                      T* r, T* p, const T* q, KState* st,
                                                      const double* __restrict__ fpart, unsigned nf) {
    constexpr int STAGE = stage_nnz<T>();
    __shared__ T prod[STAGE];
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
    const uint2 lo = bnd[bid], hi = bnd[bid + 1];
    // gated on fac_go, which only k_pcg_fac_t writes: block 0 below may set
    // done, and a block that starts after that must still update x and r
    if (!flag_load(&st->fac_go)) return;
    const T a = (T)st->alpha;
    const uint32_t r0 = lo.x, r1 = hi.x, k0 = lo.y, k1 = hi.y;
    const uint32_t cnt = k1 - k0;
    const uint32_t tid = threadIdx.x;
    // the products and the rows' operands do not depend on the tails: their
    // loads are issued first, the tails' partial sums are read meanwhile
    if (stage_fits<STAGE>(k0, cnt)) {
        constexpr int RPT = (stage_rows<T>() + kBlock - 1) / kBlock;
        uint32_t rb[RPT], re[RPT];
        T xv[RPT], rv[RPT], pv[RPT], qv[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const uint32_t row = r0 + tid + (uint32_t)j * kBlock;
            rb[j] = re[j] = 0;
            xv[j] = rv[j] = pv[j] = qv[j] = (T)0;
            if (row < r1) {
                rb[j] = rp[row] - k0;
                re[j] = rp[row + 1] - k0;
                xv[j] = x[row];
                rv[j] = r[row];
                pv[j] = p[row];
                qv[j] = q[row];
            }
        }
        if (cnt > 0) stage_products<T, STAGE, PK>(ci, cd, cbase, v, t, k0, cnt, prod);
        double s2[2] = {0.0, 0.0};
        add_partials<2>(fpart, nf, s2);
        block_sum<2>(s2);  // its barriers also order the LDS stores of prod
        if (blockIdx.x == 0 && tid == 0) FinPcgFac{st}(s2);
        const bool conv = sqrt(s2[0]) / st->normb < st->tol;
        const T b = (T)(s2[1] / st->rz_prev);
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const uint32_t row = r0 + tid + (uint32_t)j * kBlock;
            if (row < r1) {
                x[row] = xv[j] + a * pv[j];
                r[row] = rv[j] - a * qv[j];
                if (!conv) {
                    T s = (T)0;
                    for (uint32_t k = rb[j]; k < re[j]; ++k) s += prod[k];
                    p[row] = s + b * pv[j];
                }
            }
        }
    } else {
        double s = 0.0;
        for (uint32_t k = k0 + tid; k < k1; k += kBlock) s += (double)(v[k] * t[col_of<PK>(ci, cd, cbase, k)]);
        double tot[1] = {s};
        block_sum<1>(tot);
        double s2[2] = {0.0, 0.0};
        add_partials<2>(fpart, nf, s2);
        block_sum<2>(s2);
        if (blockIdx.x == 0 && tid == 0) FinPcgFac{st}(s2);
        const bool conv = sqrt(s2[0]) / st->normb < st->tol;
        const T b = (T)(s2[1] / st->rz_prev);
        if (tid == 0) {
            const T pv = p[r0];
            x[r0] += a * pv;
            r[r0] = r[r0] - a * q[r0];
            if (!conv) p[r0] = (T)tot[0] + b * pv;
        }
    }
}

// ---------------------------------------------------------------------------
// BiCGSTAB kernels
// ---------------------------------------------------------------------------
constexpr double kBT = 1e-30;

template <typename T>
__global__ void k_bi_init(const T* b, T* x, T* r, T* rt, T* p, T* v, size_t n, KState* st, double* part) {
    DONE_RET(st);
    double s = 0.0;
    GRID_LOOP(i, n) {
        const T bi = b[i];
        x[i] = (T)0;
        r[i] = bi;
        rt[i] = bi;
        p[i] = (T)0;
        v[i] = (T)0;
        s += (double)bi * (double)bi;
    }
    double w[1] = {s};
    block_partials<1>(w, part);
}
struct FinBiInit {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done) return;
        st->normb = sqrt(t[0]);
        st->normr = st->normb;
        st->it = 0;
        st->it_end = 0;
        st->iters = 0;
        st->rz = 1.0;  // rho_old
        st->alpha = 1.0;
        st->omega = 1.0;
        st->check = 0;
        st->replace = 0;
        if (st->normb < 1e-300) { st->done = 1; return; }
        if (st->normr / st->normb < st->tol) { st->done = 1; return; }
    }
};

struct FinBiRho {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done) return;
        const double rho = t[0];
        st->rho = rho;
        if (fabs(rho) < kBT) { kfail(st, SE_BI_RHO); return; }
        if (fabs(st->rz) < kBT) { kfail(st, SE_BI_RHO_OLD); return; }
        if (fabs(st->omega) < kBT) { kfail(st, SE_BI_OMEGA); return; }
        st->beta = (rho / st->rz) * (st->alpha / st->omega);
    }
};

template <typename T>
__global__ void k_bi_p(const T* r, const T* v, T* p, size_t n, KState* st) {
    DONE_RET(st);
    const T b = (T)st->beta, om = (T)st->omega;
    GRID_LOOP(i, n) p[i] = r[i] + b * (p[i] - om * v[i]);
}

struct FinBiRv {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done) return;
        st->rv = t[0];
        st->iters += 1;
        if (fabs(t[0]) < kBT) { kfail(st, SE_BI_RV); return; }
        st->alpha = st->rho / t[0];
    }
};

template <typename T>
__global__ void k_bi_s(const T* r, const T* v, T* sv, size_t n, KState* st) {
    DONE_RET(st);
    const T a = (T)st->alpha;
    GRID_LOOP(i, n) sv[i] = r[i] - a * v[i];
}

template <typename T>
__global__ void k_bi_tt(const T* t_, const T* sv, size_t n, KState* st, double* part) {
    DONE_RET(st);
    double a = 0.0, b = 0.0;
    GRID_LOOP(i, n) {
        const double ti = (double)t_[i];
        a += ti * ti;
        b += ti * (double)sv[i];
    }
    double w[2] = {a, b};
    block_partials<2>(w, part);
}
struct FinBiTt {
    KState* st;
    __device__ void operator()(double (&t)[2]) const {
        if (st->done) return;
        st->tt = t[0];
        st->ts = t[1];
        if (t[0] < kBT) { kfail(st, SE_BI_TT); return; }
        st->omega = t[1] / t[0];
    }
};

template <typename T>
__global__ void k_bi_xr(T* x, T* r, const T* y, const T* z, const T* sv, const T* t_, size_t n, KState* st,
                        double* part) {
    DONE_RET(st);
    const T a = (T)st->alpha, om = (T)st->omega;
    double rr = 0.0;
    GRID_LOOP(i, n) {
        x[i] += a * y[i] + om * z[i];
        const T ri = sv[i] - om * t_[i];
        r[i] = ri;
        rr += (double)ri * (double)ri;
    }
    double w[1] = {rr};
    block_partials<1>(w, part);
}
struct FinBiXr {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done) return;
        st->normr = sqrt(t[0]);
        st->it_end = st->it;
        if (st->normr / st->normb < st->tol) st->check = 1;
        else { st->check = 0; st->rz = st->rho; }
    }
};

// true residual b - A x  (ax = A x computed before, gated on check)
template <typename T>
__global__ void k_bi_true(const T* b, const T* ax, size_t n, KState* st, double* part) {
    DONE_RET(st);
    if (!st->check) return;
    double s = 0.0;
    GRID_LOOP(i, n) {
        const double d = (double)b[i] - (double)ax[i];
        s += d * d;
    }
    double w[1] = {s};
    block_partials<1>(w, part);
}
struct FinBiTrue {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done || !st->check) return;
        st->true_normr = sqrt(t[0]);
        if (st->true_normr / st->normb < st->tol) { st->done = 1; st->replace = 0; }
        else { st->replace = 1; st->normr = st->true_normr; }
    }
};

template <typename T>
__global__ void k_bi_replace(const T* b, const T* ax, const T* rt, T* r, size_t n, KState* st, double* part) {
    DONE_RET(st);
    if (!st->replace) return;
    double s = 0.0;
    GRID_LOOP(i, n) {
        const T ri = b[i] - ax[i];
        r[i] = ri;
        s += (double)rt[i] * (double)ri;
    }
    double w[1] = {s};
    block_partials<1>(w, part);
}
struct FinBiReplace {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done || !st->replace) return;
        if (fabs(t[0]) < kBT) { st->done = 1; return; }
        st->rz = t[0];
    }
};

__global__ void k_bi_tail(KState* st) {
    DONE_RET(st);
    st->check = 0;
    st->replace = 0;
    st->it += 1;
    if (st->it >= st->max_iter) st->done = 3;
}

// ---------------------------------------------------------------------------
// GMRES kernels; gm = [H (m+1) x m | cs m | sn m | g m+1 | y m]
// ---------------------------------------------------------------------------
struct GmLayout {
    int m;
    __host__ __device__ double* H(double* gm) const { return gm; }
    __host__ __device__ double* cs(double* gm) const { return gm + (size_t)(m + 1) * m; }
    __host__ __device__ double* sn(double* gm) const { return cs(gm) + m; }
    __host__ __device__ double* g(double* gm) const { return sn(gm) + m; }
    __host__ __device__ double* y(double* gm) const { return g(gm) + m + 1; }
    static size_t size(int m) { return (size_t)(m + 1) * m + 4 * (size_t)m + 1; }
};

template <typename T>
__global__ void k_gm_init(const T* b, T* x, size_t n, KState* st, double* part) {
    DONE_RET(st);
    double s = 0.0;
    GRID_LOOP(i, n) {
        x[i] = (T)0;
        s += (double)b[i] * (double)b[i];
    }
    double w[1] = {s};
    block_partials<1>(w, part);
}
struct FinGmInit {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done) return;
        st->normb = sqrt(t[0]);
        st->it = 0;
        st->it_end = 0;
        st->iters = 0;
        if (st->normb < 1e-300) st->done = 1;
    }
};

__global__ void k_gm_cycle_begin(KState* st) {
    DONE_RET(st);
    st->live = st->it < st->max_iter ? 1 : 0;
    st->need_true = st->it > 0 ? 1 : 0;
    st->conv = 0;
    st->j_used = 0;
}

template <typename T>
__global__ void k_gm_v0(const T* b, const T* ax, T* v0, size_t n, KState* st, double* part) {
    DONE_RET(st);
    const int nt = st->need_true;
    double s = 0.0;
    GRID_LOOP(i, n) {
        const T ri = nt ? (T)(b[i] - ax[i]) : b[i];
        v0[i] = ri;
        s += (double)ri * (double)ri;
    }
    double w[1] = {s};
    block_partials<1>(w, part);
}
struct FinGmV0 {
    KState* st;
    double* gm;
    GmLayout L;
    __device__ void operator()(double (&t)[1]) const {
        if (st->done) return;
        const int nt = st->need_true;
        const double beta = nt ? sqrt(t[0]) : st->normb;
        st->beta_g = beta;
        if (nt && beta / st->normb < st->tol) { st->done = 1; return; }
        double* g = L.g(gm);
        for (int i = 0; i <= L.m; ++i) g[i] = 0.0;
        g[0] = beta;
        double* H = L.H(gm);
        for (int i = 0; i < (L.m + 1) * L.m; ++i) H[i] = 0.0;
    }
};

template <typename T>
__global__ void k_gm_scale(T* v0, size_t n, KState* st) {
    DONE_RET(st);
    const T beta = (T)st->beta_g;
    GRID_LOOP(i, n) v0[i] = v0[i] / beta;
}

__device__ __forceinline__ bool gm_closed(const KState* st) {
    return *(volatile const int*)&st->done || !*(volatile const int*)&st->live;
}

// partial of w . V_0 (h_{0j})
template <typename T>
__global__ void k_gm_dot0(const T* w, const T* v0, size_t n, KState* st, double* part) {
    if (gm_closed(st)) return;
    double s = 0.0;
    GRID_LOOP(i, n) s += (double)w[i] * (double)v0[i];
    double a[1] = {s};
    block_partials<1>(a, part);
}
// store h_{ij}
struct FinGmH {
    KState* st;
    double* gm;
    GmLayout L;
    int i_, j;
    __device__ void operator()(double (&t)[1]) const {
        if (gm_closed(st)) return;
        L.H(gm)[i_ * L.m + j] = t[0];
    }
};

// w -= h_{i-1,j} V_{i-1}; partial of w . V_i
template <typename T>
__global__ void k_gm_axpy_dot(T* w, const T* vprev, const T* vi, size_t n, int i_, int j, KState* st, double* gm,
                              GmLayout L, double* part) {
    if (gm_closed(st)) return;
    const T h = (T)L.H(gm)[(i_ - 1) * L.m + j];
    double s = 0.0;
    GRID_LOOP(i, n) {
        const T wi = w[i] - h * vprev[i];
        w[i] = wi;
        s += (double)wi * (double)vi[i];
    }
    double a[1] = {s};
    block_partials<1>(a, part);
}

// w -= h_{jj} V_j; partial of ||w||^2
template <typename T>
__global__ void k_gm_axpy_norm(T* w, const T* vj, size_t n, int j, KState* st, double* gm, GmLayout L, double* part) {
    if (gm_closed(st)) return;
    const T h = (T)L.H(gm)[j * L.m + j];
    double s = 0.0;
    GRID_LOOP(i, n) {
        const T wi = w[i] - h * vj[i];
        w[i] = wi;
        s += (double)wi * (double)wi;
    }
    double a[1] = {s};
    block_partials<1>(a, part);
}
// h_{j+1,j} = ||w||, Givens rotations, stopping test
struct FinGmNorm {
    KState* st;
    double* gm;
    GmLayout L;
    int j;
    __device__ void operator()(double (&t)[1]) const {
        if (gm_closed(st)) return;
        const int m = L.m;
        double* H = L.H(gm);
        double* cs = L.cs(gm);
        double* sn = L.sn(gm);
        double* g = L.g(gm);
        const double hn = sqrt(t[0]);
        st->hnorm = hn;
        H[(j + 1) * m + j] = hn;
        for (int i = 0; i < j; ++i) {
            const double t1 = cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
            const double t2 = -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
            H[i * m + j] = t1;
            H[(i + 1) * m + j] = t2;
        }
        const double den = hypot(H[j * m + j], H[(j + 1) * m + j]);
        if (den < kBT) { kfail(st, SE_GM_DEN); return; }
        cs[j] = H[j * m + j] / den;
        sn[j] = H[(j + 1) * m + j] / den;
        H[j * m + j] = cs[j] * H[j * m + j] + sn[j] * H[(j + 1) * m + j];
        H[(j + 1) * m + j] = 0.0;
        const double g1 = cs[j] * g[j] + sn[j] * g[j + 1];
        const double g2 = -sn[j] * g[j] + cs[j] * g[j + 1];
        g[j] = g1;
        g[j + 1] = g2;
        st->j_used = j + 1;
        st->it_end = st->it;
        st->it += 1;
        st->iters += 1;
        if (fabs(g[j + 1]) / st->normb < st->tol) { st->conv = 1; st->live = 0; }
        else if (st->it >= st->max_iter) st->live = 0;
    }
};

// V_{j+1} = w * (1 / h_{j+1,j}) when h_{j+1,j} > 1e-30. Skipped once the
// stopping test fired: V_{j+1} is then never read.
template <typename T>
__global__ void k_gm_normalize(const T* w, T* vnext, size_t n, KState* st) {
    if (gm_closed(st)) return;
    const double hn = st->hnorm;
    if (!(hn > kBT)) return;
    const T inv = (T)(1.0 / hn);
    GRID_LOOP(i, n) vnext[i] = w[i] * inv;
}

__global__ void k_gm_backsolve(KState* st, double* gm, GmLayout L) {
    DONE_RET(st);
    const int m = L.m, ju = st->j_used;
    double* H = L.H(gm);
    double* g = L.g(gm);
    double* y = L.y(gm);
    for (int i = ju - 1; i >= 0; --i) {
        double s = g[i];
        for (int k = i + 1; k < ju; ++k) s -= H[i * m + k] * y[k];
        if (fabs(H[i * m + i]) < kBT) { kfail(st, SE_GM_TRI); return; }
        y[i] = s / H[i * m + i];
    }
}

template <typename T>
__global__ void k_gm_combine(const T* V, size_t n, T* z, KState* st, const double* gm, GmLayout L) {
    DONE_RET(st);
    const int ju = st->j_used;
    const double* y = L.y(const_cast<double*>(gm));
    GRID_LOOP(t, n) {
        T acc = (T)0;
        for (int i = 0; i < ju; ++i) acc += (T)y[i] * V[(size_t)i * n + t];
        z[t] = acc;
    }
}

template <typename T>
__global__ void k_add(T* x, const T* dz, size_t n, KState* st) {
    DONE_RET(st);
    GRID_LOOP(i, n) x[i] += dz[i];
}

__global__ void k_gm_cycle_end(KState* st) {
    DONE_RET(st);
    if (st->conv) st->done = 1;
    else if (st->it >= st->max_iter) st->done = 3;
}

// ---------------------------------------------------------------------------
// Refinement kernels
// ---------------------------------------------------------------------------
// res = b - ax, partial ||res||^2
__global__ void k_resid_norm(const double* b, const double* ax, double* res, size_t n, double* part) {
    double s = 0.0;
    GRID_LOOP(i, n) {
        const double r = b[i] - ax[i];
        res[i] = r;
        s += r * r;
    }
    double w[1] = {s};
    block_partials<1>(w, part);
}
struct FinStore {
    double* out;
    __device__ void operator()(double (&t)[1]) const { out[0] = t[0]; }
};

__global__ void k_axpy_f32(double* x, const float* d, double a, size_t n) {
    GRID_LOOP(i, n) x[i] += a * (double)d[i];
}
// the same with d in permuted order: x[perm[t]] += a d[t]
__global__ void k_axpy_f32_perm(double* x, const float* d, const uint32_t* perm, double a, size_t n) {
    GRID_LOOP(t, n) x[perm[t]] += a * (double)d[t];
}
// y[t] = scale * x[perm[t]] in float32
__global__ void k_gather_f32(float* y, const double* x, const uint32_t* perm, double scale, size_t n) {
    GRID_LOOP(t, n) y[t] = (float)(scale * x[perm[t]]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T>
void KWork<T>::ensure(int solver, uint32_t n_, int m_) {
    if (n == n_ && m == m_ && x.get()) return;
    n = n_;
    m = m_;
    x.alloc(n); r.alloc(n); z.alloc(n); p.alloc(n); q.alloc(n); b.alloc(n);
    if (solver == FEM_SOLVER_BICGSTAB) { rt.alloc(n); y.alloc(n); s.alloc(n); t.alloc(n); ax.alloc(n); }
    if (solver == FEM_SOLVER_GMRES) {
        V.alloc((size_t)(m + 1) * n); w.alloc(n); wt.alloc(n); dz.alloc(n); ax.alloc(n);
        gm.alloc(GmLayout::size(m));
    }
}
template struct KWork<double>;
template struct KWork<float>;

void GraphCache::reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    tag = -1;
}

KrylovSolver::KrylovSolver(LinOp* A, Precond* M, const fem_solve_opts& o, hipStream_t s)
    : opts(o), A_(A), M_(M), s_(s) {
    if (A->rows() != A->cols()) fail(FEM_ERR_INVALID_ARGUMENT, "the solver needs a square matrix");
    if (M && M->size() != A->rows()) fail(FEM_ERR_INVALID_ARGUMENT, "preconditioner size differs from the matrix");
    if (o.solver == FEM_SOLVER_GMRES && o.restart < 1) fail(FEM_ERR_INVALID_ARGUMENT, "GMRES restart must be >= 1");
    if (o.solver < 0 || o.solver > FEM_SOLVER_MINICG) fail(FEM_ERR_INVALID_ARGUMENT, "unknown solver kind");
    red_.init(s);
    st_.alloc(1);
    hst_.ensure(1);
    auto mode = [&](bool fp32) {
        if (!M || M->kind() == FEM_PC_NONE) return (int)PCM_NONE;
        if (fp32 ? M->diag32() != nullptr : M->diag64() != nullptr) return (int)PCM_DIAG;
        return (int)PCM_GEN;
    };
    pc_mode64_ = mode(false);
    A32_ = A_;
    M32_ = M_;
    if (o.precision == FEM_PREC_FP32_IR) {
        if (M) M_->ensure_fp32(s);
        pc_mode32_ = mode(true);
        LinOp* op = nullptr;
        Precond* pc = nullptr;
        const uint32_t* rows = nullptr;
        if (M && M->permuted32(A, op, pc, rows)) {
            // the inner solves never apply A itself in float32
            A32_ = op;
            M32_ = pc;
            perm_ = rows;
        } else {
            A_->ensure_fp32(s);
        }
    }
    if (M) M_->prepare(o.precision != FEM_PREC_FP32_IR, o.precision == FEM_PREC_FP32_IR && perm_ == nullptr, s);
    dotp_.alloc(std::max<unsigned>({1u, A_->dot_parts(), M_ ? M_->dot_parts() : 0u, A32_->dot_parts(),
                                    M32_ ? M32_->dot_parts() : 0u}));
    // PCG with M = G^T G in three launches per iteration, where both factors
    // take the CSR-stream kernels and the folded tails stay small (each block
    // of the two kernels re-adds every partial sum of the kernel before).
    // FEM_PCG_FUSED=0 in the environment keeps the six-launch form (same
    // iterates; for comparison).
    const char* fused_env = std::getenv("FEM_PCG_FUSED");
    const bool fused_ok = !(fused_env && fused_env[0] == '0');
    if (fused_ok && M && o.solver == FEM_SOLVER_PCG && M->factors(fG_, fGt_) && fG_->lanes == 0 && fGt_->lanes == 0 &&
        fG_->n == A->rows() && fG_->m == A->rows() && fGt_->n == A->rows() && fGt_->m == A->rows()) {
        auto fits = [&](const SpmvPlan& pg, const SpmvPlan& pgt, unsigned npq) {
            return pg.nblk > 0 && pgt.nblk > 0 && (uint64_t)npq * pg.nblk <= kFacFold &&
                   2ull * pg.nblk * pgt.nblk <= kFacFold;
        };
        fac64_ = fits(fG_->plan, fGt_->plan, A_->dot_parts());
        if (o.precision == FEM_PREC_FP32_IR && fG_->v32.get() && fGt_->v32.get())
            fac32_ = fits(fG_->plan32, fGt_->plan32, A_->dot_parts());
        if (fac64_ || fac32_) facp_.alloc(2ull * std::max(fG_->plan.nblk, fG_->plan32.nblk));
    }
}

KrylovSolver::~KrylovSolver() {}

template <typename T>
void KrylovSolver::precond(const T* r, T* z, Gate g) {
    if constexpr (std::is_same<T, double>::value) M_->apply(r, z, s_, g);
    else M32_->apply32(r, z, s_, g);
}
template <typename T>
void KrylovSolver::matvec(const T* x, T* y, Gate g) {
    if constexpr (std::is_same<T, double>::value) A_->apply(x, y, s_, g);
    else A32_->apply32(x, y, s_, g);
}
template <typename T>
unsigned KrylovSolver::matvec_dot(const T* x, T* y, Gate g) {
    if constexpr (std::is_same<T, double>::value) return A_->apply_dot(x, y, dotp_.get(), s_, g);
    else return A32_->apply32_dot(x, y, dotp_.get(), s_, g);
}
template <typename T>
unsigned KrylovSolver::precond_dot(const T* r, T* z, Gate g) {
    if constexpr (std::is_same<T, double>::value) return M_->apply_dot(r, z, dotp_.get(), s_, g);
    else return M32_->apply32_dot(r, z, dotp_.get(), s_, g);
}

#define FIN(nv, ...)                                          \
    do {                                                      \
        launch_finalize<nv>(part, G, __VA_ARGS__, s_);        \
        FEM_HIP(hipGetLastError());                           \
    } while (0)

#define LAUNCH(kern, grid, ...)                                                  \
    do {                                                                         \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, s_, __VA_ARGS__);  \
        FEM_HIP(hipGetLastError());                                              \
    } while (0)

// a reduction tail is folded into the next update kernel when its G blocks
// re-adding np partial sums each read at most 4M values in all (above that
// the re-reads cost more than the one-block tail launch: at 128^3, np = 3560
// and G = 2048, the folded x/r update took 17.2 us against 11.2 + 4.9 us)
static bool fold_tail(unsigned G, unsigned np) { return np > 0 && (uint64_t)np * G <= (4ull << 20); }

template <typename T>
void KrylovSolver::pcg_init(KWork<T>& W) {
    const size_t n = W.n;
    const unsigned G = red_blocks(n);
    KState* st = st_.get();
    double* part = red_.partials.get();
    const int pm = std::is_same<T, double>::value ? pc_mode64_ : pc_mode32_;
    const T* d = nullptr;
    if (pm == PCM_DIAG) {
        if constexpr (std::is_same<T, double>::value) d = M_->diag64();
        else d = M_->diag32();
    }
    LAUNCH((k_pcg_init<T>), G, W.b.get(), W.x.get(), W.r.get(), n, st, part);
    FIN(1, FinPcgInit{st});
    if (pm == PCM_GEN) {
        precond<T>(W.r.get(), W.z.get(), Gate{&st->done});
        LAUNCH((k_pcg_init_rz<T, PCM_GEN>), G, W.r.get(), W.z.get(), d, W.p.get(), n, st, part);
    } else if (pm == PCM_DIAG) {
        LAUNCH((k_pcg_init_rz<T, PCM_DIAG>), G, W.r.get(), W.z.get(), d, W.p.get(), n, st, part);
    } else {
        LAUNCH((k_pcg_init_rz<T, PCM_NONE>), G, W.r.get(), W.z.get(), d, W.p.get(), n, st, part);
    }
    FIN(1, FinPcgInitRz{st});
}

template <typename T>
void KrylovSolver::pcg_iter(KWork<T>& W) {
    const size_t n = W.n;
    const unsigned G = red_blocks(n);
    KState* st = st_.get();
    double* part = red_.partials.get();
    const int pm = std::is_same<T, double>::value ? pc_mode64_ : pc_mode32_;
    const T* d = nullptr;
    if (pm == PCM_DIAG) {
        if constexpr (std::is_same<T, double>::value) d = M_->diag64();
        else d = M_->diag32();
    }
    Gate g{&st->done};
    // q = A p and p.q: one fused kernel where the operator has one; its
    // reduction tail then runs inside the x/r update when fold_tail allows
    // (every block re-adds the partial sums), else as its own launch
    const unsigned np = matvec_dot<T>(W.p.get(), W.q.get(), g);
    if (np && (std::is_same<T, double>::value ? fac64_ : fac32_)) {
        pcg_iter_fac<T>(W, np);
        return;
    }
    if (fold_tail(G, np)) {
        if (pm == PCM_GEN) {
            LAUNCH((k_pcg_xr_f<T, PCM_GEN>), G, W.x.get(), W.r.get(), W.p.get(), W.q.get(), d, n, st, part,
                   dotp_.get(), np);
            launch_finalize<2>(part, G, FinPcgXr<PCM_GEN>{st}, s_);
            FEM_HIP(hipGetLastError());
            const unsigned nz = precond_dot<T>(W.r.get(), W.z.get(), g);
            if (fold_tail(G, nz)) {
                LAUNCH((k_pcg_p_f<T, PCM_GEN>), G, W.r.get(), W.z.get(), d, W.p.get(), n, st, dotp_.get(), nz);
            } else if (nz) {
                launch_finalize<1>(dotp_.get(), nz, FinPcgRz{st}, s_);
                FEM_HIP(hipGetLastError());
                LAUNCH((k_pcg_p<T, PCM_GEN>), G, W.r.get(), W.z.get(), d, W.p.get(), n, st);
            } else {
                precond<T>(W.r.get(), W.z.get(), g);
                LAUNCH((k_dot2<T>), G, W.r.get(), W.z.get(), n, st, part);
                FIN(1, FinPcgRz{st});
                LAUNCH((k_pcg_p<T, PCM_GEN>), G, W.r.get(), W.z.get(), d, W.p.get(), n, st);
            }
            return;
        }
        if (pm == PCM_DIAG) {
            LAUNCH((k_pcg_xr_f<T, PCM_DIAG>), G, W.x.get(), W.r.get(), W.p.get(), W.q.get(), d, n, st, part,
                   dotp_.get(), np);
            launch_finalize<2>(part, G, FinPcgXr<PCM_DIAG>{st}, s_);
            FEM_HIP(hipGetLastError());
            LAUNCH((k_pcg_p<T, PCM_DIAG>), G, W.r.get(), W.z.get(), d, W.p.get(), n, st);
        } else {
            LAUNCH((k_pcg_xr_f<T, PCM_NONE>), G, W.x.get(), W.r.get(), W.p.get(), W.q.get(), d, n, st, part,
                   dotp_.get(), np);
            launch_finalize<2>(part, G, FinPcgXr<PCM_NONE>{st}, s_);
            FEM_HIP(hipGetLastError());
            LAUNCH((k_pcg_p<T, PCM_NONE>), G, W.r.get(), W.z.get(), d, W.p.get(), n, st);
        }
        return;
    } else if (np) {
        launch_finalize<1>(dotp_.get(), np, FinPcgPq{st}, s_);
        FEM_HIP(hipGetLastError());
    } else {
        matvec<T>(W.p.get(), W.q.get(), g);
        LAUNCH((k_dot2<T>), G, W.p.get(), W.q.get(), n, st, part);
        FIN(1, FinPcgPq{st});
    }
    if (pm == PCM_GEN) {
        LAUNCH((k_pcg_xr<T, PCM_GEN>), G, W.x.get(), W.r.get(), W.p.get(), W.q.get(), d, n, st, part);
        FIN(2, FinPcgXr<PCM_GEN>{st});
        if (const unsigned nz = precond_dot<T>(W.r.get(), W.z.get(), g)) {
            launch_finalize<1>(dotp_.get(), nz, FinPcgRz{st}, s_);
            FEM_HIP(hipGetLastError());
        } else {
            precond<T>(W.r.get(), W.z.get(), g);
            LAUNCH((k_dot2<T>), G, W.r.get(), W.z.get(), n, st, part);
            FIN(1, FinPcgRz{st});
        }
        LAUNCH((k_pcg_p<T, PCM_GEN>), G, W.r.get(), W.z.get(), d, W.p.get(), n, st);
    } else if (pm == PCM_DIAG) {
        LAUNCH((k_pcg_xr<T, PCM_DIAG>), G, W.x.get(), W.r.get(), W.p.get(), W.q.get(), d, n, st, part);
        FIN(2, FinPcgXr<PCM_DIAG>{st});
        LAUNCH((k_pcg_p<T, PCM_DIAG>), G, W.r.get(), W.z.get(), d, W.p.get(), n, st);
    } else {
        LAUNCH((k_pcg_xr<T, PCM_NONE>), G, W.x.get(), W.r.get(), W.p.get(), W.q.get(), d, n, st, part);
        FIN(2, FinPcgXr<PCM_NONE>{st});
        LAUNCH((k_pcg_p<T, PCM_NONE>), G, W.r.get(), W.z.get(), d, W.p.get(), n, st);
    }
}

// the rest of a PCG iteration with M = G^T G after q = A p (whose npq
// partial sums of p.q are in dotp_): t = G (r - alpha q) into z, then the
// x/r/p update with p = G^T t + beta p
template <typename T>
void KrylovSolver::pcg_iter_fac(KWork<T>& W, unsigned npq) {
    constexpr bool D = std::is_same<T, double>::value;
    const CsrMat& Gm = *fG_;
    const CsrMat& Gt = *fGt_;
    const SpmvPlan& pg = D ? Gm.plan : Gm.plan32;
    const SpmvPlan& pgt = D ? Gt.plan : Gt.plan32;
    const T* gv;
    const T* gtv;
    if constexpr (D) {
        gv = Gm.v.get();
        gtv = Gt.v.get();
    } else {
        gv = Gm.v32.get();
        gtv = Gt.v32.get();
    }
    KState* st = st_.get();
    if (Gm.cd.get())
        LAUNCH((k_pcg_fac_t<T, true>), pg.nblk, Gm.rp.get(), Gm.ci.get(), Gm.cd.get(), Gm.cbase.get(), gv,
               pg.bnd.get(), W.r.get(), W.q.get(), W.z.get(), st, dotp_.get(), npq, facp_.get());
    else
        LAUNCH((k_pcg_fac_t<T, false>), pg.nblk, Gm.rp.get(), Gm.ci.get(), Gm.cd.get(), Gm.cbase.get(), gv,
               pg.bnd.get(), W.r.get(), W.q.get(), W.z.get(), st, dotp_.get(), npq, facp_.get());
    if (Gt.cd.get())
        LAUNCH((k_pcg_fac_p<T, true>), pgt.nblk, Gt.rp.get(), Gt.ci.get(), Gt.cd.get(), Gt.cbase.get(), gtv,
               pgt.bnd.get(), W.z.get(), W.x.get(), W.r.get(), W.p.get(), W.q.get(), st, facp_.get(), pg.nblk);
    else
        LAUNCH((k_pcg_fac_p<T, false>), pgt.nblk, Gt.rp.get(), Gt.ci.get(), Gt.cd.get(), Gt.cbase.get(), gtv,
               pgt.bnd.get(), W.z.get(), W.x.get(), W.r.get(), W.p.get(), W.q.get(), st, facp_.get(), pg.nblk);
}

template <typename T>
void KrylovSolver::bi_init(KWork<T>& W) {
    const size_t n = W.n;
    const unsigned G = red_blocks(n);
    double* part = red_.partials.get();
    KState* st = st_.get();
    LAUNCH((k_bi_init<T>), G, W.b.get(), W.x.get(), W.r.get(), W.rt.get(), W.p.get(), W.q.get(), n, st, part);
    FIN(1, FinBiInit{st});
}

template <typename T>
void KrylovSolver::bi_iter(KWork<T>& W) {
    const size_t n = W.n;
    const unsigned G = red_blocks(n);
    KState* st = st_.get();
    double* part = red_.partials.get();
    Gate g{&st->done};
    T* v = W.q.get();
    // with the identity preconditioner y = p and z = s (aliases, no copies)
    const bool has_pc = M_ && M_->kind() != FEM_PC_NONE;
    T* y = has_pc ? W.y.get() : W.p.get();
    T* z = has_pc ? W.z.get() : W.s.get();
    LAUNCH((k_dot2<T>), G, W.rt.get(), W.r.get(), n, st, part);
    FIN(1, FinBiRho{st});
    LAUNCH((k_bi_p<T>), G, W.r.get(), v, W.p.get(), n, st);
    if (has_pc) precond<T>(W.p.get(), y, g);
    matvec<T>(y, v, g);
    LAUNCH((k_dot2<T>), G, W.rt.get(), v, n, st, part);
    FIN(1, FinBiRv{st});
    LAUNCH((k_bi_s<T>), G, W.r.get(), v, W.s.get(), n, st);
    if (has_pc) precond<T>(W.s.get(), z, g);
    matvec<T>(z, W.t.get(), g);
    LAUNCH((k_bi_tt<T>), G, W.t.get(), W.s.get(), n, st, part);
    FIN(2, FinBiTt{st});
    LAUNCH((k_bi_xr<T>), G, W.x.get(), W.r.get(), y, z, W.s.get(), W.t.get(), n, st, part);
    FIN(1, FinBiXr{st});
    matvec<T>(W.x.get(), W.ax.get(), Gate{&st->done, &st->check});
    LAUNCH((k_bi_true<T>), G, W.b.get(), W.ax.get(), n, st, part);
    FIN(1, FinBiTrue{st});
    LAUNCH((k_bi_replace<T>), G, W.b.get(), W.ax.get(), W.rt.get(), W.r.get(), n, st, part);
    FIN(1, FinBiReplace{st});
    LAUNCH(k_bi_tail, 1, st);
}

template <typename T>
void KrylovSolver::gm_cycle(KWork<T>& W) {
    const size_t n = W.n;
    const unsigned G = red_blocks(n);
    const int m = W.m;
    KState* st = st_.get();
    double* part = red_.partials.get();
    GmLayout L{m};
    double* gm = W.gm.get();
    T* V = W.V.get();
    auto Vi = [&](int i) { return V + (size_t)i * n; };
    LAUNCH(k_gm_cycle_begin, 1, st);
    matvec<T>(W.x.get(), W.ax.get(), Gate{&st->done, &st->need_true});
    LAUNCH((k_gm_v0<T>), G, W.b.get(), W.ax.get(), Vi(0), n, st, part);
    FIN(1, (FinGmV0{st, gm, L}));
    LAUNCH((k_gm_scale<T>), G, Vi(0), n, st);
    const bool has_pc = M_ && M_->kind() != FEM_PC_NONE;
    for (int j = 0; j < m; ++j) {
        Gate g{&st->done, &st->live};
        const T* src = Vi(j);
        if (has_pc) {
            precond<T>(Vi(j), W.wt.get(), g);
            src = W.wt.get();
        }
        matvec<T>(src, W.w.get(), g);
        LAUNCH((k_gm_dot0<T>), G, W.w.get(), Vi(0), n, st, part);
        FIN(1, (FinGmH{st, gm, L, 0, j}));
        for (int i = 1; i <= j; ++i) {
            LAUNCH((k_gm_axpy_dot<T>), G, W.w.get(), Vi(i - 1), Vi(i), n, i, j, st, gm, L, part);
            FIN(1, (FinGmH{st, gm, L, i, j}));
        }
        LAUNCH((k_gm_axpy_norm<T>), G, W.w.get(), Vi(j), n, j, st, gm, L, part);
        FIN(1, (FinGmNorm{st, gm, L, j}));
        LAUNCH((k_gm_normalize<T>), G, W.w.get(), Vi(j + 1), n, st);
    }
    LAUNCH(k_gm_backsolve, 1, st, gm, L);
    LAUNCH((k_gm_combine<T>), G, V, n, W.z.get(), st, gm, L);
    if (has_pc) {
        precond<T>(W.z.get(), W.dz.get(), Gate{&st->done});
        LAUNCH((k_add<T>), G, W.x.get(), W.dz.get(), n, st);
    } else {
        LAUNCH((k_add<T>), G, W.x.get(), W.z.get(), n, st);
    }
    LAUNCH(k_gm_cycle_end, 1, st);
}

void KrylovSolver::replay(GraphCache& gc, int tag, const std::function<void()>& body, bool launch) {
    if (!opts.use_graph || (M_ && !M_->graph_safe())) {
        if (launch) body();
        return;
    }
    if (gc.tag != tag) {
        gc.reset();
        FEM_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            hipGraph_t tmp = nullptr;
            (void)hipStreamEndCapture(s_, &tmp);
            if (tmp) (void)hipGraphDestroy(tmp);
            throw;
        }
        FEM_HIP(hipStreamEndCapture(s_, &gc.graph));
        FEM_HIP(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0));
        gc.tag = tag;
    }
    if (launch) FEM_HIP(hipGraphLaunch(gc.exec, s_));
}

void KrylovSolver::read_state() {
    FEM_HIP(hipMemcpyAsync(hst_.p, st_.get(), sizeof(KState), hipMemcpyDeviceToHost, s_));
    FEM_HIP(hipStreamSynchronize(s_));
}

template <typename T>
int KrylovSolver::run(KWork<T>& W, double tol, int max_iter, bool soft, long long* spmv) {
    KState h{};
    h.tol = tol;
    h.max_iter = max_iter;
    h.soft = soft ? 1 : 0;
    h.minicg = opts.solver == FEM_SOLVER_MINICG ? 1 : 0;
    *hst_.p = h;
    FEM_HIP(hipMemcpyAsync(st_.get(), hst_.p, sizeof(KState), hipMemcpyHostToDevice, s_));
    const int solver = opts.solver == FEM_SOLVER_MINICG ? FEM_SOLVER_PCG : opts.solver;
    if (solver == FEM_SOLVER_PCG) pcg_init<T>(W);
    else if (solver == FEM_SOLVER_BICGSTAB) bi_init<T>(W);
    else {
        const unsigned G = red_blocks(W.n);
        double* part = red_.partials.get();
        LAUNCH((k_gm_init<T>), G, W.b.get(), W.x.get(), (size_t)W.n, st_.get(), part);
        FIN(1, FinGmInit{st_.get()});
    }
    // PCG starts from x = 0, so ||r0|| = ||b||: its first chunk is queued
    // behind the init kernels without a host read (a solve that init marks
    // done replays one chunk of gated-off kernels)
    if (solver != FEM_SOLVER_PCG || max_iter <= 0) read_state();
    if (max_iter <= 0 && solver != FEM_SOLVER_GMRES) {
        // the reference's loop body never runs: x stays 0
        return 0;
    }
    // PCG / BiCGSTAB replay a captured chunk of iterations between host
    // checks. Iterations after convergence still launch (their kernels exit
    // on the done flag), so the chunk length follows the observed
    // convergence rate: power-of-two chunks, at most `cmax`, sized to the
    // iterations the residual history predicts are left. GMRES replays one
    // restart cycle at a time.
    const int cmax = opts.check_every > 0 ? opts.check_every : (W.n < 200000 ? 32 : 16);
    GraphSet& gs = std::is_same<T, double>::value ? g64_ : g32_;
    const double rel0 = solver == FEM_SOLVER_PCG ? 1.0 : hst_.p->normb > 0 ? hst_.p->normr / hst_.p->normb : 1.0;
    const int phase = phase_++;
    const int hint = solver == FEM_SOLVER_PCG && phase < (int)hint_.size() ? hint_[phase] : 0;
    int chunk = opts.check_every > 0 ? cmax : std::min(cmax, hint > 0 ? hint : 8);
    // the length of the graph being captured or launched
    int glen = chunk;
    auto body = [&] {
        if (solver == FEM_SOLVER_PCG) for (int k = 0; k < glen; ++k) pcg_iter<T>(W);
        else if (solver == FEM_SOLVER_BICGSTAB) for (int k = 0; k < glen; ++k) bi_iter<T>(W);
        else gm_cycle<T>(W);
    };
    // Without a fixed check interval (cmax a power of two) a chunk of any
    // length runs as back-to-back launches of power-of-two graphs (13 = 8 +
    // 4 + 1) before one host read, so at most log2(cmax) + 1 graphs exist
    // and a solve of a new right-hand side, whose chunk lengths differ from
    // the previous solve's, captures none once they do.
    const bool pow2 = opts.check_every <= 0 && solver != FEM_SOLVER_GMRES;
    auto launch_chunk = [&] {
        if (solver == FEM_SOLVER_GMRES) {
            const int tag = solver * 100000 + W.m;
            replay(gs.at(tag), tag, body);
            return;
        }
        for (int p = pow2 ? cmax : chunk; p >= 1; p >>= 1) {
            if (pow2 && !(chunk & p)) continue;
            glen = p;
            const int tag = solver * 100000 + p;
            replay(gs.at(tag), tag, body);
            if (!pow2) break;
        }
    };
    // every replay runs >= 1 iteration (GMRES: one cycle of >= 1 step)
    const long long guard_max = (long long)max_iter + 4;
    long long rounds = 0;
    static const bool trace = std::getenv("FEM_TRACE_SOLVE") != nullptr;
    while (!hst_.p->done) {
        launch_chunk();
        read_state();
        if (trace)
            std::fprintf(stderr, "[solve] replay of %d: iterations %d, |r|/|b| %.3e, done %d\n",
                         solver == FEM_SOLVER_GMRES ? W.m : chunk, hst_.p->it,
                         hst_.p->normb > 0 ? hst_.p->normr / hst_.p->normb : 0.0, hst_.p->done);
        if (++rounds > guard_max) fail(FEM_ERR_RUNTIME, "solver made no progress");
        if (solver == FEM_SOLVER_GMRES || opts.check_every > 0 || hst_.p->done) continue;
        const double rel = hst_.p->normb > 0 ? hst_.p->normr / hst_.p->normb : 0.0;
        const int it = hst_.p->it;
        int next = cmax;
        if (it > 0 && rel > 0 && rel < rel0) {
            const double rate = std::pow(rel / rel0, 1.0 / it);
            const double left = std::log(tol / rel) / std::log(rate);
            if (left > 0 && left < cmax) {
                next = 1;
                while (next < cmax && next < 1.1 * left + 1) next *= 2;
            }
        }
        // the previous solve's run needed `hint` iterations: end a chunk there
        const int left_hint = hint - hst_.p->iters;
        if (left_hint > 0) next = std::min(cmax, left_hint);
        chunk = next;
    }
    read_state();
    if (solver == FEM_SOLVER_PCG && hst_.p->done == 1 && phase < 64) {
        if ((int)hint_.size() <= phase) hint_.resize(phase + 1, 0);
        hint_[phase] = hst_.p->iters;
        // capture now the graphs a repeated solve launches (the power-of-two
        // parts of the cmax-long chunks and of the remainder), so that solve
        // captures nothing
        if (pow2)
            for (const int c : {std::min(cmax, hint_[phase]), hint_[phase] % cmax})
                for (int p = 1; p <= c; p <<= 1) {
                    if (!(c & p)) continue;
                    glen = p;
                    const int tag = solver * 100000 + p;
                    replay(gs.at(tag), tag, body, false);
                }
    }
    if (hst_.p->done == 2) {
        char ctx[128];
        std::snprintf(ctx, sizeof ctx, " (iteration %d, ||r||/||b|| = %.3e)", hst_.p->it,
                      hst_.p->normb > 0 ? hst_.p->normr / hst_.p->normb : 0.0);
        fail(FEM_ERR_RUNTIME, std::string(solver_err_msg(hst_.p->err)) + ctx);
    }
    if (spmv) *spmv += hst_.p->iters;
    return hst_.p->iters;
}

double KrylovSolver::true_residual(const double* b, const double* x, double* work) {
    const size_t n = A_->rows();
    A_->apply(x, work, s_);
    if (!tr_out_.get()) tr_out_.alloc(1);
    {
        const unsigned G = red_blocks(n);
        double* part = red_.partials.get();
        LAUNCH(k_resid_norm, G, b, work, work, n, part);
        FIN(1, FinStore{tr_out_.get()});
    }
    double h = 0.0;
    tr_out_.download(&h, 1, s_);
    FEM_HIP(hipStreamSynchronize(s_));
    return std::sqrt(h);
}

// A preconditioner with an iteration budget (the learned SPAI) gets one
// attempt of at most that many iterations; if it does not converge (or breaks
// down) the system is solved again from x0 = 0 with its fallback
// preconditioner, and stats->fallback is set. The fallback's iteration count
// is handed back to the preconditioner, which can make it the budget of later
// solves (a learned solve that needs more iterations than the fallback did is
// a regression).
void KrylovSolver::solve(const double* b, double* x, fem_solve_stats* stats) {
    const int budget = M_ ? M_->budget() : 0;
    if (budget <= 0) {
        solve_once(b, x, stats);
        return;
    }
    Timer tm;
    const int saved = opts.max_iter;
    opts.max_iter = std::min(saved, budget);
    fem_solve_stats S{};
    bool ok = false;
    try {
        solve_once(b, x, &S);
        ok = S.converged == 1;
    } catch (const Error& e) {
        opts.max_iter = saved;
        if (e.code != FEM_ERR_RUNTIME) throw;
    }
    opts.max_iter = saved;
    if (!ok) {
        const long long first = S.spmv_count;
        if (!fb_) fb_.reset(new KrylovSolver(A_, M_->fallback(s_), opts, s_));
        fb_->solve(b, x, &S);
        if (S.converged) M_->fallback_ran(S.inner_iterations);
        S.fallback = 1;
        S.spmv_count += first;
    }
    S.solve_ms = tm.ms();
    if (stats) *stats = S;
}

void KrylovSolver::solve_once(const double* b, double* x, fem_solve_stats* stats) {
    phase_ = 0;
    const uint32_t n = A_->rows();
    const int m = opts.solver == FEM_SOLVER_GMRES ? opts.restart : 0;
    fem_solve_stats S{};
    Timer tm;
    if (opts.precision == FEM_PREC_FP64) {
        w64_.ensure(opts.solver, n, m);
        FEM_HIP(hipMemcpyAsync(w64_.b.get(), b, (size_t)n * 8, hipMemcpyDeviceToDevice, s_));
        S.inner_iterations = run<double>(w64_, opts.tol, opts.max_iter, false, &S.spmv_count);
        // mini_cg returns the raw loop index (test_BlockMatrix_mpi.cpp:157-182)
        S.iterations = opts.solver == FEM_SOLVER_MINICG ? hst_.p->it : hst_.p->it_end;
        S.converged = hst_.p->done == 1 ? 1 : 0;
        FEM_HIP(hipMemcpyAsync(x, w64_.x.get(), (size_t)n * 8, hipMemcpyDeviceToDevice, s_));
        FEM_HIP(hipStreamSynchronize(s_));
        S.solve_ms = tm.ms();
    } else {
        // float32 Krylov inside float64 iterative refinement
        w32_.ensure(opts.solver, n, m);
        if (x64_.n != n) { x64_.alloc(n); ax64_.alloc(n); res64_.alloc(n); }
        if (!nrm_out_.get()) nrm_out_.alloc(1);
        DevBuf<double>& nrm = nrm_out_;
        vec_fill<double>(x64_.get(), 0.0, n, s_);
        FEM_HIP(hipMemcpyAsync(res64_.get(), b, (size_t)n * 8, hipMemcpyDeviceToDevice, s_));
        // ||b||
        vec_fill<double>(ax64_.get(), 0.0, n, s_);
        {
            const unsigned G = red_blocks(n);
            double* part = red_.partials.get();
            LAUNCH(k_resid_norm, G, b, ax64_.get(), res64_.get(), (size_t)n, part);
            FIN(1, FinStore{nrm.get()});
        }
        double h = 0.0;
        nrm.download(&h, 1, s_);
        FEM_HIP(hipStreamSynchronize(s_));
        const double normb = std::sqrt(h);
        double normr = normb;
        int total = 0;
        int refine = 0;
        const int max_ref = opts.max_refine > 0 ? opts.max_refine : 50;
        bool conv = false;
        if (normb < 1e-300) conv = true;
        while (!conv && refine < max_ref && total < opts.max_iter) {
            if (normr / normb < opts.tol) { conv = true; break; }
            if (perm_) LAUNCH(k_gather_f32, red_blocks(n), w32_.b.get(), res64_.get(), perm_, 1.0 / normr, (size_t)n);
            else vec_convert<float, double>(w32_.b.get(), res64_.get(), 1.0 / normr, n, s_);
            const double want = std::max(opts.inner_tol, 0.5 * opts.tol * normb / normr);
            total += run<float>(w32_, want, opts.max_iter - total, true, &S.spmv_count);
            // A breakdown of the float32 inner solve (p.Ap <= 0, r.z <= 0, ...)
            // ends it softly only once it has reduced its scaled residual
            // below kSoftExit: there it is float32 round-off near convergence,
            // and the refinement step continues from the iterate reached.
            // Earlier it means A or M is not what the method needs (not SPD
            // for PCG), and the solve fails with the error the float64 solve
            // raises.
            if (hst_.p->err < 0) {
                const double inner_rel = hst_.p->normb > 0 ? hst_.p->normr / hst_.p->normb : 1.0;
                if (!(inner_rel < kSoftExit)) {
                    char ctx[160];
                    std::snprintf(ctx, sizeof ctx, " (float32 inner solve, refinement step %d, iteration %d)",
                                  refine + 1, hst_.p->it);
                    fail(FEM_ERR_RUNTIME, std::string(solver_err_msg(-hst_.p->err)) + ctx);
                }
            }
            if (perm_) LAUNCH(k_axpy_f32_perm, red_blocks(n), x64_.get(), w32_.x.get(), perm_, normr, (size_t)n);
            else LAUNCH(k_axpy_f32, red_blocks(n), x64_.get(), w32_.x.get(), normr, (size_t)n);
            A_->apply(x64_.get(), ax64_.get(), s_);
            S.spmv_count += 1;
            {
                const unsigned G = red_blocks(n);
                double* part = red_.partials.get();
                LAUNCH(k_resid_norm, G, b, ax64_.get(), res64_.get(), (size_t)n, part);
                FIN(1, FinStore{nrm.get()});
            }
            nrm.download(&h, 1, s_);
            FEM_HIP(hipStreamSynchronize(s_));
            const double nr = std::sqrt(h);
            ++refine;
            if (std::getenv("FEMGPU_DEBUG"))
                std::fprintf(stderr, "[femgpu] refine %d: inner tol %.2e iters %d -> ||r||/||b|| %.3e (inner done %d err %d)\n",
                             refine, want, total, nr / normb, hst_.p->done, hst_.p->err);
            if (!(nr < normr)) {
                // the step made the residual larger (or NaN): take it back
                // and keep the last good iterate and its residual norm
                if (perm_)
                    LAUNCH(k_axpy_f32_perm, red_blocks(n), x64_.get(), w32_.x.get(), perm_, -normr, (size_t)n);
                else LAUNCH(k_axpy_f32, red_blocks(n), x64_.get(), w32_.x.get(), -normr, (size_t)n);
                break;
            }
            if (!(nr < normr * 0.999)) { normr = nr; break; }  // stagnation in float32
            normr = nr;
        }
        if (normr / normb < opts.tol) conv = true;
        S.inner_iterations = total;
        S.iterations = total;
        S.refinements = refine;
        S.converged = conv ? 1 : 0;
        FEM_HIP(hipMemcpyAsync(x, x64_.get(), (size_t)n * 8, hipMemcpyDeviceToDevice, s_));
        FEM_HIP(hipStreamSynchronize(s_));
        S.solve_ms = tm.ms();
    }
    if (stats) {
        if (st_work_.n != n) {
            st_work_.alloc(n);
            st_zero_.alloc(n);
            st_out_.alloc(1);
            vec_fill<double>(st_zero_.get(), 0.0, n, s_);
        }
        double* work = st_work_.get();
        const double normb = [&] {
            const unsigned G = red_blocks(n);
            double* part = red_.partials.get();
            LAUNCH(k_resid_norm, G, b, st_zero_.get(), work, (size_t)n, part);
            FIN(1, FinStore{st_out_.get()});
            double hh = 0.0;
            st_out_.download(&hh, 1, s_);
            FEM_HIP(hipStreamSynchronize(s_));
            return std::sqrt(hh);
        }();
        const double tr = true_residual(b, x, work);
        S.true_relres = normb > 0 ? tr / normb : 0.0;
        *stats = S;
    }
}

}  // namespace fem

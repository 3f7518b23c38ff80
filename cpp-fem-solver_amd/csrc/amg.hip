// amg.hip — algebraic multigrid preconditioners.
//   "amg" (classical C/F):      ode-solver/src/math/array.cpp:3165-3398
//   smoothed aggregation (SA):  ode-solver/include/math/blockMatrix.hpp:868-1179,
//                               applied to scalar CSR through a 1x1 condensation
//                               (ode-solver/test/test_BlockMatrix_schur3d.cpp:143-146)
//
// The hierarchy is built once (amg_setup.cpp); every level lives on the
// device as CSR (A_k, P_k, R_k = P_k^T) with SpMV plans. One V-cycle is a
// fixed sequence of launches (captured into the solver's hipGraph):
//   pre-smooth -> residual f - A u (fused SpMV epilogue) -> restrict R r
//   -> ... -> coarsest: u = C f with C the explicit coarse solution operator
//   (float64 in both precisions) -> prolong u += P u_c (fused) -> post-smooth.
// Smoothers:
//   C/F Jacobi   u <- u + w (f - A u) / d          (array.cpp:3355-3369)
//   SA  Jacobi   u <- u + w (d^{-1} (f - A u))      (blockMatrix.hpp:1010-1024)
//   SGS          multicolour Gauss-Seidel, colours ascending before the coarse
//                correction and descending after it (symmetric V-cycle), on
//                levels that colour into at most kMaxGsColours colours; other
//                levels are smoothed by damped Jacobi with weight
//                (4/3) / rho(D^{-1} A), rho the Gershgorin bound
// Every level starts from u = 0, as the reference's V-cycle does.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "amg.hpp"
#include "precond.hpp"
#include "reduce.hpp"

namespace fem {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// first Jacobi sweep from u = 0 (the SpMV of a zero vector is exactly 0)
template <typename T, bool INV>
__global__ void k_jacobi0(const T* __restrict__ f, const T* __restrict__ d, T* __restrict__ u, uint32_t n, T w,
                          Gate g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T fi = f[i], di = d[i];  // read while the gate's flag is in flight
    if (g.closed()) return;
    u[i] = jac0<T, INV>(fi, di, w);
}

// Multicolour Gauss-Seidel on a colour-grouped copy of the level matrix:
// row t of the grouped CSR (gp, gci, gv) is row rows[t] of A, the rows of
// each colour are contiguous, gd[t] = a_ii. In the one-workgroup kernel for
// small levels LANES lanes share one row (the row's products summed by a
// butterfly over those lanes).
//   u_i += w (f_i - (A u)_i) / a_ii   for the rows t in [lo, hi)
template <typename T, int LANES>
__device__ __forceinline__ void gs_rows(uint32_t t, uint32_t hi, int lane, const uint32_t* __restrict__ rows,
                                        const uint32_t* __restrict__ gp, const uint32_t* __restrict__ gci,
                                        const T* __restrict__ gv, const T* __restrict__ gd, const T* __restrict__ f,
                                        T* u, T w) {
    T s = (T)0;
    if (t < hi)
        for (uint32_t p = gp[t] + lane; p < gp[t + 1]; p += LANES) s += gv[p] * u[gci[p]];
    s = group_sum<LANES>(s);
    if (t < hi && lane == 0) {
        const uint32_t i = rows[t];
        u[i] = u[i] + (w * (f[i] - s)) / gd[t];
    }
}

// all colours of `sweeps` sweeps in one workgroup (levels of <= 8192 rows)
template <typename T, int LANES>
__global__ __launch_bounds__(1024) void k_gs_small(int ncol, const uint32_t* __restrict__ cptr,
                                                   const uint32_t* __restrict__ rows, const uint32_t* __restrict__ gp,
                                                   const uint32_t* __restrict__ gci, const T* __restrict__ gv,
                                                   const T* __restrict__ gd, const T* __restrict__ f, T* u, T w,
                                                   int sweeps, int reverse, Gate g) {
    if (g.closed()) return;
    constexpr uint32_t kRowsPerPass = 1024 / LANES;
    const int lane = threadIdx.x % LANES;
    for (int sw = 0; sw < sweeps; ++sw) {
        for (int cc = 0; cc < ncol; ++cc) {
            const int c = reverse ? ncol - 1 - cc : cc;
            const uint32_t lo = cptr[c], hi = cptr[c + 1];
            for (uint32_t base = lo; base < hi; base += kRowsPerPass)
                gs_rows<T, LANES>(base + threadIdx.x / LANES, hi, lane, rows, gp, gci, gv, gd, f, u, w);
            __syncthreads();
        }
    }
}

// One colour of a large level, blocked like the CSR-stream SpMV over the
// colour-grouped matrix: a block stages the products of its rows' entries
// in LDS (coalesced loads of gci/gv), then one thread per row sums them and
// applies the update to u[rows[t]]. blk[] holds row boundaries of the
// grouped matrix cut at colour boundaries; the launch covers one colour.
// Rows of one colour share no entries, so blocks never read a u value that
// another block of the same launch writes (the diagonal entry is read in
// the staging phase, before this block's own update).
// MODE 0 updates u; MODE 1 writes the residual res[i] = f_i - (A u)_i of the
// rows instead, and blocks from `zero_from` on write res = 0 without reading
// anything (the last colour right after a forward sweep with weight 1, whose
// residual is zero in exact arithmetic).
// DOT (MODE 0): the block also writes part[bid] = sum over its rows of
// f_i * u_i (float64, the new u_i): with f = r on the finest level and the
// last pass of each colour, the partial sums of r.z.
template <typename T, int MODE, bool PK, bool DOT = false>
__global__ __launch_bounds__(kBlock) void k_gs_stream(const uint2* __restrict__ blk,
                                                      const uint32_t* __restrict__ rows,
                                                      const uint32_t* __restrict__ gp,
                                                      const uint32_t* __restrict__ gci,
                                                      const uint16_t* __restrict__ gcd,
                                                      const uint32_t* __restrict__ gcbase, const T* __restrict__ gv,
                                                      const T* __restrict__ gd, const T* __restrict__ f, T* u, T w,
                                                      T* res, uint32_t zero_from, Gate g,
                                                      double* __restrict__ part) {
    constexpr int STAGE = stage_nnz<T>();
    __shared__ T prod[STAGE];
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
    // the block's bounds are read while the gate's flag is in flight (reads
    // only: the gate guards the stores)
    const uint2 lo = blk[bid], hi = blk[bid + 1];
    const uint32_t r0 = lo.x, r1 = hi.x, k0 = lo.y, k1 = hi.y;
    if (g.closed()) return;
    const uint32_t t = threadIdx.x;
    if (MODE == 1 && bid >= zero_from) {
        for (uint32_t row = r0 + t; row < r1; row += kBlock) res[rows[row]] = (T)0;
        return;
    }
    const uint32_t cnt = k1 - k0;
    if (stage_fits<STAGE>(k0, cnt)) {
        // a block holds at most stage_rows<T>() rows: RPT per thread, whose
        // bounds and operands are loaded before the staging
        constexpr int RPT = (stage_rows<T>() + kBlock - 1) / kBlock;
        uint32_t rb[RPT], re[RPT], ri[RPT];
        T fi[RPT], di[RPT], ui[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const uint32_t row = r0 + t + (uint32_t)j * kBlock;
            rb[j] = re[j] = ri[j] = 0;
            fi[j] = di[j] = ui[j] = (T)0;
            if (row < r1) {
                rb[j] = gp[row] - k0;
                re[j] = gp[row + 1] - k0;
                ri[j] = rows[row];
                fi[j] = f[ri[j]];
                if (MODE == 0) {
                    di[j] = gd[row];
                    ui[j] = u[ri[j]];
                }
            }
        }
        if (cnt > 0) stage_products<T, STAGE, PK>(gci, gcd, gcbase, gv, u, k0, cnt, prod);
        __syncthreads();
        double dsum[1] = {0.0};
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const uint32_t row = r0 + t + (uint32_t)j * kBlock;
            if (row < r1) {
                T s = (T)0;
                for (uint32_t k = rb[j]; k < re[j]; ++k) s += prod[k];
                if (MODE == 0) {
                    const T un = ui[j] + (w * (fi[j] - s)) / di[j];
                    u[ri[j]] = un;
                    if (DOT) dsum[0] += (double)fi[j] * (double)un;
                } else {
                    res[ri[j]] = fi[j] - s;
                }
            }
        }
        if constexpr (DOT) {
            block_sum<1>(dsum);
            if (t == 0) part[bid] = dsum[0];
        }
    } else {
        double s = 0.0;
        for (uint32_t k = k0 + t; k < k1; k += kBlock) s += (double)(gv[k] * u[col_of<PK>(gci, gcd, gcbase, k)]);
        double tot[1] = {s};
        block_sum<1>(tot);
        if (t == 0) {  // a block of one long row
            const uint32_t i = rows[r0];
            if (MODE == 0) {
                const T un = u[i] + (w * (f[i] - (T)tot[0])) / gd[r0];
                u[i] = un;
                if (DOT) part[bid] = (double)f[i] * (double)un;
            } else {
                res[i] = f[i] - (T)tot[0];
            }
        }
    }
}

// One pass over the colour-ordered form of a level (rows in colour order,
// the off-diagonal entries O with columns given as positions, the diagonal
// d): a block stages the products of its rows' entries in LDS like the
// CSR-stream SpMV, then one thread per row finishes the row.
//   MODE 0: u_t = (f_t - sum_k o_tk u_k) / d_t   Gauss-Seidel with weight 1
//           (u_t itself is not read; rows of one colour read only other
//           colours' entries, so blocks never read what this launch writes);
//           DOT: part[bid] = sum of f_t u_t over the block's rows (new u_t)
//   MODE 1: y_t = f_t - (sum_k o_tk u_k + d_t u_t)   residual
//   MODE 2: y_t = sum_k o_tk u_k + d_t u_t           product y = A u;
//           DOT: part[bid] = sum of u_t y_t
template <typename T, int MODE, bool PK, bool DOT>
__global__ __launch_bounds__(kBlock) void k_gsp(const uint2* __restrict__ blk, const uint32_t* __restrict__ rp,
                                                const uint32_t* __restrict__ ci, const uint16_t* __restrict__ cd,
                                                const uint32_t* __restrict__ cbase, const T* __restrict__ v,
                                                const T* __restrict__ d, const T* __restrict__ f, T* u,
                                                T* __restrict__ y, Gate g, double* __restrict__ part) {
    constexpr int STAGE = stage_nnz<T>();
    __shared__ T prod[STAGE];
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
    const uint2 lo = blk[bid], hi = blk[bid + 1];
    const uint32_t r0 = lo.x, r1 = hi.x, k0 = lo.y, k1 = hi.y;
    if (g.closed()) return;
    const uint32_t t = threadIdx.x;
    const uint32_t cnt = k1 - k0;
    if (stage_fits<STAGE>(k0, cnt)) {
        constexpr int RPT = (stage_rows<T>() + kBlock - 1) / kBlock;
        uint32_t rb[RPT], re[RPT];
        T fi[RPT], di[RPT], ui[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const uint32_t row = r0 + t + (uint32_t)j * kBlock;
            rb[j] = re[j] = 0;
            fi[j] = di[j] = ui[j] = (T)0;
            if (row < r1) {
                rb[j] = rp[row] - k0;
                re[j] = rp[row + 1] - k0;
                di[j] = d[row];
                if (MODE != 2) fi[j] = f[row];
                if (MODE != 0) ui[j] = u[row];
            }
        }
        if (cnt > 0) stage_products<T, STAGE, PK>(ci, cd, cbase, v, u, k0, cnt, prod);
        __syncthreads();
        double dsum[1] = {0.0};
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const uint32_t row = r0 + t + (uint32_t)j * kBlock;
            if (row < r1) {
                T sg = (T)0;
                for (uint32_t k = rb[j]; k < re[j]; ++k) sg += prod[k];
                if constexpr (MODE == 0) {
                    const T un = (fi[j] - sg) / di[j];
                    u[row] = un;
                    if (DOT) dsum[0] += (double)fi[j] * (double)un;
                } else if constexpr (MODE == 1) {
                    y[row] = fi[j] - (sg + di[j] * ui[j]);
                } else {
                    const T yv = sg + di[j] * ui[j];
                    y[row] = yv;
                    if (DOT) dsum[0] += (double)ui[j] * (double)yv;
                }
            }
        }
        if constexpr (DOT) {
            block_sum<1>(dsum);
            if (t == 0) part[bid] = dsum[0];
        }
    } else {
        double sd = 0.0;
        for (uint32_t k = k0 + t; k < k1; k += kBlock) sd += (double)(v[k] * u[col_of<PK>(ci, cd, cbase, k)]);
        double tot[1] = {sd};
        block_sum<1>(tot);
        if (t == 0) {  // a block of one long row
            const T sg = (T)tot[0];
            if constexpr (MODE == 0) {
                const T un = (f[r0] - sg) / d[r0];
                u[r0] = un;
                if (DOT) part[bid] = (double)f[r0] * (double)un;
            } else if constexpr (MODE == 1) {
                y[r0] = f[r0] - (sg + d[r0] * u[r0]);
            } else {
                const T yv = sg + d[r0] * u[r0];
                y[r0] = yv;
                if (DOT) part[bid] = (double)u[r0] * (double)yv;
            }
        }
    }
}

// The start of a forward sweep from u = 0 in the colour-ordered form: the
// first colour's rows (positions below c0_end) get u = f / d, the rows at
// positions from z_lo on are set to zero (the colours a pass reads before
// their own pass writes them); n is the end of the launch.
template <typename T>
__global__ void k_gsp_first(uint32_t c0_end, uint32_t z_lo, uint32_t n, const T* __restrict__ d,
                            const T* __restrict__ f, T* __restrict__ u, Gate g) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (g.closed()) return;
    if (t < c0_end) u[t] = f[t] / d[t];
    else if (t >= z_lo) u[t] = (T)0;
}

// The first colour of a sweep from u = 0 and the zero start for the other
// colours in one pass: grouped rows t < c0_end get u = w f / a_ii (the
// Gauss-Seidel update with every neighbour still zero), the rest u = 0.
template <typename T>
__global__ void k_gs_first(const uint32_t* __restrict__ rows, uint32_t n, uint32_t c0_end, const T* __restrict__ gd,
                           const T* __restrict__ f, T* __restrict__ u, T w, Gate g) {
    if (g.closed()) return;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = rows[t];
    u[i] = t < c0_end ? (w * f[i]) / gd[t] : (T)0;
}

// u = C f with C dense (row-major, float64); f, u of type T. One wavefront
// per row: coalesced reads of the row, lane-strided partial sums in float64,
// butterfly total.
template <typename T>
__global__ __launch_bounds__(256) void k_coarse(const double* __restrict__ C, uint32_t nc, const T* __restrict__ f,
                                                T* __restrict__ u, Gate g) {
    extern __shared__ double fs[];
    for (uint32_t j = threadIdx.x; j < nc; j += blockDim.x) fs[j] = (double)f[j];
    __syncthreads();
    if (g.closed()) return;  // uniform: tested after the staging, before any store to u
    const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= nc) return;
    const double* row = C + (size_t)i * nc;
    double s = 0.0;
    for (uint32_t j = lane; j < nc; j += 64) s += row[j] * fs[j];
    s = wave_sum(s);
    if (lane == 0) u[i] = (T)s;
}

// ---------------------------------------------------------------------------
// Fused tail of the V-cycle on the level above a dense coarsest level, when
// that level smooths by damped Jacobi with two pre-sweeps: per V-cycle two
// launches replace its residual, the restriction, the coarse solve, the
// prolongation and the first post-sweep. With S = 2w Dj^-1 - w^2 Dj^-1 A
// Dj^-1 the operator of the two pre-sweeps from zero (Dj^-1: division by
// the smoother's diagonal d, or multiplication by its inverse), the coarse
// right-hand side is R (f - A S f) = Mpre f, and with C the coarse solution
// operator the prolongation and the first post-sweep give
// u = S f + W fc,  out = u + w Dj^-1 (f - A S f - AW fc),
// W = P C and AW = A P C (set up in float64 from the float64 coarse
// operator; float vectors read float32 copies of Mpre, W and AW, and every
// product with fc is summed in float64).
// ---------------------------------------------------------------------------
template <typename T, bool INV>
__device__ __forceinline__ T jdiv(T x, T d) {
    if constexpr (INV) return d * x;
    else return x / d;
}
// w Dj^-1 r with the rounding of the Jacobi epilogues
template <typename T, bool INV>
__device__ __forceinline__ T jstep(T w, T r, T d) {
    if constexpr (INV) return w * (d * r);
    else return (w * r) / d;
}

// blocks [0, nc nq): the coarse right-hand side Mpre f in nq chunks of
// kTailChunk columns, fcp[q nc + c] the float64 partial of chunk q of coarse
// row c; the rest: the two pre-sweeps from zero, LANES lanes per row (the
// float arithmetic of EPI_JACOBI0 / EPI_JACOBI0_INV). Loads are issued in
// batches before their values are used.
constexpr int kTailBatch = 4;      // sparse-row loads per lane and batch
constexpr int kTailRowBatch = 8;   // Mpre row loads per thread
constexpr uint32_t kTailChunk = kTailRowBatch * kBlock;  // Mpre columns per block
constexpr int kTailWBatch = 16;    // W / AW loads per lane and batch
template <typename T, bool INV, int LANES>
__global__ __launch_bounds__(kBlock) void k_tail_pre(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                                     const T* __restrict__ v, const T* __restrict__ f,
                                                     const T* __restrict__ d, T w, uint32_t n, T* __restrict__ upre,
                                                     const T* __restrict__ mpre, uint32_t nc,
                                                     double* __restrict__ fcp, Gate g) {
    constexpr int U = kTailBatch;
    const uint32_t nq = (n + kTailChunk - 1) / kTailChunk;
    if (blockIdx.x < nc * nq) {
        // a chunk of kTailChunk columns of a row of Mpre (the longest work
        // of the launch: its blocks come first), one batch of loads
        constexpr int UC = kTailRowBatch;
        const uint32_t c = blockIdx.x / nq, q = blockIdx.x % nq;
        const T* row = mpre + (size_t)c * n;
        double t[1] = {0.0};
        const uint32_t j0 = q * kTailChunk + threadIdx.x;
        T mv[UC], fv[UC];
#pragma unroll
        for (int u = 0; u < UC; ++u) {
            const uint32_t j = j0 + u * kBlock;
            mv[u] = j < n ? row[j] : (T)0;
            fv[u] = j < n ? f[j] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < UC; ++u) t[0] += (double)mv[u] * (double)fv[u];
        block_sum<1>(t);
        if (g.closed()) return;  // uniform in the block, after its barriers
        if (threadIdx.x == 0) fcp[(size_t)q * nc + c] = t[0];
        return;
    }
    const uint32_t row = ((blockIdx.x - nc * nq) * (uint32_t)kBlock + threadIdx.x) / LANES;
    const int lane = threadIdx.x % LANES;
    uint32_t k0 = 0, k1 = 0;
    if (row < n) {
        k0 = rp[row];
        k1 = rp[row + 1];
    }
    if (g.closed()) return;
    T s = (T)0;
    for (uint32_t kb = k0 + lane; kb < k1; kb += U * LANES) {
        uint32_t cv[U];
        T av[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t k = kb + u * LANES;
            cv[u] = k < k1 ? ci[k] : 0u;
            av[u] = k < k1 ? v[k] : (T)0;
        }
        T fv[U], dv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            fv[u] = f[cv[u]];
            dv[u] = d[cv[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (kb + u * LANES < k1) s += av[u] * jac0<T, INV>(fv[u], dv[u], w);
    }
    s = group_sum<LANES>(s);
    if (row < n && lane == 0) upre[row] = jac0<T, INV>(f[row], d[row], w) + jstep<T, INV>(w, f[row] - s, d[row]);
}

// the prolongation u = upre + W fc and the first post-sweep in one pass:
// out = u + w Dj^-1 (f - A upre - AW fc); LANES lanes per row; W and AW in
// WT (float64, or float32 copies for float vectors), their products with fc
// summed in float64
template <typename T, typename WT, bool INV, int LANES>
__global__ __launch_bounds__(kBlock) void k_tail_post(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                                      const T* __restrict__ v, const T* __restrict__ f,
                                                      const T* __restrict__ d, T w, uint32_t n,
                                                      const T* __restrict__ upre, const WT* __restrict__ W,
                                                      const WT* __restrict__ AW, uint32_t nc,
                                                      const double* __restrict__ fcp, uint32_t nq,
                                                      T* __restrict__ out, Gate g) {
    constexpr int U = kTailBatch, UW = kTailWBatch;
    const uint32_t row = (blockIdx.x * (uint32_t)kBlock + threadIdx.x) / LANES;
    const int lane = threadIdx.x % LANES;
    const bool on = row < n;
    const uint32_t rr = on ? row : 0u;  // rows past the end read row 0 and store nothing
    const uint32_t k0 = rp[rr], k1 = on ? rp[rr + 1] : k0;
    const WT* wr = W + (size_t)rr * nc;
    const WT* ar = AW + (size_t)rr * nc;
    T s1 = (T)0;
    double s2 = 0.0, s3 = 0.0;
    // fc into LDS (its chunk partials added in chunk order), loads issued
    // before the row's own
    extern __shared__ double fs[];
    for (uint32_t c = threadIdx.x; c < nc; c += blockDim.x) {
        double x = fcp[c];
        for (uint32_t q = 1; q < nq; ++q) x += fcp[(size_t)q * nc + c];
        fs[c] = x;
    }
    WT wv[UW], avv[UW];
    // a batch of this row's W and AW entries
    auto load = [&](uint32_t cb) {
#pragma unroll
        for (int u = 0; u < UW; ++u) {
            const uint32_t c = cb + u * LANES;
            wv[u] = c < nc ? wr[c] : (WT)0;
            avv[u] = c < nc ? ar[c] : (WT)0;
        }
    };
    auto add = [&](uint32_t cb) {
#pragma unroll
        for (int u = 0; u < UW; ++u) {
            const uint32_t c = cb + u * LANES;
            const double x = c < nc ? fs[c] : 0.0;
            s2 += (double)wv[u] * x;
            s3 += (double)avv[u] * x;
        }
    };
    load((uint32_t)lane);  // issued before the sparse part
    for (uint32_t kb = k0 + lane; kb < k1; kb += U * LANES) {
        uint32_t cv[U];
        T av[U], uv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t k = kb + u * LANES;
            cv[u] = k < k1 ? ci[k] : 0u;
            av[u] = k < k1 ? v[k] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) uv[u] = upre[cv[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) s1 += av[u] * uv[u];
    }
    __syncthreads();  // fs complete
    add((uint32_t)lane);
    for (uint32_t cb = lane + UW * LANES; cb < nc; cb += UW * LANES) {
        load(cb);
        add(cb);
    }
    s1 = group_sum<LANES>(s1);
    s2 = group_sum<LANES>(s2);
    s3 = group_sum<LANES>(s3);
    if (g.closed()) return;
    if (on && lane == 0) {
        const double u = (double)upre[row] + s2;
        const double r = (double)f[row] - ((double)s1 + s3);
        out[row] = (T)(u + jstep<double, INV>((double)w, r, (double)d[row]));
    }
}

// set-up products of the fused tail, float64 and deterministic (every
// output element is one thread's fixed-order sum); dense n x nc matrices are
// row-major, so the products with the sparse n x n operator are sparse times
// dense: the transposes of R A and R A Dj^-1 A are At R^T and At (...).
// X (n x nc) := R^T for the CSR matrix R (nc x n); X zeroed before
__global__ void k_dense_of_csr_t(const uint32_t* rp, const uint32_t* ci, const double* v, uint32_t nc, double* X) {
    const uint32_t c = blockIdx.x;
    if (c >= nc) return;
    for (uint32_t p = rp[c] + threadIdx.x; p < rp[c + 1]; p += blockDim.x) X[(size_t)ci[p] * nc + c] = v[p];
}
// Y = P X (P: n x m CSR, X: m x nc dense), one block per row of P; row i
// then divided by d[i] (INV: multiplied by it) when d is not null
template <bool INV>
__global__ void k_csr_times_dense(const uint32_t* rp, const uint32_t* ci, const double* v, uint32_t n,
                                  const double* X, uint32_t nc, const double* d, double* Y) {
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    for (uint32_t c = threadIdx.x; c < nc; c += blockDim.x) {
        double s = 0.0;
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) s += v[p] * X[(size_t)ci[p] * nc + c];
        if (d) s = jdiv<double, INV>(s, d[i]);
        Y[(size_t)i * nc + c] = s;
    }
}
// Mpre = (X1 - 2w Xa + w^2 Dj^-1 X4)^T for n x nc inputs, written nc x n
template <bool INV>
__global__ void k_tail_mpre(const double* X1, const double* Xa, const double* X4, const double* d, double w,
                            uint32_t n, uint32_t nc, double* M) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)n * nc) return;
    const uint32_t c = (uint32_t)(t / n), j = (uint32_t)(t % n);  // output order: coalesced stores
    const size_t q = (size_t)j * nc + c;
    M[t] = X1[q] - 2.0 * w * Xa[q] + w * w * jdiv<double, INV>(X4[q], d[j]);
}

template <typename T>
__global__ void k_zero(T* u, uint32_t n, Gate g) {
    if (g.closed()) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = (T)0;
}

static inline dim3 g1(uint32_t n) { return dim3((n + 255) / 256); }

// ---------------------------------------------------------------------------
// preconditioner
// ---------------------------------------------------------------------------
struct PcAmg;
// Q A Q^T on the colour-ordered form of an AMG hierarchy's finest level
struct PermOp : LinOp {
    AmgLevelDev* D = nullptr;
    uint32_t rows() const override { return D->n; }
    uint32_t cols() const override { return D->n; }
    template <typename T>
    unsigned run(const T* x, T* y, double* part, hipStream_t s, Gate g) const;
    void apply(const double* x, double* y, hipStream_t s, Gate g = {}) const override { run<double>(x, y, nullptr, s, g); }
    void apply32(const float* x, float* y, hipStream_t s, Gate g = {}) const override { run<float>(x, y, nullptr, s, g); }
    unsigned apply_dot(const double* x, double* y, double* part, hipStream_t s, Gate g = {}) const override {
        return run<double>(x, y, part, s, g);
    }
    unsigned apply32_dot(const float* x, float* y, double* part, hipStream_t s, Gate g = {}) const override {
        return run<float>(x, y, part, s, g);
    }
    unsigned dot_parts() const override { return std::max(D->O.plan.nblk, D->O.plan32.nblk); }
    void ensure_fp32(hipStream_t) override {}
    uint64_t bytes_per_apply(bool fp32) const override {
        const uint64_t sv = fp32 ? 4 : 8;
        return D->O.nnz * (sv + 4) + 4ull * (D->n + 1) + 4ull * D->n * sv;  // entries, d, x, y (x gathered once)
    }
};
// Q M^{-1} Q^T: the V-cycle with the finest level in colour order (float32)
struct PermPc : Precond {
    PcAmg* M = nullptr;
    int kind() const override;
    uint32_t size() const override;
    void apply(const double*, double*, hipStream_t, Gate = {}) override {
        fail(FEM_ERR_RUNTIME, "AMG: the colour-ordered form runs in float32 only");
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g = {}) override;
    bool has_fp32() const override { return true; }
    unsigned apply32_dot(const float* r, float* z, double* part, hipStream_t s, Gate g = {}) override;
    unsigned dot_parts() const override;
    double apply_bytes(int s) override;
};

struct PcAmg : Precond {
    int kind_;
    bool device_setup = false;  // hierarchy built on the device
    uint64_t src_uid = 0;       // the matrix the hierarchy was built from (CsrMat::uid)
    PermOp pop_;
    PermPc ppc_;
    std::vector<std::vector<uint8_t>> splits;  // C/F: split per coarsened level
    bool inv_diag;  // SA smoother multiplies by the inverse diagonal
    int smoother, npre, npost;
    double omega;
    std::vector<std::unique_ptr<AmgLevelDev>> lv;
    DevBuf<double> coarse;  // nc x nc
    uint32_t nco = 0;
    bool fp32 = false;
    bool plain0_ = true;  // the finest level's grouped copy (and shadows) built
    // fused tail on level lv.size() - 2 (k_tail_pre / k_tail_post)
    bool tail_ = false;
    DevBuf<double> tW_, tAW_, tM_;  // W = P C, AW = A P C (n x nco), Mpre (nco x n)
    DevBuf<float> tM32_, tW32_, tAW32_;
    DevBuf<double> tFcp_;  // the coarse right-hand side in column-chunk partials

    PcAmg(int kind, AmgDevHierarchy&& H, const fem_pc_params& p, hipStream_t s)
        : kind_(kind), inv_diag(kind == FEM_PC_SAAMG || kind == FEM_PC_BSAAMG), smoother(p.smoother), npre(p.npre),
          npost(p.npost), omega(p.omega) {
        const size_t L = H.lv.size();
        levels = (int)L;
        SetupTrace tr("AMG levels", s);
        lv = std::move(H.lv);
        for (size_t k = 0; k < L; ++k) {
            AmgLevelDev& D = *lv[k];
            lvl_rows.push_back(D.n);
            lvl_nnz.push_back(D.A.nnz);
            D.A.stream = s;
            if (k > 0) D.A.use_vector_rows();  // coarse Galerkin operators
            if (k + 1 < L) {
                // transfer operators: summation order is free, long rows get
                // lane groups
                D.P.use_vector_rows();
                D.R.use_vector_rows();
                D.u.alloc(D.n);
                D.u2.alloc(D.n);
                D.res.alloc(D.n);
                if (smoother == FEM_SMOOTH_SGS) {
                    // a float32 hierarchy whose finest level may take the
                    // colour-ordered form builds that level's grouped copy
                    // (and its float32 shadows) only when a solve or an apply
                    // needs the level in its own order (ensure_plain)
                    const bool lazy = k == 0 && p.fp32 && npost >= 1 && D.n > 8192 && D.omega == 1.0 && npre >= 1 &&
                                      perm_env();
                    sgs_setup_device(D, s, !lazy);
                    if (lazy && D.use_gs) {
                        if (D.ncol >= 2) plain0_ = false;
                        else sgs_grouped_device(D, s);
                    }
                    if (residual_shortcut(D)) {
                        const uint32_t lo = D.h_cptr[D.ncol - 1], hi = D.h_cptr[D.ncol];
                        drop_columns_device(D.R, D.crows.get() + lo, hi - lo, D.Rz, s);
                        D.Rz.use_vector_rows();
                    }
                    tr.lap("smoother colouring and grouped copy");
                }
            } else {
                D.u.alloc(D.n);
            }
            if (k > 0) D.f.alloc(D.n);
        }
        nco = lv.back()->n;
        coarse.upload(H.coarse, s);
        FEM_HIP(hipStreamSynchronize(s));
        tr.lap("vectors, coarsest operator");
        build_tail(s);
        if (tail_) tr.lap("fused tail operators");
    }

    // The fused tail applies on the level above a dense coarsest level that
    // smooths by damped Jacobi with two pre-sweeps and at least one
    // post-sweep, while its dense operators stay small (FEM_AMG_TAIL=0 in
    // the environment keeps the separate launches, for comparison).
    // Level j (>= 1) takes its first damped Jacobi sweep from zero from the
    // restriction that writes its right-hand side (EPI_STORE_J0 /
    // EPI_STORE_J0_INV: o = jac0(f, d)) wherever that sweep would otherwise
    // be a launch of its own (not on levels whose first two sweeps run as one
    // launch, nor on the fused tail's level).
    // FEM_AMG_J0=0 in the environment keeps the sweep's own launch (the
    // results are the same bit for bit; for comparison).
    bool j0_from_restriction(size_t j) const {
        if (j == 0 || j + 1 >= lv.size() || npre < 1 || !j0_env_) return false;
        const AmgLevelDev& D = *lv[j];
        if (smoother == FEM_SMOOTH_SGS && D.use_gs) return false;
        return !fuse_first2(D, npre);
    }
    const bool j0_env_ = [] {
        const char* e = std::getenv("FEM_AMG_J0");
        return !(e && e[0] == '0');
    }();
    // the restriction's epilogue into level j (its right-hand side f, and
    // the first sweep into u when j0_from_restriction(j))
    template <typename T>
    int restrict_epi(size_t j, SpmvArgs<T>& a) {
        if (!j0_from_restriction(j)) return EPI_STORE;
        AmgLevelDev& C = *lv[j];
        Bufs<T> BC = bufs<T>(C);
        a.d = jd(C, BC);
        a.w = jwt(C);
        a.o = BC.u;
        return jinv() ? EPI_STORE_J0_INV : EPI_STORE_J0;
    }
    static constexpr uint64_t kTailDense = 2ull << 20;  // entries per dense tail operator
    static double tail_chunks(const AmgLevelDev& D) { return (double)((D.n + kTailChunk - 1) / kTailChunk); }
    bool tail_ok() const {
        const size_t L = lv.size();
        if (L < 3 || npre != 2 || npost < 1) return false;
        const AmgLevelDev& D = *lv[L - 2];
        if (smoother == FEM_SMOOTH_SGS && D.use_gs) return false;
        if (!fuse_first2(D, npre) || (uint64_t)D.n * nco > kTailDense || D.nc != nco) return false;
        const char* e = std::getenv("FEM_AMG_TAIL");
        return !(e && e[0] == '0');
    }
    void build_tail(hipStream_t s) {
        if (!tail_ok()) return;
        AmgLevelDev& D = *lv[lv.size() - 2];
        const uint32_t n = D.n, nc = nco;
        const size_t mn = (size_t)n * nc;
        const double* d = smoother == FEM_SMOOTH_SGS ? D.diag.get() : D.d.get();
        const bool inv = jinv();
        const double w = jwt(D);
        CsrMat At;
        transpose_device(D.A, At, s);
        // transposed: X1 = R^T, Xa = Dj^-1 At X1, X4 = At Xa; Mpre = R - 2w
        // Xa^T + w^2 (Dj^-1 X4)^T
        DevBuf<double> X1(mn), Xa(mn), X4(mn);
        FEM_HIP(hipMemsetAsync(X1.get(), 0, mn * sizeof(double), s));
        hipLaunchKernelGGL(k_dense_of_csr_t, dim3(nc), dim3(256), 0, s, D.R.rp.get(), D.R.ci.get(), D.R.v.get(), nc,
                           X1.get());
        if (inv)
            hipLaunchKernelGGL(k_csr_times_dense<true>, dim3(n), dim3(256), 0, s, At.rp.get(), At.ci.get(),
                               At.v.get(), n, X1.get(), nc, d, Xa.get());
        else
            hipLaunchKernelGGL(k_csr_times_dense<false>, dim3(n), dim3(256), 0, s, At.rp.get(), At.ci.get(),
                               At.v.get(), n, X1.get(), nc, d, Xa.get());
        hipLaunchKernelGGL(k_csr_times_dense<false>, dim3(n), dim3(256), 0, s, At.rp.get(), At.ci.get(), At.v.get(),
                           n, Xa.get(), nc, (const double*)nullptr, X4.get());
        tM_.alloc(mn);
        const dim3 G((unsigned)((mn + 255) / 256));
        if (inv)
            hipLaunchKernelGGL(k_tail_mpre<true>, G, dim3(256), 0, s, X1.get(), Xa.get(), X4.get(), d, w, n, nc,
                               tM_.get());
        else
            hipLaunchKernelGGL(k_tail_mpre<false>, G, dim3(256), 0, s, X1.get(), Xa.get(), X4.get(), d, w, n, nc,
                               tM_.get());
        // W = P C, AW = A W
        tW_.alloc(mn);
        tAW_.alloc(mn);
        tFcp_.alloc((size_t)((n + kTailChunk - 1) / kTailChunk) * nc);
        hipLaunchKernelGGL(k_csr_times_dense<false>, dim3(n), dim3(256), 0, s, D.P.rp.get(), D.P.ci.get(),
                           D.P.v.get(), n, coarse.get(), nc, (const double*)nullptr, tW_.get());
        hipLaunchKernelGGL(k_csr_times_dense<false>, dim3(n), dim3(256), 0, s, D.A.rp.get(), D.A.ci.get(),
                           D.A.v.get(), n, tW_.get(), nc, (const double*)nullptr, tAW_.get());
        FEM_HIP(hipGetLastError());
        FEM_HIP(hipStreamSynchronize(s));  // the temporaries end with this call
        tail_ = true;
    }
    template <typename T, bool INV, int LANES>
    void tail_launch(bool pre, AmgLevelDev& D, const T* f, const T* d, const T* upre_in, T* out, hipStream_t s,
                     Gate g) {
        constexpr bool F64 = std::is_same<T, double>::value;
        const T* v;
        const T* mp;
        if constexpr (F64) {
            v = D.A.v.get();
            mp = tM_.get();
        } else {
            v = D.A.v32.get();
            mp = tM32_.get();
        }
        if (!v || !mp) fail(FEM_ERR_RUNTIME, "AMG: fused tail without its operators in this precision");
        const T w = (T)jwt(D);
        const uint32_t nrb = (uint32_t)(((uint64_t)D.n * LANES + kBlock - 1) / kBlock);
        const uint32_t nq = (D.n + kTailChunk - 1) / kTailChunk;
        if (pre) {
            hipLaunchKernelGGL((k_tail_pre<T, INV, LANES>), dim3(nrb + nco * nq), dim3(kBlock), 0, s, D.A.rp.get(),
                               D.A.ci.get(), v, f, d, w, D.n, out, mp, nco, tFcp_.get(), g);
        } else {
            const T* W;
            const T* AW;
            if constexpr (F64) {
                W = tW_.get();
                AW = tAW_.get();
            } else {
                W = tW32_.get();
                AW = tAW32_.get();
            }
            if (!W || !AW) fail(FEM_ERR_RUNTIME, "AMG: fused tail without its operators in this precision");
            hipLaunchKernelGGL((k_tail_post<T, T, INV, LANES>), dim3(nrb), dim3(kBlock), (size_t)nco * sizeof(double), s,
                               D.A.rp.get(),
                               D.A.ci.get(), v, f, d, w, D.n, upre_in, W, AW, nco, (const double*)tFcp_.get(), nq,
                               out, g);
        }
        FEM_HIP(hipGetLastError());
    }
    // pre: u_pre into out and the coarse right-hand side into tFcp_; post:
    // the prolongation of its coarse solution and the first post-sweep from
    // upre into out
    template <typename T>
    void tail(bool pre, AmgLevelDev& D, const T* f, const T* d, const T* upre, T* out, hipStream_t s, Gate g) {
        const bool wide = D.A.lanes == 64;
        if (jinv()) {
            if (wide) tail_launch<T, true, 64>(pre, D, f, d, upre, out, s, g);
            else tail_launch<T, true, 32>(pre, D, f, d, upre, out, s, g);
        } else {
            if (wide) tail_launch<T, false, 64>(pre, D, f, d, upre, out, s, g);
            else tail_launch<T, false, 32>(pre, D, f, d, upre, out, s, g);
        }
    }
    int kind() const override { return kind_; }
    uint32_t size() const override { return lv[0]->n; }
    const CsrMat* amg_matrix(int l, int which) const override {
        if (l < 0 || l >= (int)lv.size()) return nullptr;
        const CsrMat& M = which == 0 ? lv[l]->A : which == 1 ? lv[l]->P : lv[l]->R;
        return M.n ? &M : nullptr;
    }
    const std::vector<uint8_t>* amg_split(int l) const override {
        return l >= 0 && l < (int)splits.size() ? &splits[l] : nullptr;
    }
    bool has_fp32() const override { return fp32; }
    // float32 shadows of a level's operators in their own order (A, P, R, Rz
    // and the grouped copy with its blocking)
    static void plain_shadows32(AmgLevelDev& D, bool last, hipStream_t s) {
        D.A.ensure_fp32(s);
        if (last) return;
        D.P.ensure_fp32(s);
        D.R.ensure_fp32(s);
        if (D.Rz.n) D.Rz.ensure_fp32(s);
        if (D.gv.get()) {
            const size_t nz = D.gv.n;
            D.gv32.alloc(nz);
            vec_convert<float, double>(D.gv32.get(), D.gv.get(), 1.0, nz, s);
            sgs_blocks32(D, s);
        }
    }
    void ensure_fp32(hipStream_t s) override {
        if (fp32) return;
        SetupTrace tr("AMG float32", s);
        for (size_t k = 0; k < lv.size(); ++k) {
            auto& D = *lv[k];
            const bool last = k + 1 == lv.size();
            if (k > 0 || plain0_) plain_shadows32(D, last, s);
            if (!last) {
                D.d32.alloc(D.n);
                vec_convert<float, double>(D.d32.get(), D.d.get(), 1.0, D.n, s);
                if (D.diag.get()) {
                    D.diag32.alloc(D.n);
                    vec_convert<float, double>(D.diag32.get(), D.diag.get(), 1.0, D.n, s);
                }
                if (D.gd.get()) {
                    D.gd32.alloc(D.n);
                    vec_convert<float, double>(D.gd32.get(), D.gd.get(), 1.0, D.n, s);
                }
                D.u232.alloc(D.n);
                D.res32.alloc(D.n);
            }
            D.u32.alloc(D.n);
            if (k > 0) D.f32.alloc(D.n);
            tr.lap(k == 0 ? "shadows, level 0" : "shadows, level >= 1");
        }
        if (tail_) {
            // float32 copies: the products with fc are still summed in
            // float64 (the coarse correction was rounded to float32 before,
            // as u_c and through P)
            tM32_.alloc(tM_.n);
            vec_convert<float, double>(tM32_.get(), tM_.get(), 1.0, tM_.n, s);
            tW32_.alloc(tW_.n);
            vec_convert<float, double>(tW32_.get(), tW_.get(), 1.0, tW_.n, s);
            tAW32_.alloc(tAW_.n);
            vec_convert<float, double>(tAW32_.get(), tAW_.get(), 1.0, tAW_.n, s);
        }
        fp32 = true;
        if (!plain0_ && !perm_ok()) ensure_plain(s);  // not expected: the form's conditions held at setup
        if (perm_ok()) {
            perm_setup_device(*lv[0], s);
            pop_.D = lv[0].get();
            ppc_.M = this;
            tr.lap("colour-ordered form");
        }
        FEM_HIP(hipStreamSynchronize(s));
    }

    // The float32 inner solves run on the colour-ordered form of the finest
    // level when its Gauss-Seidel passes have weight 1 and take the residual
    // shortcut (FEM_AMG_PERM=0 in the environment keeps A's own order, for
    // comparison). In that order every colour is a contiguous range of the
    // vectors (no row list, no strided access), the operator's diagonal is
    // kept apart (the off-diagonal entries of a row all sit in other colours,
    // so 16-bit column packing applies), the Gauss-Seidel update with weight
    // 1 does not read the row's own old value, and the prolongation fills
    // only the rows the post-smoothing reads before rewriting them (the last
    // colour's rows are rewritten first).
    bool perm_ok() const {
        if (smoother != FEM_SMOOTH_SGS || lv.size() < 2 || npost < 1) return false;
        const AmgLevelDev& D = *lv[0];
        if (!D.use_gs || D.ncol < 2 || !residual_shortcut(D) || !D.Rz.n) return false;
        return perm_env();
    }
    static bool perm_env() {
        const char* e = std::getenv("FEM_AMG_PERM");
        return !(e && e[0] == '0');
    }
    // the finest level in its own order (its grouped copy and, for a float32
    // hierarchy, that level's float32 shadows), built on first need; never
    // while the stream is being captured into a graph (a solver prepares it
    // before it captures anything)
    void ensure_plain(hipStream_t s) {
        if (plain0_) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        FEM_HIP(hipStreamIsCapturing(s, &cs));
        if (cs != hipStreamCaptureStatusNone)
            fail(FEM_ERR_RUNTIME, "AMG: the finest level's own-order data are built outside graph capture");
        AmgLevelDev& D = *lv[0];
        sgs_grouped_device(D, s);
        if (fp32) plain_shadows32(D, lv.size() == 1, s);
        FEM_HIP(hipStreamSynchronize(s));
        plain0_ = true;
    }
    void prepare(bool plain64, bool plain32, hipStream_t s) override {
        if (plain64 || plain32) ensure_plain(s);
    }
    bool permuted32(const LinOp* A, LinOp*& op, Precond*& pc, const uint32_t*& rows) override {
        if (!lv[0]->perm || src_uid == 0) return false;
        const CsrMat* c = dynamic_cast<const CsrMat*>(A);
        if (!c || c->uid != src_uid) return false;
        op = &pop_;
        pc = &ppc_;
        rows = lv[0]->crows.get();
        return true;
    }

    template <typename T>
    struct Bufs {
        T *f, *u, *u2, *res;
        const T *d, *diag;
    };
    template <typename T>
    Bufs<T> bufs(AmgLevelDev& D) {
        if constexpr (std::is_same<T, double>::value)
            return {D.f.get(), D.u.get(), D.u2.get(), D.res.get(), D.d.get(), D.diag.get()};
        else
            return {D.f32.get(), D.u32.get(), D.u232.get(), D.res32.get(), D.d32.get(), D.diag32.get()};
    }

    // Jacobi sweeps; `cur` holds the iterate (nullptr = zero); returns the
    // buffer holding the result. `a`/`b` are the two buffers to alternate.
    // `inv`: d holds 1/a_ii (else a_ii); `wt`: the damping weight
    template <typename T>
    T* jacobi(AmgLevelDev& D, const T* f, const T* d, bool inv, double wt, T* cur, T* a, T* b, int sweeps,
              hipStream_t s, Gate g) {
        const T w = (T)wt;
        for (int k = 0; k < sweeps; ++k) {
            T* out = (cur == a) ? b : a;
            if (!cur && fuse_first2(D, sweeps)) {
                // the first two sweeps in one launch (the second forms the
                // first's values from f and d where it reads them); the
                // result lands where the second sweep's would
                T* out2 = (out == a) ? b : a;
                SpmvArgs<T> ar;
                ar.f = f;
                ar.d = d;
                ar.w = wt;
                csr_spmv<T>(D.A, f, out2, inv ? EPI_JACOBI0_INV : EPI_JACOBI0, ar, s, g);
                cur = out2;
                ++k;
                continue;
            }
            if (!cur) {
                if (inv) hipLaunchKernelGGL((k_jacobi0<T, true>), g1(D.n), dim3(256), 0, s, f, d, out, D.n, w, g);
                else hipLaunchKernelGGL((k_jacobi0<T, false>), g1(D.n), dim3(256), 0, s, f, d, out, D.n, w, g);
            } else {
                SpmvArgs<T> ar;
                ar.f = f;
                ar.u = cur;
                ar.d = d;
                ar.w = wt;
                csr_spmv<T>(D.A, cur, out, inv ? EPI_JACOBI_INV : EPI_JACOBI, ar, s, g);
            }
            cur = out;
        }
        return cur;
    }

    // the CSR-stream blocks of the grouped matrix for sweeps in type T
    template <typename T>
    static const std::vector<uint32_t>& cblk(const AmgLevelDev& D) {
        return std::is_same<T, double>::value ? D.h_cblk : D.h_cblk32;
    }
    template <typename T>
    static const uint2* gblk(const AmgLevelDev& D) {
        return std::is_same<T, double>::value ? D.gblk.get() : D.gblk32.get();
    }

    // `from_zero`: u has not been initialised; the sweep starts from u = 0
    // `part` (non-null): the last sweep's passes also write the partial sums
    // of f.u, colour c's blocks at part + cblk[c] (stream kernels only)
    template <typename T, int LANES>
    void gs_launch(AmgLevelDev& D, const T* gv, const T* gd, const T* f, T* u, int sweeps, bool reverse,
                   bool from_zero, hipStream_t s, Gate g, double* part) {
        const T w = (T)D.omega;
        if (D.n <= 8192) {
            if (from_zero) hipLaunchKernelGGL((k_zero<T>), g1(D.n), dim3(256), 0, s, u, D.n, g);
            hipLaunchKernelGGL((k_gs_small<T, LANES>), dim3(1), dim3(1024), 0, s, D.ncol, D.cptr.get(),
                               D.crows.get(), D.gp.get(), D.gci.get(), gv, gd, f, u, w, sweeps, reverse ? 1 : 0, g);
            return;
        }
        for (int sw = 0; sw < sweeps; ++sw)
            for (int cc = 0; cc < D.ncol; ++cc) {
                const int c = reverse ? D.ncol - 1 - cc : cc;
                if (from_zero && sw == 0 && cc == 0) {
                    const uint32_t lo = D.h_cptr[c], hi = D.h_cptr[c + 1];
                    if (lo != 0) fail(FEM_ERR_RUNTIME, "AMG: a sweep from zero must start with the first colour");
                    hipLaunchKernelGGL((k_gs_first<T>), g1(D.n), dim3(256), 0, s, D.crows.get(), D.n, hi, gd, f, u, w,
                                       g);
                    continue;
                }
                const uint32_t b0 = cblk<T>(D)[c], b1 = cblk<T>(D)[c + 1];
                if (b0 >= b1) continue;
                const dim3 grid(b1 - b0), block(kBlock);
                const uint2* blk = gblk<T>(D) + b0;
                if (part && sw == sweeps - 1) {
                    if (D.gcd.get())
                        hipLaunchKernelGGL((k_gs_stream<T, 0, true, true>), grid, block, 0, s, blk, D.crows.get(),
                                           D.gp.get(), D.gci.get(), D.gcd.get(), D.gcbase.get(), gv, gd, f, u, w,
                                           (T*)nullptr, 0u, g, part + b0);
                    else
                        hipLaunchKernelGGL((k_gs_stream<T, 0, false, true>), grid, block, 0, s, blk, D.crows.get(),
                                           D.gp.get(), D.gci.get(), D.gcd.get(), D.gcbase.get(), gv, gd, f, u, w,
                                           (T*)nullptr, 0u, g, part + b0);
                } else if (D.gcd.get()) {
                    hipLaunchKernelGGL((k_gs_stream<T, 0, true>), grid, block, 0, s, blk, D.crows.get(), D.gp.get(),
                                       D.gci.get(), D.gcd.get(), D.gcbase.get(), gv, gd, f, u, w, (T*)nullptr, 0u, g,
                                       (double*)nullptr);
                } else {
                    hipLaunchKernelGGL((k_gs_stream<T, 0, false>), grid, block, 0, s, blk, D.crows.get(), D.gp.get(),
                                       D.gci.get(), D.gcd.get(), D.gcbase.get(), gv, gd, f, u, w, (T*)nullptr, 0u, g,
                                       (double*)nullptr);
                }
            }
        if (from_zero && sweeps == 0) hipLaunchKernelGGL((k_zero<T>), g1(D.n), dim3(256), 0, s, u, D.n, g);
    }

    template <typename T>
    void gs(AmgLevelDev& D, const T* f, T* u, int sweeps, bool reverse, hipStream_t s, Gate g,
            bool from_zero = false, double* part = nullptr) {
        if (sweeps <= 0 && !from_zero) return;
        const T *gv, *gd;
        if constexpr (std::is_same<T, double>::value) {
            gv = D.gv.get();
            gd = D.gd.get();
        } else {
            gv = D.gv32.get();
            gd = D.gd32.get();
        }
        switch (D.lanes) {
            case 1: gs_launch<T, 1>(D, gv, gd, f, u, sweeps, reverse, from_zero, s, g, part); break;
            case 4: gs_launch<T, 4>(D, gv, gd, f, u, sweeps, reverse, from_zero, s, g, part); break;
            case 8: gs_launch<T, 8>(D, gv, gd, f, u, sweeps, reverse, from_zero, s, g, part); break;
            case 16: gs_launch<T, 16>(D, gv, gd, f, u, sweeps, reverse, from_zero, s, g, part); break;
            case 32: gs_launch<T, 32>(D, gv, gd, f, u, sweeps, reverse, from_zero, s, g, part); break;
            default: gs_launch<T, 64>(D, gv, gd, f, u, sweeps, reverse, from_zero, s, g, part); break;
        }
    }

    // residual f - A u after a forward Gauss-Seidel sweep with weight 1 on a
    // large level: the last colour's rows are zero, the others are computed
    // on the colour-grouped matrix. With Rz (R without those columns) the
    // zeros are not written: only Rz reads this residual.
    bool residual_shortcut(const AmgLevelDev& D) const {
        return D.use_gs && D.n > 8192 && D.omega == 1.0 && npre >= 1 && D.ncol >= 2;
    }
    template <typename T>
    bool gs_residual(AmgLevelDev& D, const T* f, T* u, T* res, hipStream_t s, Gate g) {
        if (!residual_shortcut(D)) return false;
        const T* gv;
        if constexpr (std::is_same<T, double>::value) gv = D.gv.get();
        else gv = D.gv32.get();
        const uint32_t zero_from = cblk<T>(D)[D.ncol - 1];
        const uint32_t nb = D.Rz.n ? zero_from : cblk<T>(D)[D.ncol];
        if (nb == 0) return true;
        if (D.gcd.get())
            hipLaunchKernelGGL((k_gs_stream<T, 1, true>), dim3(nb), dim3(kBlock), 0, s, gblk<T>(D), D.crows.get(),
                               D.gp.get(), D.gci.get(), D.gcd.get(), D.gcbase.get(), gv, (const T*)nullptr, f, u, (T)1,
                               res, zero_from, g, (double*)nullptr);
        else
            hipLaunchKernelGGL((k_gs_stream<T, 1, false>), dim3(nb), dim3(kBlock), 0, s, gblk<T>(D), D.crows.get(),
                               D.gp.get(), D.gci.get(), D.gcd.get(), D.gcbase.get(), gv, (const T*)nullptr, f, u, (T)1,
                               res, zero_from, g, (double*)nullptr);
        return true;
    }

    // Jacobi data of a level: the SGS smoother's fallback divides by a_ii
    // with the level's Gershgorin weight; the Jacobi smoother uses d and the
    // configured (or robust) weight
    template <typename T>
    const T* jd(AmgLevelDev&, const Bufs<T>& B) const { return smoother == FEM_SMOOTH_SGS ? B.diag : B.d; }
    bool jinv() const { return smoother == FEM_SMOOTH_SGS ? false : inv_diag; }
    double jwt(const AmgLevelDev& D) const { return smoother == FEM_SMOOTH_SGS ? D.jw : D.omega; }

    // part (non-null, dot_fused()): the finest level's last smoothing sweep
    // also writes the partial sums of r.z
    // sweeps over the colour-ordered form of level D (weight 1), colours
    // ascending or descending; `from_zero`: u has not been initialised;
    // `part`: the last sweep's passes write the partial sums of f.u
    void gsp(AmgLevelDev& D, const float* f, float* u, int sweeps, bool reverse, bool from_zero, hipStream_t s,
             Gate g, double* part) {
        const CsrMat& O = D.O;
        const float* v = O.v32.get();
        const float* dg = D.gd32.get();
        for (int sw = 0; sw < sweeps; ++sw)
            for (int cc = 0; cc < D.ncol; ++cc) {
                const int c = reverse ? D.ncol - 1 - cc : cc;
                if (from_zero && sw == 0 && cc == 0) {
                    const uint32_t c0 = D.h_cptr[1];
                    const uint32_t zlo = D.ncol > 2 ? D.h_cptr[2] : D.n;
                    const uint32_t end = D.ncol > 2 ? D.n : c0;
                    hipLaunchKernelGGL((k_gsp_first<float>), g1(end), dim3(256), 0, s, c0, zlo, end, dg, f, u, g);
                    continue;
                }
                const uint32_t b0 = D.h_oblk32[c], b1 = D.h_oblk32[c + 1];
                if (b0 >= b1) continue;
                const dim3 grid(b1 - b0), block(kBlock);
                const uint2* blk = D.oblk32.get() + b0;
                if (part && sw == sweeps - 1) {
                    if (O.cd.get())
                        hipLaunchKernelGGL((k_gsp<float, 0, true, true>), grid, block, 0, s, blk, O.rp.get(),
                                           O.ci.get(), O.cd.get(), O.cbase.get(), v, dg, f, u, (float*)nullptr, g,
                                           part + b0);
                    else
                        hipLaunchKernelGGL((k_gsp<float, 0, false, true>), grid, block, 0, s, blk, O.rp.get(),
                                           O.ci.get(), O.cd.get(), O.cbase.get(), v, dg, f, u, (float*)nullptr, g,
                                           part + b0);
                } else if (O.cd.get()) {
                    hipLaunchKernelGGL((k_gsp<float, 0, true, false>), grid, block, 0, s, blk, O.rp.get(), O.ci.get(),
                                       O.cd.get(), O.cbase.get(), v, dg, f, u, (float*)nullptr, g, (double*)nullptr);
                } else {
                    hipLaunchKernelGGL((k_gsp<float, 0, false, false>), grid, block, 0, s, blk, O.rp.get(),
                                       O.ci.get(), O.cd.get(), O.cbase.get(), v, dg, f, u, (float*)nullptr, g,
                                       (double*)nullptr);
                }
            }
    }
    // residual of every colour but the last after a forward sweep (the last
    // colour's residual is zero and Rzp does not read it)
    void gsp_residual(AmgLevelDev& D, const float* f, float* u, float* res, hipStream_t s, Gate g) {
        const CsrMat& O = D.O;
        const uint32_t nb = D.h_oblk32[D.ncol - 1];
        if (nb == 0) return;
        if (O.cd.get())
            hipLaunchKernelGGL((k_gsp<float, 1, true, false>), dim3(nb), dim3(kBlock), 0, s, D.oblk32.get(),
                               O.rp.get(), O.ci.get(), O.cd.get(), O.cbase.get(), O.v32.get(), D.gd32.get(), f, u,
                               res, g, (double*)nullptr);
        else
            hipLaunchKernelGGL((k_gsp<float, 1, false, false>), dim3(nb), dim3(kBlock), 0, s, D.oblk32.get(),
                               O.rp.get(), O.ci.get(), O.cd.get(), O.cbase.get(), O.v32.get(), D.gd32.get(), f, u,
                               res, g, (double*)nullptr);
    }

    // `perm`: r and z are in the colour order of the finest level's
    // colour-ordered form (float32 only)
    template <typename T>
    void vcycle(const T* r, T* z, hipStream_t s, Gate g, double* part = nullptr, bool perm = false) {
        const size_t L = lv.size();
        if (perm && (!std::is_same<T, float>::value || !lv[0]->perm))
            fail(FEM_ERR_RUNTIME, "AMG: no colour-ordered float32 form of this hierarchy");
        std::vector<T*> ucur(L, nullptr), xa(L, nullptr), xb(L, nullptr);
        std::vector<const T*> fk(L, nullptr);
        fk[0] = r;
        for (size_t k = 0; k + 1 < L; ++k) {
            AmgLevelDev& D = *lv[k];
            Bufs<T> B = bufs<T>(D);
            const T* f = fk[k];
            T* u;
            if constexpr (std::is_same<T, float>::value) {
                if (perm && k == 0) {
                    gsp(D, f, z, npre, false, true, s, g, nullptr);
                    gsp_residual(D, f, z, B.res, s, g);
                    AmgLevelDev& C = *lv[1];
                    Bufs<T> BC = bufs<T>(C);
                    SpmvArgs<T> ra;
                    const int repi = restrict_epi<T>(1, ra);
                    csr_spmv<T>(D.Rzp, B.res, BC.f, repi, ra, s, g);
                    ucur[0] = z;
                    fk[1] = BC.f;
                    continue;
                }
            }
            if (tail_ && k + 2 == L) {
                // the two pre-sweeps (into u2, where the fused sweeps put
                // them) and the coarse right-hand side in one launch
                xa[k] = B.u;
                xb[k] = B.u2;
                tail<T>(true, D, f, jd(D, B), nullptr, B.u2, s, g);
                ucur[k] = B.u2;
                continue;
            }
            if (smoother == FEM_SMOOTH_SGS && D.use_gs) {
                u = (k == 0) ? z : B.u;
                gs<T>(D, f, u, npre, false, s, g, true);
            } else {
                // Jacobi alternates between two buffers; on level 0 pick the
                // first one so that the last post-smoothing sweep writes z
                T* X = B.u;
                T* Y = B.u2;
                if (k == 0) {
                    const int toggles = (npre >= 1 ? npre - 1 : 0) + npost;
                    X = (toggles % 2 == 0) ? z : B.u2;
                    Y = (toggles % 2 == 0) ? B.u2 : z;
                }
                xa[k] = X;
                xb[k] = Y;
                if (npre > 0 && j0_from_restriction(k)) {
                    // the first sweep is in u (= X) already
                    u = npre > 1 ? jacobi<T>(D, f, jd(D, B), jinv(), jwt(D), X, X, Y, npre - 1, s, g) : X;
                } else if (npre > 0) {
                    u = jacobi<T>(D, f, jd(D, B), jinv(), jwt(D), nullptr, X, Y, npre, s, g);
                } else {
                    hipLaunchKernelGGL((k_zero<T>), g1(D.n), dim3(256), 0, s, X, D.n, g);
                    u = X;
                }
            }
            ucur[k] = u;
            SpmvArgs<T> ar;
            ar.f = f;
            const bool shortcut = smoother == FEM_SMOOTH_SGS && gs_residual<T>(D, f, u, B.res, s, g);
            if (!shortcut) csr_spmv<T>(D.A, u, B.res, EPI_RESID, ar, s, g);
            AmgLevelDev& C = *lv[k + 1];
            Bufs<T> BC = bufs<T>(C);
            SpmvArgs<T> ra;
            const int repi = restrict_epi<T>(k + 1, ra);
            csr_spmv<T>(shortcut && D.Rz.n ? D.Rz : D.R, B.res, BC.f, repi, ra, s, g);
            fk[k + 1] = BC.f;
        }
        // coarsest level: u = C f (dense, float64 accumulation; inside the
        // fused tail's second launch when it applies)
        if (!tail_) {
            AmgLevelDev& C = *lv[L - 1];
            Bufs<T> BC = bufs<T>(C);
            T* u = (L == 1) ? z : BC.u;
            const size_t sh = (size_t)nco * sizeof(double);
            hipLaunchKernelGGL((k_coarse<T>), dim3((nco + 3) / 4), dim3(256), sh, s, coarse.get(), nco, fk[L - 1], u,
                               g);
            ucur[L - 1] = u;
        }
        for (int64_t k = (int64_t)L - 2; k >= 0; --k) {
            AmgLevelDev& D = *lv[k];
            Bufs<T> B = bufs<T>(D);
            T* u = ucur[k];
            SpmvArgs<T> ar;
            ar.u = u;
            if constexpr (std::is_same<T, float>::value) {
                if (perm && k == 0) {
                    csr_spmv<T>(D.Pp, ucur[1], u, EPI_ADD, ar, s, g);
                    gsp(D, fk[0], u, npost, true, false, s, g, part);
                    continue;
                }
            }
            if (tail_ && k + 2 == (int64_t)L) {
                tail<T>(false, D, fk[k], jd(D, B), u, xa[k], s, g);
                ucur[k] = npost > 1 ? jacobi<T>(D, fk[k], jd(D, B), jinv(), jwt(D), xa[k], xa[k], xb[k], npost - 1, s, g)
                                    : xa[k];
                continue;
            }
            csr_spmv<T>(D.P, ucur[k + 1], u, EPI_ADD, ar, s, g);
            if (smoother == FEM_SMOOTH_SGS && D.use_gs) gs<T>(D, fk[k], u, npost, true, s, g, false, k == 0 ? part : nullptr);
            else u = jacobi<T>(D, fk[k], jd(D, B), jinv(), jwt(D), u, xa[k], xb[k], npost, s, g);
            ucur[k] = u;
        }
        FEM_HIP(hipGetLastError());
    }

    void apply(const double* r, double* z, hipStream_t s, Gate g) override {
        ensure_plain(s);
        vcycle<double>(r, z, s, g);
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override {
        ensure_plain(s);
        vcycle<float>(r, z, s, g);
    }

    // r.z from the V-cycle itself: the finest level's post-smoothing ends
    // with one stream pass per colour, which reads f = r and writes each
    // row's final z (rows of one colour are written by that colour's pass
    // only), so those passes' blocks give the partial sums in block order.
    bool dot_fused() const {
        if (smoother != FEM_SMOOTH_SGS || npost < 1 || lv.size() < 2) return false;
        const AmgLevelDev& D = *lv[0];
        return D.use_gs && D.n > 8192 && !D.h_cblk.empty();
    }
    unsigned dot_parts() const override { return dot_fused() ? lv[0]->h_cblk.back() : 0; }
    unsigned apply_dot(const double* r, double* z, double* part, hipStream_t s, Gate g) override {
        ensure_plain(s);
        if (!dot_fused()) return 0;
        vcycle<double>(r, z, s, g, part);
        return lv[0]->h_cblk.back();
    }
    unsigned apply32_dot(const float* r, float* z, double* part, hipStream_t s, Gate g) override {
        // the float32 blocking never has more blocks than the float64 one
        // that sized the partial buffer; checked here all the same
        ensure_plain(s);
        if (!dot_fused() || lv[0]->h_cblk32.empty() || lv[0]->h_cblk32.back() > lv[0]->h_cblk.back()) return 0;
        vcycle<float>(r, z, s, g, part);
        return lv[0]->h_cblk32.back();
    }

    // Algorithmic bytes of one V-cycle, kernel by kernel in the order vcycle
    // launches them (s-byte values, uint32 indices).
    // (s = 4 with the colour-ordered form built: that form, the one float32
    // inner solves run)
    double apply_bytes(int s) override { return bytes_of(s, s == 4 && lv[0]->perm); }
    double bytes_of(int s, bool perm) {
        const size_t L = lv.size();
        double b = 0.0;
        for (size_t k = 0; k + 1 < L; ++k) {
            AmgLevelDev& D = *lv[k];
            // the restriction's second output: d read, the first sweep written
            const double j0x = j0_from_restriction(k + 1) ? 2.0 * lv[k + 1]->n * s : 0.0;
            if (perm && k == 0) {
                b += gsp_bytes(D, s, npre, true) + gsp_residual_bytes(D, s) + csr_bytes(D.Rzp, s, 0) + j0x;
                continue;
            }
            if (tail_ && k + 2 == L) {
                // the fused pre-sweeps (f gathered, d, u_pre) and Mpre f
                // (its float64 chunk partials)
                b += csr_bytes(D.A, s, 1) + (double)nco * D.n * s + 8.0 * nco * tail_chunks(D);
                continue;
            }
            if (smoother == FEM_SMOOTH_SGS && D.use_gs) b += gs_bytes(D, s, npre, true);
            else if (j0_from_restriction(k)) b += jacobi_bytes(D, s, npre - 1, false);
            else b += npre > 0 ? jacobi_bytes(D, s, npre, true) : (double)D.n * s;
            const bool shortcut = smoother == FEM_SMOOTH_SGS && residual_shortcut(D);
            b += shortcut ? gs_residual_bytes(D, s) : csr_bytes(D.A, s, 1);
            b += csr_bytes(shortcut && D.Rz.n ? D.Rz : D.R, s, 0) + j0x;
        }
        if (!tail_) b += (double)nco * nco * 8 + 2.0 * nco * s;  // dense coarsest operator, f, u
        for (int64_t k = (int64_t)L - 2; k >= 0; --k) {
            AmgLevelDev& D = *lv[k];
            if (tail_ && k + 2 == (int64_t)L) {
                // u_pre gathered through A, W and AW, f, d, fc and the
                // output; then the remaining post-sweeps
                b += csr_bytes(D.A, s, 2) + 2.0 * D.n * nco * s + 8.0 * nco * tail_chunks(D);
                for (int sw = 1; sw < npost; ++sw) b += csr_bytes(D.A, s, 2);
                continue;
            }
            if (perm && k == 0) {
                b += csr_bytes(D.Pp, s, 1) + gsp_bytes(D, s, npost, false);
                continue;
            }
            b += csr_bytes(D.P, s, 1);  // u += P e reads u
            b += (smoother == FEM_SMOOTH_SGS && D.use_gs) ? gs_bytes(D, s, npost, false)
                                                          : jacobi_bytes(D, s, npost, false);
        }
        return b;
    }

    // sweep 0 from zero reads f, d and writes u; a sweep from u is an SpMV
    // whose epilogue reads f and d
    // The first two Jacobi sweeps from zero run as one launch on levels of at
    // most kFuseRows rows, where a launch costs more than its work (on a
    // 188,709-row level at 128^3 the fused float32 sweep took 25.7 us
    // against 4.7 + 16.4 us: two gathers and a division per entry; on the
    // 5,537-row level 5.3 against 4.8 + 4.7 us).
    static constexpr uint32_t kFuseRows = 65536;
    static bool fuse_first2(const AmgLevelDev& D, int sweeps) { return sweeps >= 2 && D.n <= kFuseRows; }
    static double jacobi_bytes(const AmgLevelDev& D, int s, int sweeps, bool from_zero) {
        double b = 0.0;
        int k = 0;
        if (from_zero && fuse_first2(D, sweeps)) {
            b += csr_bytes(D.A, s, 1);  // the first two sweeps: f gathered, d, the output
            k = 2;
        }
        for (; k < sweeps; ++k) b += (from_zero && k == 0) ? 3.0 * D.n * s : csr_bytes(D.A, s, 2);
        return b;
    }

    // entries of the colour-grouped matrix in grouped rows [lo, hi)
    static uint64_t grouped_nnz(const AmgLevelDev& D, uint32_t lo, uint32_t hi) {
        if (!D.h_gp.empty()) return (uint64_t)D.h_gp.at(hi) - D.h_gp.at(lo);  // the host copy sgs_setup_device keeps
        // no grouped copy (yet): for [lo, hi) at colour bounds, the
        // colour-ordered form's off-diagonal entries plus one diagonal per row
        uint64_t z = hi - lo;
        for (int c = 0; c < D.ncol; ++c)
            if (D.h_cptr[c] >= lo && D.h_cptr[c + 1] <= hi) z += D.h_onnz.at(c);
        return z;
    }

    // one colour pass over r rows holding z entries: the rows' matrix part
    // and row list, u gathered once, f, the diagonal and the new u per row
    static double colour_bytes(const AmgLevelDev& D, int s, uint64_t z, uint32_t r) {
        return (double)z * (s + 4) + 4.0 * (r + 1) + 4.0 * r + (double)D.n * s + 3.0 * r * s;
    }

    double gs_bytes(const AmgLevelDev& D, int s, int sweeps, bool from_zero) const {
        if (sweeps <= 0 && !from_zero) return 0.0;
        if (D.n <= 8192)  // one workgroup runs every sweep; each re-reads the level
            return (from_zero ? (double)D.n * s : 0.0) + sweeps * colour_bytes(D, s, D.A.nnz, D.n);
        if (sweeps == 0) return (double)D.n * s;  // k_zero
        double b = 0.0;
        for (int sw = 0; sw < sweeps; ++sw)
            for (int c = 0; c < D.ncol; ++c) {
                const uint32_t lo = D.h_cptr[c], hi = D.h_cptr[c + 1];
                if (from_zero && sw == 0 && c == 0) {
                    b += 4.0 * D.n + 2.0 * (hi - lo) * s + (double)D.n * s;  // k_gs_first: row list, f, d; u
                    continue;
                }
                if (hi > lo) b += colour_bytes(D, s, grouped_nnz(D, lo, hi), hi - lo);
            }
        return b;
    }

    // colour-ordered passes: the colour's entries and row bounds, d, f and
    // the new u per row, the other colours' u gathered once; the first pass
    // from zero reads f and d and writes u for the first colour, zeros from
    // the third colour on
    double gsp_bytes(const AmgLevelDev& D, int s, int sweeps, bool from_zero) const {
        double b = 0.0;
        for (int sw = 0; sw < sweeps; ++sw)
            for (int c = 0; c < D.ncol; ++c) {
                const double r = (double)D.h_cptr[c + 1] - D.h_cptr[c];
                if (from_zero && sw == 0 && c == 0) {
                    b += 3.0 * r * s + (D.ncol > 2 ? (double)(D.n - D.h_cptr[2]) * s : 0.0);
                    continue;
                }
                if (r > 0) b += (double)D.h_onnz[c] * (s + 4) + 4.0 * (r + 1) + 3.0 * r * s + (D.n - r) * s;
            }
        return b;
    }
    // residual of the colours before the last: their entries and bounds, d,
    // f, u and the residual per row, the last colour's u gathered once
    double gsp_residual_bytes(const AmgLevelDev& D, int s) const {
        const double r = D.h_cptr[D.ncol - 1];
        double z = 0.0;
        for (int c = 0; c + 1 < D.ncol; ++c) z += (double)D.h_onnz[c];
        return z * (s + 4) + 4.0 * (r + 1) + 4.0 * r * s + (D.n - r) * s;
    }

    // the residual of every colour but the last (a zero written for each of
    // the last colour's rows unless Rz skips them): matrix part, row list,
    // u gathered once, f read and the residual written per row
    double gs_residual_bytes(const AmgLevelDev& D, int s) const {
        const uint32_t r = D.h_cptr[D.ncol - 1];
        double b = (double)grouped_nnz(D, 0, r) * (s + 4) + 4.0 * (r + 1) + 4.0 * r + (double)D.n * s + 2.0 * r * s;
        if (!D.Rz.n) b += (4.0 + s) * (D.n - r);
        return b;
    }
};

template <typename T>
unsigned PermOp::run(const T* x, T* y, double* part, hipStream_t s, Gate g) const {
    const CsrMat& O = D->O;
    constexpr bool F64 = std::is_same<T, double>::value;
    const SpmvPlan& plan = F64 ? O.plan : O.plan32;
    const T* v;
    const T* dg;
    if constexpr (F64) {
        v = O.v.get();
        dg = D->gd.get();
    } else {
        v = O.v32.get();
        dg = D->gd32.get();
    }
    if (!v || !dg || plan.nblk == 0) fail(FEM_ERR_RUNTIME, "AMG: colour-ordered operator not built");
    const dim3 grid(plan.nblk), block(kBlock);
    T* xu = const_cast<T*>(x);  // MODE 2 only reads it
    if (part) {
        if (O.cd.get())
            hipLaunchKernelGGL((k_gsp<T, 2, true, true>), grid, block, 0, s, plan.bnd.get(), O.rp.get(), O.ci.get(),
                               O.cd.get(), O.cbase.get(), v, dg, (const T*)nullptr, xu, y, g, part);
        else
            hipLaunchKernelGGL((k_gsp<T, 2, false, true>), grid, block, 0, s, plan.bnd.get(), O.rp.get(), O.ci.get(),
                               O.cd.get(), O.cbase.get(), v, dg, (const T*)nullptr, xu, y, g, part);
    } else {
        if (O.cd.get())
            hipLaunchKernelGGL((k_gsp<T, 2, true, false>), grid, block, 0, s, plan.bnd.get(), O.rp.get(), O.ci.get(),
                               O.cd.get(), O.cbase.get(), v, dg, (const T*)nullptr, xu, y, g, (double*)nullptr);
        else
            hipLaunchKernelGGL((k_gsp<T, 2, false, false>), grid, block, 0, s, plan.bnd.get(), O.rp.get(),
                               O.ci.get(), O.cd.get(), O.cbase.get(), v, dg, (const T*)nullptr, xu, y, g,
                               (double*)nullptr);
    }
    FEM_HIP(hipGetLastError());
    return part ? plan.nblk : 0;
}

int PermPc::kind() const { return M->kind(); }
uint32_t PermPc::size() const { return M->size(); }
void PermPc::apply32(const float* r, float* z, hipStream_t s, Gate g) { M->vcycle<float>(r, z, s, g, nullptr, true); }
unsigned PermPc::apply32_dot(const float* r, float* z, double* part, hipStream_t s, Gate g) {
    M->vcycle<float>(r, z, s, g, part, true);
    return M->lv[0]->h_oblk32.back();
}
unsigned PermPc::dot_parts() const { return M->lv[0]->h_oblk32.back(); }
double PermPc::apply_bytes(int s) { return M->bytes_of(s, true); }

std::unique_ptr<Precond> make_amg_cf(CsrMat& A, const fem_pc_params& p, hipStream_t s) {
    if (A.n != A.m) fail(FEM_ERR_INVALID_ARGUMENT, "AMG needs a square matrix");
    std::vector<std::vector<uint8_t>> splits;
    AmgDevHierarchy H;
    if (rows_sorted_device(A, s)) {
        H = build_cf_hierarchy_device(A, p, s, &splits);
    } else {
        // rows with unsorted or repeated columns: the same operator with its
        // columns sorted and repeats summed
        std::vector<uint32_t> rp, ci;
        std::vector<double> v;
        condense_device(A, 1, A.n, A.m, rp, ci, v, s);
        CsrMat B;
        B.set_host(A.n, A.m, std::move(rp), std::move(ci), std::move(v), s);
        H = build_cf_hierarchy_device(B, p, s, &splits);
    }
    auto pc = std::unique_ptr<PcAmg>(new PcAmg(FEM_PC_AMG, std::move(H), p, s));
    pc->src_uid = A.uid;
    pc->splits = std::move(splits);
    pc->device_setup = true;
    return pc;
}

std::unique_ptr<Precond> make_saamg(CsrMat& A, const fem_pc_params& p, hipStream_t s) {
    if (A.n != A.m) fail(FEM_ERR_INVALID_ARGUMENT, "Block AMG needs a square matrix");
    SetupTrace tr("saamg setup", s);
    // the robust variant is built on the device when asked for and the rows
    // are sorted and duplicate free (the host path condenses them first)
    bool dev = p.p_omega <= 0.0 && p.setup_device != 0 && rows_sorted_device(A, s);
    AmgDevHierarchy H;
    if (dev) {
        H = build_sa_hierarchy_device(A, p, s);
    } else {
        A.sync_host();
        H = upload_hierarchy(build_sa_hierarchy(A, p), s);
    }
    tr.lap(dev ? "hierarchy (device)" : "hierarchy (host)");
    auto pc = std::unique_ptr<PcAmg>(new PcAmg(FEM_PC_SAAMG, std::move(H), p, s));
    pc->src_uid = A.uid;
    tr.lap("levels ready");
    pc->device_setup = dev;
    return pc;
}

}  // namespace fem

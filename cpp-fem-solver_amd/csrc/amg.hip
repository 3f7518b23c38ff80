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
    const T r = fi - (T)0;
    if constexpr (INV) u[i] = (T)0 + w * (di * r);
    else u[i] = (T)0 + (w * r) / di;
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
    if (cnt <= (uint32_t)STAGE) {
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
struct PcAmg : Precond {
    int kind_;
    bool device_setup = false;  // hierarchy built on the device
    std::vector<std::vector<uint8_t>> splits;  // C/F: split per coarsened level
    bool inv_diag;  // SA smoother multiplies by the inverse diagonal
    int smoother, npre, npost;
    double omega;
    std::vector<std::unique_ptr<AmgLevelDev>> lv;
    DevBuf<double> coarse;  // nc x nc
    uint32_t nco = 0;
    bool fp32 = false;

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
                    sgs_setup_device(D, s);
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
    void ensure_fp32(hipStream_t s) override {
        if (fp32) return;
        for (size_t k = 0; k < lv.size(); ++k) {
            auto& D = *lv[k];
            D.A.ensure_fp32(s);
            if (k + 1 < lv.size()) {
                D.P.ensure_fp32(s);
                D.R.ensure_fp32(s);
                if (D.Rz.n) D.Rz.ensure_fp32(s);
                D.d32.alloc(D.n);
                vec_convert<float, double>(D.d32.get(), D.d.get(), 1.0, D.n, s);
                if (D.diag.get()) {
                    D.diag32.alloc(D.n);
                    vec_convert<float, double>(D.diag32.get(), D.diag.get(), 1.0, D.n, s);
                }
                if (D.gv.get()) {
                    const size_t nz = D.gv.n;
                    D.gv32.alloc(nz);
                    vec_convert<float, double>(D.gv32.get(), D.gv.get(), 1.0, nz, s);
                    D.gd32.alloc(D.n);
                    vec_convert<float, double>(D.gd32.get(), D.gd.get(), 1.0, D.n, s);
                    sgs_blocks32(D, s);
                }
                D.u232.alloc(D.n);
                D.res32.alloc(D.n);
            }
            D.u32.alloc(D.n);
            if (k > 0) D.f32.alloc(D.n);
        }
        fp32 = true;
        FEM_HIP(hipStreamSynchronize(s));
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
    template <typename T>
    void vcycle(const T* r, T* z, hipStream_t s, Gate g, double* part = nullptr) {
        const size_t L = lv.size();
        std::vector<T*> ucur(L, nullptr), xa(L, nullptr), xb(L, nullptr);
        std::vector<const T*> fk(L, nullptr);
        fk[0] = r;
        for (size_t k = 0; k + 1 < L; ++k) {
            AmgLevelDev& D = *lv[k];
            Bufs<T> B = bufs<T>(D);
            const T* f = fk[k];
            T* u;
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
                if (npre > 0) {
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
            csr_spmv<T>(shortcut && D.Rz.n ? D.Rz : D.R, B.res, BC.f, EPI_STORE, {}, s, g);
            fk[k + 1] = BC.f;
        }
        // coarsest level: u = C f (dense, float64 accumulation)
        {
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
            csr_spmv<T>(D.P, ucur[k + 1], u, EPI_ADD, ar, s, g);
            if (smoother == FEM_SMOOTH_SGS && D.use_gs) gs<T>(D, fk[k], u, npost, true, s, g, false, k == 0 ? part : nullptr);
            else u = jacobi<T>(D, fk[k], jd(D, B), jinv(), jwt(D), u, xa[k], xb[k], npost, s, g);
            ucur[k] = u;
        }
        FEM_HIP(hipGetLastError());
    }

    void apply(const double* r, double* z, hipStream_t s, Gate g) override { vcycle<double>(r, z, s, g); }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override { vcycle<float>(r, z, s, g); }

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
        if (!dot_fused()) return 0;
        vcycle<double>(r, z, s, g, part);
        return lv[0]->h_cblk.back();
    }
    unsigned apply32_dot(const float* r, float* z, double* part, hipStream_t s, Gate g) override {
        // the float32 blocking never has more blocks than the float64 one
        // that sized the partial buffer; checked here all the same
        if (!dot_fused() || lv[0]->h_cblk32.empty() || lv[0]->h_cblk32.back() > lv[0]->h_cblk.back()) return 0;
        vcycle<float>(r, z, s, g, part);
        return lv[0]->h_cblk32.back();
    }

    // Algorithmic bytes of one V-cycle, kernel by kernel in the order vcycle
    // launches them (s-byte values, uint32 indices).
    double apply_bytes(int s) override {
        const size_t L = lv.size();
        double b = 0.0;
        for (size_t k = 0; k + 1 < L; ++k) {
            AmgLevelDev& D = *lv[k];
            if (smoother == FEM_SMOOTH_SGS && D.use_gs) b += gs_bytes(D, s, npre, true);
            else b += npre > 0 ? jacobi_bytes(D, s, npre, true) : (double)D.n * s;
            const bool shortcut = smoother == FEM_SMOOTH_SGS && residual_shortcut(D);
            b += shortcut ? gs_residual_bytes(D, s) : csr_bytes(D.A, s, 1);
            b += csr_bytes(shortcut && D.Rz.n ? D.Rz : D.R, s, 0);
        }
        b += (double)nco * nco * 8 + 2.0 * nco * s;  // dense coarsest operator, f, u
        for (int64_t k = (int64_t)L - 2; k >= 0; --k) {
            AmgLevelDev& D = *lv[k];
            b += csr_bytes(D.P, s, 1);  // u += P e reads u
            b += (smoother == FEM_SMOOTH_SGS && D.use_gs) ? gs_bytes(D, s, npost, false)
                                                          : jacobi_bytes(D, s, npost, false);
        }
        return b;
    }

    // sweep 0 from zero reads f, d and writes u; a sweep from u is an SpMV
    // whose epilogue reads f and d
    static double jacobi_bytes(const AmgLevelDev& D, int s, int sweeps, bool from_zero) {
        double b = 0.0;
        for (int k = 0; k < sweeps; ++k) b += (from_zero && k == 0) ? 3.0 * D.n * s : csr_bytes(D.A, s, 2);
        return b;
    }

    // entries of the colour-grouped matrix in grouped rows [lo, hi)
    static uint64_t grouped_nnz(const AmgLevelDev& D, uint32_t lo, uint32_t hi) {
        uint32_t a = 0, e = 0;
        FEM_HIP(hipMemcpy(&a, D.gp.get() + lo, sizeof(uint32_t), hipMemcpyDeviceToHost));
        FEM_HIP(hipMemcpy(&e, D.gp.get() + hi, sizeof(uint32_t), hipMemcpyDeviceToHost));
        return (uint64_t)e - a;
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
    tr.lap("levels ready");
    pc->device_setup = dev;
    return pc;
}

}  // namespace fem

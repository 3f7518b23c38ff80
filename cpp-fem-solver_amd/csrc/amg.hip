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
//                correction and descending after it (symmetric V-cycle)
// Every level starts from u = 0, as the reference's V-cycle does.
#include <algorithm>
#include <cmath>

#include "amg.hpp"
#include "precond.hpp"

namespace fem {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// first Jacobi sweep from u = 0 (the SpMV of a zero vector is exactly 0)
template <typename T, bool INV>
__global__ void k_jacobi0(const T* __restrict__ f, const T* __restrict__ d, T* __restrict__ u, uint32_t n, T w,
                          Gate g) {
    if (g.closed()) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T r = f[i] - (T)0;
    if constexpr (INV) u[i] = (T)0 + w * (d[i] * r);
    else u[i] = (T)0 + (w * r) / d[i];
}

// one multicolour Gauss-Seidel colour: u_i += w (f_i - (A u)_i) / a_ii
template <typename T>
__global__ void k_gs_colour(uint32_t lo, uint32_t hi, const uint32_t* __restrict__ rows,
                            const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                            const T* __restrict__ v, const T* __restrict__ d, const T* __restrict__ f, T* u, T w,
                            Gate g) {
    if (g.closed()) return;
    const uint32_t t = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= hi) return;
    const uint32_t i = rows[t];
    T s = (T)0;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) s += v[p] * u[ci[p]];
    u[i] = u[i] + (w * (f[i] - s)) / d[i];
}

// all colours of `sweeps` Gauss-Seidel sweeps in one workgroup (small levels)
template <typename T>
__global__ __launch_bounds__(1024) void k_gs_small(int ncol, const uint32_t* __restrict__ cptr,
                                                   const uint32_t* __restrict__ rows, const uint32_t* __restrict__ rp,
                                                   const uint32_t* __restrict__ ci, const T* __restrict__ v,
                                                   const T* __restrict__ d, const T* __restrict__ f, T* u, T w,
                                                   int sweeps, int reverse, Gate g) {
    if (g.closed()) return;
    for (int sw = 0; sw < sweeps; ++sw) {
        for (int cc = 0; cc < ncol; ++cc) {
            const int c = reverse ? ncol - 1 - cc : cc;
            for (uint32_t t = cptr[c] + threadIdx.x; t < cptr[c + 1]; t += blockDim.x) {
                const uint32_t i = rows[t];
                T s = (T)0;
                for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) s += v[p] * u[ci[p]];
                u[i] = u[i] + (w * (f[i] - s)) / d[i];
            }
            __syncthreads();
        }
    }
}

// u = C f with C dense (row-major, float64); f, u of type T
template <typename T>
__global__ __launch_bounds__(256) void k_coarse(const double* __restrict__ C, uint32_t nc, const T* __restrict__ f,
                                                T* __restrict__ u, Gate g) {
    if (g.closed()) return;
    extern __shared__ double fs[];
    for (uint32_t j = threadIdx.x; j < nc; j += blockDim.x) fs[j] = (double)f[j];
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
        double s = 0.0;
        const double* row = C + (size_t)i * nc;
        for (uint32_t j = 0; j < nc; ++j) s += row[j] * fs[j];
        u[i] = (T)s;
    }
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
struct AmgLevelDev {
    CsrMat A, P, R;
    uint32_t n = 0, nc = 0;
    double omega = 0.0;  // smoother weight
    DevBuf<double> d;      // C/F: diagonal; SA: inverse diagonal
    DevBuf<float> d32;
    DevBuf<double> f, u, u2, res;
    DevBuf<float> f32, u32, u232, res32;
    // multicolour ordering for the SGS smoother
    int ncol = 0;
    std::vector<uint32_t> h_cptr;
    DevBuf<uint32_t> cptr, crows;
    DevBuf<double> diag;  // a_ii for SGS
    DevBuf<float> diag32;
};

struct PcAmg : Precond {
    int kind_;
    bool inv_diag;  // SA smoother multiplies by the inverse diagonal
    int smoother, npre, npost;
    double omega;
    std::vector<std::unique_ptr<AmgLevelDev>> lv;
    DevBuf<double> coarse;  // nc x nc
    uint32_t nco = 0;
    bool fp32 = false;

    PcAmg(int kind, AmgHierarchy&& H, const fem_pc_params& p, hipStream_t s)
        : kind_(kind), inv_diag(kind == FEM_PC_SAAMG || kind == FEM_PC_BSAAMG), smoother(p.smoother), npre(p.npre),
          npost(p.npost), omega(p.omega) {
        const size_t L = H.A.size();
        levels = (int)L;
        for (size_t k = 0; k < L; ++k) {
            auto D = std::make_unique<AmgLevelDev>();
            HCsr& A = H.A[k];
            D->n = A.n;
            D->omega = k < H.omega.size() ? H.omega[k] : p.omega;
            lvl_rows.push_back(A.n);
            lvl_nnz.push_back(A.v.size());
            D->A.set_host(A.n, A.m, std::move(A.rp), std::move(A.ci), std::move(A.v), s);
            D->A.stream = s;
            if (k + 1 < L) {
                HCsr& P = H.P[k];
                D->nc = P.m;
                HCsr R = transpose(P);
                D->P.set_host(P.n, P.m, std::move(P.rp), std::move(P.ci), std::move(P.v), s);
                D->R.set_host(R.n, R.m, std::move(R.rp), std::move(R.ci), std::move(R.v), s);
                D->d.upload(H.d[k], s);
                D->u.alloc(D->n);
                D->u2.alloc(D->n);
                D->res.alloc(D->n);
                if (smoother == FEM_SMOOTH_SGS) {
                    D->ncol = (int)H.cptr[k].size() - 1;
                    D->h_cptr = H.cptr[k];
                    D->cptr.upload(H.cptr[k], s);
                    D->crows.upload(H.crows[k], s);
                    D->diag.upload(H.diag[k], s);
                }
            } else {
                D->u.alloc(D->n);
            }
            if (k > 0) D->f.alloc(D->n);
            lv.push_back(std::move(D));
        }
        nco = lv.back()->n;
        coarse.upload(H.coarse, s);
        FEM_HIP(hipStreamSynchronize(s));
    }
    int kind() const override { return kind_; }
    uint32_t size() const override { return lv[0]->n; }
    bool has_fp32() const override { return fp32; }
    void ensure_fp32(hipStream_t s) override {
        if (fp32) return;
        for (size_t k = 0; k < lv.size(); ++k) {
            auto& D = *lv[k];
            D.A.ensure_fp32(s);
            if (k + 1 < lv.size()) {
                D.P.ensure_fp32(s);
                D.R.ensure_fp32(s);
                D.d32.alloc(D.n);
                vec_convert<float, double>(D.d32.get(), D.d.get(), 1.0, D.n, s);
                if (D.diag.get()) {
                    D.diag32.alloc(D.n);
                    vec_convert<float, double>(D.diag32.get(), D.diag.get(), 1.0, D.n, s);
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
    template <typename T>
    T* jacobi(AmgLevelDev& D, const T* f, const T* d, T* cur, T* a, T* b, int sweeps, hipStream_t s, Gate g) {
        const T w = (T)D.omega;
        for (int k = 0; k < sweeps; ++k) {
            T* out = (cur == a) ? b : a;
            if (!cur) {
                if (inv_diag) hipLaunchKernelGGL((k_jacobi0<T, true>), g1(D.n), dim3(256), 0, s, f, d, out, D.n, w, g);
                else hipLaunchKernelGGL((k_jacobi0<T, false>), g1(D.n), dim3(256), 0, s, f, d, out, D.n, w, g);
            } else {
                SpmvArgs<T> ar;
                ar.f = f;
                ar.u = cur;
                ar.d = d;
                ar.w = D.omega;
                csr_spmv<T>(D.A, cur, out, inv_diag ? EPI_JACOBI_INV : EPI_JACOBI, ar, s, g);
            }
            cur = out;
        }
        return cur;
    }

    template <typename T>
    void gs(AmgLevelDev& D, const T* f, const T* diag, T* u, int sweeps, bool reverse, hipStream_t s, Gate g) {
        if (sweeps <= 0) return;
        const T w = (T)D.omega;
        const T* vals;
        if constexpr (std::is_same<T, double>::value) vals = D.A.v.get();
        else vals = D.A.v32.get();
        if (D.n <= 8192) {
            hipLaunchKernelGGL((k_gs_small<T>), dim3(1), dim3(1024), 0, s, D.ncol, D.cptr.get(), D.crows.get(),
                               D.A.rp.get(), D.A.ci.get(), vals, diag, f, u, w, sweeps, reverse ? 1 : 0, g);
            return;
        }
        for (int sw = 0; sw < sweeps; ++sw)
            for (int cc = 0; cc < D.ncol; ++cc) {
                const int c = reverse ? D.ncol - 1 - cc : cc;
                const uint32_t lo = D.h_cptr[c], hi = D.h_cptr[c + 1];
                if (lo >= hi) continue;
                hipLaunchKernelGGL((k_gs_colour<T>), g1(hi - lo), dim3(256), 0, s, lo, hi, D.crows.get(),
                                   D.A.rp.get(), D.A.ci.get(), vals, diag, f, u, w, g);
            }
    }

    template <typename T>
    void vcycle(const T* r, T* z, hipStream_t s, Gate g) {
        const size_t L = lv.size();
        std::vector<T*> ucur(L, nullptr), xa(L, nullptr), xb(L, nullptr);
        std::vector<const T*> fk(L, nullptr);
        fk[0] = r;
        for (size_t k = 0; k + 1 < L; ++k) {
            AmgLevelDev& D = *lv[k];
            Bufs<T> B = bufs<T>(D);
            const T* f = fk[k];
            T* u;
            if (smoother == FEM_SMOOTH_SGS) {
                u = (k == 0) ? z : B.u;
                hipLaunchKernelGGL((k_zero<T>), g1(D.n), dim3(256), 0, s, u, D.n, g);
                gs<T>(D, f, B.diag, u, npre, false, s, g);
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
                    u = jacobi<T>(D, f, B.d, nullptr, X, Y, npre, s, g);
                } else {
                    hipLaunchKernelGGL((k_zero<T>), g1(D.n), dim3(256), 0, s, X, D.n, g);
                    u = X;
                }
            }
            ucur[k] = u;
            SpmvArgs<T> ar;
            ar.f = f;
            csr_spmv<T>(D.A, u, B.res, EPI_RESID, ar, s, g);
            AmgLevelDev& C = *lv[k + 1];
            Bufs<T> BC = bufs<T>(C);
            csr_spmv<T>(D.R, B.res, BC.f, EPI_STORE, {}, s, g);
            fk[k + 1] = BC.f;
        }
        // coarsest level: u = C f (dense, float64 accumulation)
        {
            AmgLevelDev& C = *lv[L - 1];
            Bufs<T> BC = bufs<T>(C);
            T* u = (L == 1) ? z : BC.u;
            const size_t sh = (size_t)nco * sizeof(double);
            hipLaunchKernelGGL((k_coarse<T>), dim3((nco + 255) / 256), dim3(256), sh, s, coarse.get(), nco, fk[L - 1],
                               u, g);
            ucur[L - 1] = u;
        }
        for (int64_t k = (int64_t)L - 2; k >= 0; --k) {
            AmgLevelDev& D = *lv[k];
            Bufs<T> B = bufs<T>(D);
            T* u = ucur[k];
            SpmvArgs<T> ar;
            ar.u = u;
            csr_spmv<T>(D.P, ucur[k + 1], u, EPI_ADD, ar, s, g);
            if (smoother == FEM_SMOOTH_SGS) gs<T>(D, fk[k], B.diag, u, npost, true, s, g);
            else u = jacobi<T>(D, fk[k], B.d, u, xa[k], xb[k], npost, s, g);
            ucur[k] = u;
        }
        FEM_HIP(hipGetLastError());
    }

    void apply(const double* r, double* z, hipStream_t s, Gate g) override { vcycle<double>(r, z, s, g); }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override { vcycle<float>(r, z, s, g); }
};

std::unique_ptr<Precond> make_amg_cf(CsrMat& A, const fem_pc_params& p, hipStream_t s) {
    if (A.n != A.m) fail(FEM_ERR_INVALID_ARGUMENT, "AMG needs a square matrix");
    A.sync_host();
    AmgHierarchy H = build_cf_hierarchy(A, p);
    return std::unique_ptr<Precond>(new PcAmg(FEM_PC_AMG, std::move(H), p, s));
}

std::unique_ptr<Precond> make_saamg(CsrMat& A, const fem_pc_params& p, hipStream_t s) {
    if (A.n != A.m) fail(FEM_ERR_INVALID_ARGUMENT, "Block AMG needs a square matrix");
    A.sync_host();
    AmgHierarchy H = build_sa_hierarchy(A, p);
    return std::unique_ptr<Precond>(new PcAmg(FEM_PC_SAAMG, std::move(H), p, s));
}

}  // namespace fem

// bsr.hip — block CSR matrices (block::BlockMatrix<double,BS>) on the device.
//   BlockMatrix, matvec, transpose, condense   ode-solver/include/math/blockMatrix.hpp:121-315
//   make_block_ilu0_preconditioner             blockMatrix.hpp:454-623
//   "blockjacobi" (BlockMatrix::gmres)         blockMatrix.hpp:714-735
//   block-Jacobi of the BlockMatrix tests      ode-solver/test/test_BlockMatrix_mpi.cpp:132-155
//
// Layout as the reference: row-major BS x BS blocks stored contiguously per
// nonzero block, node-major vectors (component c of block row b at b*BS + c).
// matvec: one thread per scalar row (b, r); for every block of row b the
// inner product over c is formed from 0.0 and then added to y[b*BS + r], the
// reference's operation order. Block ILU(0) is factored on the device level by
// level (one thread per block row, the reference's per-row arithmetic) and
// applied with level-scheduled block sweeps.
#include <algorithm>
#include <cmath>

#include "amg.hpp"
#include "handles.hpp"
#include "precond.hpp"

namespace fem {

struct BsrMat : LinOp {
    int bs = 1;
    uint32_t nbr = 0, nbc = 0;
    uint64_t nnzb = 0;
    DevBuf<uint32_t> rp, ci;
    DevBuf<double> v;
    DevBuf<float> v32;
    std::vector<uint32_t> h_rp, h_ci;
    std::vector<double> h_v;
    hipStream_t stream = nullptr;

    uint32_t rows() const override { return nbr * (uint32_t)bs; }
    uint32_t cols() const override { return nbc * (uint32_t)bs; }
    void apply(const double* x, double* y, hipStream_t s, Gate g = {}) const override;
    void apply32(const float* x, float* y, hipStream_t s, Gate g = {}) const override;
    void ensure_fp32(hipStream_t s) override {
        if (v32.get() || !v.n) return;
        v32.alloc(v.n);
        vec_convert<float, double>(v32.get(), v.get(), 1.0, v.n, s);
    }
    uint64_t bytes_per_apply(bool fp32) const override {
        const uint64_t sv = fp32 ? 4 : 8;
        return nnzb * ((uint64_t)bs * bs * sv + 4) + 4ull * (nbr + 1) + (uint64_t)(nbr + nbc) * bs * sv;
    }
    void set_host(int bs_, uint32_t nbr_, uint32_t nbc_, std::vector<uint32_t> rp_, std::vector<uint32_t> ci_,
                  std::vector<double> v_, hipStream_t s) {
        bs = bs_;
        nbr = nbr_;
        nbc = nbc_;
        nnzb = ci_.size();
        stream = s;
        rp.upload(rp_, s);
        ci.upload(ci_.empty() ? std::vector<uint32_t>{0} : ci_, s);
        v.upload(v_.empty() ? std::vector<double>{0.0} : v_, s);
        v.n = v_.size();
        h_rp = std::move(rp_);
        h_ci = std::move(ci_);
        h_v = std::move(v_);
        v32.release();
        FEM_HIP(hipStreamSynchronize(s));
    }
};

template <typename T, int BS>
__global__ __launch_bounds__(256) void k_bsr_mv(uint32_t nbr, const uint32_t* __restrict__ rp,
                                                const uint32_t* __restrict__ ci, const T* __restrict__ v,
                                                const T* __restrict__ x, T* __restrict__ y, Gate g) {
    if (g.closed()) return;
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nbr * BS) return;
    const uint32_t b = row / BS, r = row % BS;
    T yr = (T)0;
    for (uint32_t p = rp[b]; p < rp[b + 1]; ++p) {
        const T* B = v + ((size_t)p * BS + r) * BS;
        const T* xj = x + (size_t)ci[p] * BS;
        T acc = (T)0;
#pragma unroll
        for (int c = 0; c < BS; ++c) acc += B[c] * xj[c];
        yr += acc;
    }
    y[row] = yr;
}

template <typename T>
static void bsr_mv(const BsrMat& A, const T* vals, const T* x, T* y, hipStream_t s, Gate g) {
    const uint32_t n = A.nbr * A.bs;
    if (!n) return;
    const dim3 G((n + 255) / 256), B(256);
    switch (A.bs) {
        case 1: hipLaunchKernelGGL((k_bsr_mv<T, 1>), G, B, 0, s, A.nbr, A.rp.get(), A.ci.get(), vals, x, y, g); break;
        case 2: hipLaunchKernelGGL((k_bsr_mv<T, 2>), G, B, 0, s, A.nbr, A.rp.get(), A.ci.get(), vals, x, y, g); break;
        case 3: hipLaunchKernelGGL((k_bsr_mv<T, 3>), G, B, 0, s, A.nbr, A.rp.get(), A.ci.get(), vals, x, y, g); break;
        case 4: hipLaunchKernelGGL((k_bsr_mv<T, 4>), G, B, 0, s, A.nbr, A.rp.get(), A.ci.get(), vals, x, y, g); break;
        default: fail(FEM_ERR_INVALID_ARGUMENT, "block size must be 1..4");
    }
    FEM_HIP(hipGetLastError());
}

void BsrMat::apply(const double* x, double* y, hipStream_t s, Gate g) const { bsr_mv<double>(*this, v.get(), x, y, s, g); }
void BsrMat::apply32(const float* x, float* y, hipStream_t s, Gate g) const {
    if (!v32.get()) fail(FEM_ERR_RUNTIME, "float32 copy of the block matrix was not built");
    bsr_mv<float>(*this, v32.get(), x, y, s, g);
}

// ---------------------------------------------------------------------------
// small dense block helpers (device); the reference's operation order
// ---------------------------------------------------------------------------
// Gauss elimination with partial pivoting on [M | I] then back substitution
// (blk_invert, blockMatrix.hpp:70-100); false on a pivot below 1e-300
template <int BS>
__device__ bool d_blk_invert(const double* M, double* Mi) {
    double a[BS][2 * BS];
    for (int i = 0; i < BS; ++i)
        for (int j = 0; j < BS; ++j) { a[i][j] = M[i * BS + j]; a[i][BS + j] = (i == j) ? 1.0 : 0.0; }
    for (int k = 0; k < BS; ++k) {
        int pv = k;
        for (int i = k + 1; i < BS; ++i)
            if (fabs(a[i][k]) > fabs(a[pv][k])) pv = i;
        if (fabs(a[pv][k]) < 1e-300) return false;
        if (pv != k)
            for (int j = 0; j < 2 * BS; ++j) { const double t = a[k][j]; a[k][j] = a[pv][j]; a[pv][j] = t; }
        for (int i = k + 1; i < BS; ++i) {
            const double f = a[i][k] / a[k][k];
            for (int j = k; j < 2 * BS; ++j) a[i][j] -= f * a[k][j];
        }
    }
    for (int i = BS - 1; i >= 0; --i) {
        for (int j = i + 1; j < BS; ++j)
            for (int c = 0; c < BS; ++c) a[i][BS + c] -= a[i][j] * a[j][BS + c];
        for (int c = 0; c < BS; ++c) a[i][BS + c] /= a[i][i];
    }
    for (int i = 0; i < BS; ++i)
        for (int j = 0; j < BS; ++j) Mi[i * BS + j] = a[i][BS + j];
    return true;
}

// C -= A*B  (blk_matmat_accum, blockMatrix.hpp:37-46)
template <int BS>
__device__ __forceinline__ void d_mm_sub(const double* A, const double* B, double* C) {
    for (int r = 0; r < BS; ++r)
        for (int c = 0; c < BS; ++c) {
            double s = 0.0;
            for (int k = 0; k < BS; ++k) s += A[r * BS + k] * B[k * BS + c];
            C[r * BS + c] -= s;
        }
}
// C := C*B  (blk_right_mul, blockMatrix.hpp:48-59)
template <int BS>
__device__ __forceinline__ void d_right_mul(double* C, const double* B) {
    double t[BS * BS];
    for (int r = 0; r < BS; ++r)
        for (int c = 0; c < BS; ++c) {
            double s = 0.0;
            for (int k = 0; k < BS; ++k) s += C[r * BS + k] * B[k * BS + c];
            t[r * BS + c] = s;
        }
    for (int i = 0; i < BS * BS; ++i) C[i] = t[i];
}

// ---------------------------------------------------------------------------
// block-Jacobi (inverse diagonal blocks cached at setup)
// ---------------------------------------------------------------------------
template <int BS>
__global__ void k_bjac_setup(uint32_t nb, const double* v, const uint32_t* dptr, double* inv, int* bad) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    if (!d_blk_invert<BS>(v + (size_t)dptr[b] * BS * BS, inv + (size_t)b * BS * BS)) atomicMin(bad, (int)b);
}

template <typename T, int BS>
__global__ void k_bjac_apply(uint32_t nb, const T* __restrict__ inv, const T* __restrict__ r, T* __restrict__ z,
                             Gate g) {
    if (g.closed()) return;
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nb * BS) return;
    const uint32_t b = row / BS, rr = row % BS;
    const T* M = inv + (size_t)b * BS * BS + rr * BS;
    const T* x = r + (size_t)b * BS;
    T a = (T)0;
#pragma unroll
    for (int c = 0; c < BS; ++c) a += M[c] * x[c];
    z[row] = a;
}

template <int BS>
struct PcBJacobi : Precond {
    uint32_t nb;
    DevBuf<double> inv;
    DevBuf<float> inv32;
    PcBJacobi(const BsrMat& A, hipStream_t s) : nb(A.nbr) {
        if (A.nbr != A.nbc) fail(FEM_ERR_INVALID_ARGUMENT, "block-Jacobi needs a square block matrix");
        std::vector<uint32_t> dp(nb, UINT32_MAX);
        for (uint32_t b = 0; b < nb; ++b)
            for (uint32_t p = A.h_rp[b]; p < A.h_rp[b + 1]; ++p)
                if (A.h_ci[p] == b) { dp[b] = p; break; }
        for (uint32_t b = 0; b < nb; ++b)
            if (dp[b] == UINT32_MAX) fail(FEM_ERR_RUNTIME, "block-Jacobi: block row " + std::to_string(b) + " has no diagonal block");
        DevBuf<uint32_t> ddp;
        ddp.upload(dp, s);
        inv.alloc((size_t)nb * BS * BS);
        DevBuf<int> bad(1);
        const int big = 0x7fffffff;
        FEM_HIP(hipMemcpyAsync(bad.get(), &big, sizeof(int), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL((k_bjac_setup<BS>), dim3((nb + 255) / 256), dim3(256), 0, s, nb, A.v.get(), ddp.get(),
                           inv.get(), bad.get());
        FEM_HIP(hipGetLastError());
        int hb = 0;
        FEM_HIP(hipMemcpyAsync(&hb, bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        if (hb != big) fail(FEM_ERR_RUNTIME, "block-Jacobi: singular diagonal block " + std::to_string(hb));
    }
    int kind() const override { return FEM_PC_BJACOBI; }
    uint32_t size() const override { return nb * BS; }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override {
        hipLaunchKernelGGL((k_bjac_apply<double, BS>), dim3((nb * BS + 255) / 256), dim3(256), 0, s, nb, inv.get(), r, z, g);
        FEM_HIP(hipGetLastError());
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override {
        hipLaunchKernelGGL((k_bjac_apply<float, BS>), dim3((nb * BS + 255) / 256), dim3(256), 0, s, nb, inv32.get(), r, z,
                           g);
        FEM_HIP(hipGetLastError());
    }
    bool has_fp32() const override { return inv32.get() != nullptr; }
    void ensure_fp32(hipStream_t s) override {
        if (inv32.get()) return;
        inv32.alloc(inv.n);
        vec_convert<float, double>(inv32.get(), inv.get(), 1.0, inv.n, s);
    }
};

// ---------------------------------------------------------------------------
// block ILU(0)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t d_find(const uint32_t* rp, const uint32_t* ci, uint32_t row, uint32_t col) {
    uint32_t lo = rp[row], hi = rp[row + 1];
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (ci[mid] == col) return mid;
        if (ci[mid] < col) lo = mid + 1;
        else hi = mid;
    }
    return UINT32_MAX;
}

template <int BS>
__global__ void k_bilu_level(uint32_t a, uint32_t b, const uint32_t* __restrict__ rows, const uint32_t* __restrict__ Lp,
                             const uint32_t* __restrict__ Lc, double* Lv, const uint32_t* __restrict__ Up,
                             const uint32_t* __restrict__ Uc, double* Uv, const uint32_t* __restrict__ Ud, double* inv,
                             int* bad) {
    const uint32_t t = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b) return;
    constexpr int BV = BS * BS;
    const uint32_t I = rows[t];
    const uint32_t Ls = Lp[I], Le = Lp[I + 1], Us = Up[I], Ue = Up[I + 1];
    for (uint32_t pL = Ls; pL < Le; ++pL) {
        const uint32_t K = Lc[pL];
        double* LIK = Lv + (size_t)pL * BV;
        for (uint32_t q = Ls; q < pL; ++q) {
            const uint32_t pJK = d_find(Up, Uc, Lc[q], K);
            if (pJK != UINT32_MAX) d_mm_sub<BS>(Lv + (size_t)q * BV, Uv + (size_t)pJK * BV, LIK);
        }
        d_right_mul<BS>(LIK, inv + (size_t)K * BV);
        uint32_t pi = Us, pk = Up[K];
        const uint32_t ek = Up[K + 1];
        while (pi < Ue && pk < ek) {
            if (Uc[pi] == Uc[pk]) { d_mm_sub<BS>(LIK, Uv + (size_t)pk * BV, Uv + (size_t)pi * BV); ++pi; ++pk; }
            else if (Uc[pi] < Uc[pk]) ++pi;
            else ++pk;
        }
    }
    if (!d_blk_invert<BS>(Uv + (size_t)Ud[I] * BV, inv + (size_t)I * BV)) atomicMin(bad, (int)I);
}

template <typename T, int BS>
__global__ void k_bilu_fw(uint32_t a, uint32_t b, const uint32_t* __restrict__ rows, const uint32_t* __restrict__ Lp,
                          const uint32_t* __restrict__ Lc, const T* __restrict__ Lv, const T* __restrict__ rhs, T* y,
                          Gate g) {
    if (g.closed()) return;
    const uint32_t t = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b) return;
    const uint32_t I = rows[t];
    T yi[BS];
#pragma unroll
    for (int c = 0; c < BS; ++c) yi[c] = rhs[(size_t)I * BS + c];
    for (uint32_t p = Lp[I]; p < Lp[I + 1]; ++p) {
        const T* L = Lv + (size_t)p * BS * BS;
        const T* yj = y + (size_t)Lc[p] * BS;
        T yjv[BS];
#pragma unroll
        for (int c = 0; c < BS; ++c) yjv[c] = yj[c];
#pragma unroll
        for (int r = 0; r < BS; ++r) {
            T s = (T)0;
#pragma unroll
            for (int c = 0; c < BS; ++c) s += L[r * BS + c] * yjv[c];
            yi[r] -= s;
        }
    }
#pragma unroll
    for (int c = 0; c < BS; ++c) y[(size_t)I * BS + c] = yi[c];
}

template <typename T, int BS>
__global__ void k_bilu_bw(uint32_t a, uint32_t b, const uint32_t* __restrict__ rows, const uint32_t* __restrict__ Up,
                          const uint32_t* __restrict__ Uc, const T* __restrict__ Uv, const T* __restrict__ inv,
                          const T* __restrict__ y, T* z, Gate g) {
    if (g.closed()) return;
    const uint32_t t = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b) return;
    const uint32_t I = rows[t];
    T tmp[BS];
#pragma unroll
    for (int c = 0; c < BS; ++c) tmp[c] = y[(size_t)I * BS + c];
    for (uint32_t p = Up[I]; p < Up[I + 1]; ++p) {
        const uint32_t J = Uc[p];
        if (J == I) continue;
        const T* U = Uv + (size_t)p * BS * BS;
        const T* zj = z + (size_t)J * BS;
        T zjv[BS];
#pragma unroll
        for (int c = 0; c < BS; ++c) zjv[c] = zj[c];
#pragma unroll
        for (int r = 0; r < BS; ++r) {
            T s = (T)0;
#pragma unroll
            for (int c = 0; c < BS; ++c) s += U[r * BS + c] * zjv[c];
            tmp[r] -= s;
        }
    }
    const T* M = inv + (size_t)I * BS * BS;
#pragma unroll
    for (int r = 0; r < BS; ++r) {
        T s = (T)0;
#pragma unroll
        for (int c = 0; c < BS; ++c) s += M[r * BS + c] * tmp[c];
        z[(size_t)I * BS + r] = s;
    }
}

template <int BS>
struct PcBILU0 : Precond {
    static constexpr int BV = BS * BS;
    uint32_t nb;
    std::vector<uint32_t> flp, blp;  // level pointers (forward, backward)
    DevBuf<uint32_t> frows, brows, Lp, Lc, Up, Uc;
    DevBuf<double> Lv, Uv, inv, y;
    DevBuf<float> Lv32, Uv32, inv32, y32;

    static std::vector<uint32_t> order_levels(uint32_t n, const std::vector<uint32_t>& rp,
                                              const std::vector<uint32_t>& ci, bool fwd, std::vector<uint32_t>& lp,
                                              bool skip_diag) {
        std::vector<uint32_t> lev(n, 0);
        uint32_t mx = 0;
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t i = fwd ? k : n - 1 - k;
            uint32_t l = 0;
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
                if (skip_diag && ci[p] == i) continue;
                l = std::max(l, lev[ci[p]] + 1);
            }
            lev[i] = l;
            mx = std::max(mx, l);
        }
        lp.assign(n ? mx + 2 : 1, 0);
        for (uint32_t i = 0; i < n; ++i) lp[lev[i] + 1]++;
        for (size_t l = 0; l + 1 < lp.size(); ++l) lp[l + 1] += lp[l];
        std::vector<uint32_t> rows(n), off(lp.begin(), lp.end() - 1);
        for (uint32_t i = 0; i < n; ++i) rows[off[lev[i]]++] = i;
        return rows;
    }

    PcBILU0(const BsrMat& A, hipStream_t s) : nb(A.nbr) {
        if (A.nbr != A.nbc) fail(FEM_ERR_INVALID_ARGUMENT, "Block-ILU(0) needs a square block matrix");
        const auto& rp = A.h_rp;
        const auto& ci = A.h_ci;
        std::vector<uint32_t> lp(nb + 1, 0), up(nb + 1, 0), lc, uc, lmap, umap, ud(nb, UINT32_MAX);
        for (uint32_t I = 0; I < nb; ++I) {
            for (uint32_t p = rp[I]; p < rp[I + 1]; ++p) {
                const uint32_t J = ci[p];
                if (J < I) { lc.push_back(J); lmap.push_back(p); }
                else {
                    if (J == I) ud[I] = (uint32_t)uc.size();
                    uc.push_back(J);
                    umap.push_back(p);
                }
            }
            lp[I + 1] = (uint32_t)lc.size();
            up[I + 1] = (uint32_t)uc.size();
        }
        for (uint32_t I = 0; I < nb; ++I)
            if (ud[I] == UINT32_MAX)
                fail(FEM_ERR_RUNTIME, "Block-ILU(0) failed: block row " + std::to_string(I) + " has no diagonal block");
        std::vector<double> hl(std::max<size_t>(lc.size(), 1) * BV), hu(std::max<size_t>(uc.size(), 1) * BV);
        for (size_t k = 0; k < lc.size(); ++k) std::copy_n(&A.h_v[(size_t)lmap[k] * BV], BV, &hl[k * BV]);
        for (size_t k = 0; k < uc.size(); ++k) std::copy_n(&A.h_v[(size_t)umap[k] * BV], BV, &hu[k * BV]);
        const auto fr = order_levels(nb, lp, lc, true, flp, false);
        const auto br = order_levels(nb, up, uc, false, blp, true);
        levels = (int)flp.size() - 1;
        frows.upload(fr, s);
        brows.upload(br, s);
        Lp.upload(lp, s);
        Up.upload(up, s);
        Lc.upload(lc.empty() ? std::vector<uint32_t>{0} : lc, s);
        Uc.upload(uc, s);
        Lv.upload(hl, s);
        Uv.upload(hu, s);
        inv.alloc((size_t)nb * BV);
        DevBuf<uint32_t> dud;
        dud.upload(ud, s);
        DevBuf<int> bad(1);
        const int big = 0x7fffffff;
        FEM_HIP(hipMemcpyAsync(bad.get(), &big, sizeof(int), hipMemcpyHostToDevice, s));
        for (size_t l = 0; l + 1 < flp.size(); ++l) {
            const uint32_t a = flp[l], b = flp[l + 1];
            if (a >= b) continue;
            hipLaunchKernelGGL((k_bilu_level<BS>), dim3((b - a + 127) / 128), dim3(128), 0, s, a, b, frows.get(),
                               Lp.get(), Lc.get(), Lv.get(), Up.get(), Uc.get(), Uv.get(), dud.get(), inv.get(),
                               bad.get());
        }
        FEM_HIP(hipGetLastError());
        int hb = 0;
        FEM_HIP(hipMemcpyAsync(&hb, bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        if (hb != big)
            fail(FEM_ERR_RUNTIME, "Block-ILU(0) breakdown: singular diagonal block U_" + std::to_string(hb));
        y.alloc((size_t)nb * BS);
    }
    int kind() const override { return FEM_PC_BILU0; }
    uint32_t size() const override { return nb * BS; }
    template <typename T>
    void run(const T* r, T* z, const T* lv, const T* uv, const T* iv, T* yy, hipStream_t s, Gate g) {
        for (size_t l = 0; l + 1 < flp.size(); ++l) {
            const uint32_t a = flp[l], b = flp[l + 1];
            if (a < b)
                hipLaunchKernelGGL((k_bilu_fw<T, BS>), dim3((b - a + 255) / 256), dim3(256), 0, s, a, b, frows.get(),
                                   Lp.get(), Lc.get(), lv, r, yy, g);
        }
        for (size_t l = 0; l + 1 < blp.size(); ++l) {
            const uint32_t a = blp[l], b = blp[l + 1];
            if (a < b)
                hipLaunchKernelGGL((k_bilu_bw<T, BS>), dim3((b - a + 255) / 256), dim3(256), 0, s, a, b, brows.get(),
                                   Up.get(), Uc.get(), uv, iv, yy, z, g);
        }
        FEM_HIP(hipGetLastError());
    }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override {
        run<double>(r, z, Lv.get(), Uv.get(), inv.get(), y.get(), s, g);
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override {
        run<float>(r, z, Lv32.get(), Uv32.get(), inv32.get(), y32.get(), s, g);
    }
    bool has_fp32() const override { return y32.get() != nullptr; }
    void ensure_fp32(hipStream_t s) override {
        if (y32.get()) return;
        auto cv = [&](DevBuf<double>& a, DevBuf<float>& b) {
            b.alloc(a.n);
            vec_convert<float, double>(b.get(), a.get(), 1.0, a.n, s);
        };
        cv(Lv, Lv32);
        cv(Uv, Uv32);
        cv(inv, inv32);
        y32.alloc((size_t)nb * BS);
    }
};

template <template <int> class PC>
static std::unique_ptr<Precond> make_bs(const BsrMat& A, hipStream_t s) {
    switch (A.bs) {
        case 1: return std::unique_ptr<Precond>(new PC<1>(A, s));
        case 2: return std::unique_ptr<Precond>(new PC<2>(A, s));
        case 3: return std::unique_ptr<Precond>(new PC<3>(A, s));
        case 4: return std::unique_ptr<Precond>(new PC<4>(A, s));
        default: fail(FEM_ERR_INVALID_ARGUMENT, "block size must be 1..4");
    }
}

struct PcIdentity : Precond {
    uint32_t n;
    explicit PcIdentity(uint32_t n_) : n(n_) {}
    int kind() const override { return FEM_PC_NONE; }
    uint32_t size() const override { return n; }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override { launch_copy(r, z, n, s, g); }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override { launch_copy(r, z, n, s, g); }
    bool has_fp32() const override { return true; }
};

// condense (blockMatrix.hpp:269-315): block column j/BS, duplicates summed in
// scalar CSR order, block columns sorted and unique, explicit zeros kept
std::unique_ptr<Precond> make_block_saamg(const BsrMat& A, const fem_pc_params& p, hipStream_t s) {
    if (A.bs != 1) fail(FEM_ERR_INVALID_ARGUMENT, "block SA-AMG on the device supports BS = 1 (scalar condensation)");
    CsrMat C;
    C.set_host(A.nbr, A.nbc, A.h_rp, A.h_ci, A.h_v, s);
    return make_saamg(C, p, s);
}

}  // namespace fem

using namespace fem;

struct fem_bsr_s {
    hipStream_t stream = nullptr;
    BsrMat A;
    DevBuf<double> xs, ys;
    ~fem_bsr_s() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

int fem_bsr_create(int bs, uint32_t nbr, uint32_t nbc, const uint32_t* row_ptr, const uint32_t* col_idx,
                   const double* values, fem_bsr* out) {
    return api_guard([&] {
        if (bs < 1 || bs > 4) fail(FEM_ERR_INVALID_ARGUMENT, "block size must be 1..4");
        if (!row_ptr || row_ptr[0] != 0) fail(FEM_ERR_INVALID_ARGUMENT, "row_ptr must start at 0");
        for (uint32_t i = 0; i < nbr; ++i)
            if (row_ptr[i + 1] < row_ptr[i]) fail(FEM_ERR_INVALID_ARGUMENT, "row_ptr must be non-decreasing");
        const uint32_t nz = row_ptr[nbr];
        for (uint32_t k = 0; k < nz; ++k)
            if (col_idx[k] >= nbc) fail(FEM_ERR_OUT_OF_RANGE, "block column index out of range");
        auto h = std::make_unique<fem_bsr_s>();
        h->stream = new_stream();
        h->A.set_host(bs, nbr, nbc, std::vector<uint32_t>(row_ptr, row_ptr + nbr + 1),
                      std::vector<uint32_t>(col_idx, col_idx + nz),
                      std::vector<double>(values, values + (size_t)nz * bs * bs), h->stream);
        *out = h.release();
    });
}

int fem_bsr_condense(fem_csr A, int bs, uint32_t nbr, uint32_t nbc, fem_bsr* out) {
    return api_guard([&] {
        if (bs < 1 || bs > 4) fail(FEM_ERR_INVALID_ARGUMENT, "block size must be 1..4");
        std::vector<uint32_t> rp, ci;
        std::vector<double> v;
        condense_device(A->A, bs, nbr, nbc, rp, ci, v, A->stream);
        auto h = std::make_unique<fem_bsr_s>();
        h->stream = new_stream();
        h->A.set_host(bs, nbr, nbc, std::move(rp), std::move(ci), std::move(v), h->stream);
        *out = h.release();
    });
}

int fem_bsr_info(fem_bsr A, int* bs, uint32_t* nbr, uint32_t* nbc, uint64_t* nnzb) {
    return api_guard([&] {
        if (bs) *bs = A->A.bs;
        if (nbr) *nbr = A->A.nbr;
        if (nbc) *nbc = A->A.nbc;
        if (nnzb) *nnzb = A->A.nnzb;
    });
}

int fem_bsr_export(fem_bsr A, uint32_t* row_ptr, uint32_t* col_idx, double* values) {
    return api_guard([&] {
        if (row_ptr) std::copy(A->A.h_rp.begin(), A->A.h_rp.end(), row_ptr);
        if (col_idx) std::copy(A->A.h_ci.begin(), A->A.h_ci.end(), col_idx);
        if (values) std::copy(A->A.h_v.begin(), A->A.h_v.end(), values);
    });
}

// transpose (blockMatrix.hpp:237-267): block pattern transposed by counting
// sort, each block transposed
int fem_bsr_transpose(fem_bsr A, fem_bsr* out) {
    return api_guard([&] {
        const BsrMat& M = A->A;
        const int bs = M.bs, BV = bs * bs;
        std::vector<uint32_t> rp(M.nbc + 1, 0), ci(M.nnzb);
        std::vector<double> v(M.h_v.size());
        for (uint64_t p = 0; p < M.nnzb; ++p) rp[M.h_ci[p] + 1]++;
        for (uint32_t j = 0; j < M.nbc; ++j) rp[j + 1] += rp[j];
        std::vector<uint32_t> off(rp.begin(), rp.end() - 1);
        for (uint32_t i = 0; i < M.nbr; ++i)
            for (uint32_t p = M.h_rp[i]; p < M.h_rp[i + 1]; ++p) {
                const uint32_t q = off[M.h_ci[p]]++;
                ci[q] = i;
                for (int r = 0; r < bs; ++r)
                    for (int c = 0; c < bs; ++c) v[(size_t)q * BV + r * bs + c] = M.h_v[(size_t)p * BV + c * bs + r];
            }
        auto h = std::make_unique<fem_bsr_s>();
        h->stream = new_stream();
        h->A.set_host(bs, M.nbc, M.nbr, std::move(rp), std::move(ci), std::move(v), h->stream);
        *out = h.release();
    });
}

int fem_bsr_matvec(fem_bsr A, const double* x, double* y) {
    return api_guard([&] {
        const uint32_t n = A->A.rows(), m = A->A.cols();
        if (A->xs.n < m) A->xs.alloc(std::max<uint32_t>(m, 1));
        if (A->ys.n < n) A->ys.alloc(std::max<uint32_t>(n, 1));
        A->xs.upload(x, m, A->stream);
        A->A.apply(A->xs.get(), A->ys.get(), A->stream);
        A->ys.download(y, n, A->stream);
        FEM_HIP(hipStreamSynchronize(A->stream));
    });
}

int fem_bsr_matvec_device(fem_bsr A, const void* x, void* y, int fp32) {
    return api_guard([&] {
        if (fp32) {
            A->A.ensure_fp32(A->stream);
            A->A.apply32((const float*)x, (float*)y, A->stream);
        } else {
            A->A.apply((const double*)x, (double*)y, A->stream);
        }
        FEM_HIP(hipStreamSynchronize(A->stream));
    });
}

int fem_bsr_precond_create(fem_bsr A, int kind, const fem_pc_params* params, fem_pc* out) {
    return api_guard([&] {
        fem_pc_params p;
        if (params) p = *params;
        else fem_pc_params_default(kind, &p);
        auto h = std::make_unique<fem_pc_s>();
        h->stream = A->stream;
        Timer tm;
        switch (kind) {
            case FEM_PC_NONE: h->pc.reset(new PcIdentity(A->A.rows())); break;
            case FEM_PC_BJACOBI: h->pc = make_bs<PcBJacobi>(A->A, A->stream); break;
            case FEM_PC_BILU0: h->pc = make_bs<PcBILU0>(A->A, A->stream); break;
            case FEM_PC_BSAAMG: h->pc = make_block_saamg(A->A, p, A->stream); break;
            default: fail(FEM_ERR_INVALID_ARGUMENT, "unknown block preconditioner kind " + std::to_string(kind));
        }
        if (p.fp32) h->pc->ensure_fp32(A->stream);
        FEM_HIP(hipStreamSynchronize(A->stream));
        h->pc->setup_ms = tm.ms();
        *out = h.release();
    });
}

int fem_bsr_solve(fem_bsr A, fem_pc M, const double* b, double* x, const fem_solve_opts* opts, fem_solve_stats* stats) {
    return api_guard([&] {
        fem_solver_s h;
        fem_solver_make(&A->A, A->stream, M, opts, &h);
        fem_solver_run(&h, b, x, stats);
    });
}

int fem_bsr_destroy(fem_bsr A) {
    return api_guard([&] { delete A; });
}

}  // extern "C"

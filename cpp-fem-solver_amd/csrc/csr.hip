// csr.hip — device CSR matrix and SpMV (sparse::SMatrix::spmv,
// ode-solver/src/math/array.cpp:3418-3442).
//
// SpMV kernel ("CSR-stream"): the host cuts the rows into consecutive ranges
// whose nonzeros fit a block's LDS staging area (at most kStageNnz nonzeros
// and kStageRows rows). A 256-thread block streams its contiguous slice of
// values/col_idx with fully coalesced loads, forms each product v*x[col] and
// stages it in LDS; then one thread per row adds its row's products in
// stored column order starting from 0.0. The products are rounded before the
// sum exactly as the reference's `sum += v * x` (no fused multiply-add: the
// library is built with -ffp-contract=off), so FP64 results are bit-identical
// to the reference's sequential row sum. Rows longer than the staging area
// get a block each and are reduced with a strided sum + block tree.
#include <mutex>
#include <cstring>
#include <algorithm>
#include <atomic>

#include "internal.hpp"
#include "reduce.hpp"

namespace fem {

void plan_rows(const uint32_t* h_rp, uint32_t lo, uint32_t hi, std::vector<uint32_t>& b, uint32_t max_nnz,
               uint32_t max_rows) {
    if (b.empty()) b.push_back(lo);
    // a block's entries, counted from its first entry rounded down to a
    // multiple of 4 (the staging's groups), fit max_nnz; a row that does not
    // fit alone gets a block of its own
    // (the span grows with the block's last row, so the block's end is found
    // by binary search on the row pointers)
    uint32_t start = lo;
    while (start < hi) {
        if (stage_span(h_rp[start], h_rp[start + 1]) > max_nnz) {
            ++start;
            b.push_back(start);
            continue;
        }
        const uint64_t lim64 = (uint64_t)(h_rp[start] & ~3u) + max_nnz;
        const uint32_t lim = lim64 > 0xffffffffull ? 0xffffffffu : (uint32_t)lim64;
        const uint32_t emax = (uint32_t)std::min<uint64_t>(hi, (uint64_t)start + max_rows);
        // the last end in [start + 1, emax] with rp[end] <= lim
        const uint32_t end = (uint32_t)(std::upper_bound(h_rp + start + 1, h_rp + emax + 1, lim) - h_rp) - 1;
        b.push_back(end);
        start = end;
    }
}

// Nonzeros per CSR-stream block: the LDS stage size, halved (down to 512)
// while the matrix would give fewer than min_blocks blocks. A launch of fewer
// blocks than the GPU holds at once runs one round of blocks and is bound by
// a block's chain of dependent loads; smaller blocks shorten that chain.
// Measured at 128^3 (SA-AMG V-cycle) and 96^3 (learned SPAI PCG): SpMV plans
// are best with min_blocks 1024, the Gauss-Seidel colour sweeps with 2048.
std::vector<uint2> block_bounds(const std::vector<uint32_t>& b, const uint32_t* h_rp) {
    std::vector<uint2> o(b.size());
    for (size_t i = 0; i < b.size(); ++i) o[i] = make_uint2(b[i], h_rp[b[i]]);
    return o;
}

// minimum block count of a matrix's SpMV plans (float64 and float32): 512,
// 2048 and 4096 measured no better at 128^3 (SA-AMG PCG) and 96^3 (learned
// SPAI PCG, whose folded reduction tails grow with the block count)
constexpr uint64_t kPlanMinBlocks = 1024;

uint32_t block_nnz(uint64_t nnz, uint32_t stage, uint64_t min_blocks) {
    uint32_t m = stage;
    while (m > 512 && nnz / m < min_blocks) m /= 2;
    return m;
}

void SpmvPlan::build(const uint32_t* h_rp, uint32_t n, hipStream_t s, uint32_t max_nnz, uint32_t max_rows) {
    std::vector<uint32_t> b;
    b.reserve(n / 128 + 4);
    b.push_back(0);
    plan_rows(h_rp, 0, n, b, max_nnz, max_rows);
    nblk = (uint32_t)b.size() - 1;
    bnd.upload(block_bounds(b, h_rp), s);
    FEM_HIP(hipStreamSynchronize(s));  // b ends with this call
}

// the vector the product reads at column c: x[c], or for the fused first
// two Jacobi sweeps the first sweep's value jac0(f[c], d[c])
template <typename T, int EPI>
__device__ __forceinline__ T spmv_x(const T* x, const SpmvArgs<T>& a, uint32_t c) {
    if constexpr (EPI == EPI_JACOBI0) return jac0<T, false>(a.f[c], a.d[c], (T)a.w);
    else if constexpr (EPI == EPI_JACOBI0_INV) return jac0<T, true>(a.f[c], a.d[c], (T)a.w);
    else return x[c];
}

template <typename T, int EPI>
__device__ __forceinline__ T spmv_epilogue(uint32_t row, T s, const SpmvArgs<T>& a) {
    if constexpr (EPI == EPI_STORE || EPI == EPI_STORE_J0 || EPI == EPI_STORE_J0_INV) return s;
    else if constexpr (EPI == EPI_RESID) return a.f[row] - s;
    else if constexpr (EPI == EPI_ADD) return a.u[row] + s;
    else if constexpr (EPI == EPI_JACOBI) return a.u[row] + ((T)a.w * (a.f[row] - s)) / a.d[row];
    else if constexpr (EPI == EPI_JACOBI_INV) return a.u[row] + (T)a.w * (a.d[row] * (a.f[row] - s));
    else if constexpr (EPI == EPI_JACOBI0)
        return jac0<T, false>(a.f[row], a.d[row], (T)a.w) + ((T)a.w * (a.f[row] - s)) / a.d[row];
    else return jac0<T, true>(a.f[row], a.d[row], (T)a.w) + (T)a.w * (a.d[row] * (a.f[row] - s));
}

// the second output of EPI_STORE_J0 / EPI_STORE_J0_INV: the first Jacobi
// sweep from zero with right-hand side y
template <typename T, int EPI>
__device__ __forceinline__ void spmv_second(uint32_t row, T y, T d, const SpmvArgs<T>& a) {
    if constexpr (EPI == EPI_STORE_J0) a.o[row] = jac0<T, false>(y, d, (T)a.w);
    else if constexpr (EPI == EPI_STORE_J0_INV) a.o[row] = jac0<T, true>(y, d, (T)a.w);
}

// Epilogue operands of one row, loaded before the block stages its products
// (their latency overlaps the staging instead of following the barrier).
template <typename T, int EPI>
struct EpiIn {
    T f = (T)0, u = (T)0, d = (T)0;
    __device__ __forceinline__ void load(uint32_t row, const SpmvArgs<T>& a) {
        constexpr bool J0 = EPI == EPI_JACOBI0 || EPI == EPI_JACOBI0_INV;
        if constexpr (EPI == EPI_RESID || EPI == EPI_JACOBI || EPI == EPI_JACOBI_INV || J0) f = a.f[row];
        if constexpr (EPI == EPI_ADD || EPI == EPI_JACOBI || EPI == EPI_JACOBI_INV) u = a.u[row];
        if constexpr (EPI == EPI_JACOBI || EPI == EPI_JACOBI_INV || J0 || EPI == EPI_STORE_J0 ||
                      EPI == EPI_STORE_J0_INV)
            d = a.d[row];
    }
    __device__ __forceinline__ T apply(T s, const SpmvArgs<T>& a) const {
        if constexpr (EPI == EPI_STORE || EPI == EPI_STORE_J0 || EPI == EPI_STORE_J0_INV) return s;
        else if constexpr (EPI == EPI_RESID) return f - s;
        else if constexpr (EPI == EPI_ADD) return u + s;
        else if constexpr (EPI == EPI_JACOBI) return u + ((T)a.w * (f - s)) / d;
        else if constexpr (EPI == EPI_JACOBI_INV) return u + (T)a.w * (d * (f - s));
        else if constexpr (EPI == EPI_JACOBI0) return jac0<T, false>(f, d, (T)a.w) + ((T)a.w * (f - s)) / d;
        else return jac0<T, true>(f, d, (T)a.w) + (T)a.w * (d * (f - s));
    }
};

// DOT = 1: the block also writes part[bid] = sum over its rows of
// x[i] * y[i] (float64; square matrices); DOT = 2: the sum of y[i]^2
template <typename T, int EPI, bool PK, int DOT = 0>
__global__ __launch_bounds__(kBlock) void k_spmv(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                                 const uint16_t* __restrict__ cd,
                                                 const uint32_t* __restrict__ cbase, const T* __restrict__ v,
                                                 const T* __restrict__ x, T* __restrict__ y,
                                                 const uint2* __restrict__ bnd, SpmvArgs<T> a, Gate g,
                                                 double* __restrict__ part = nullptr) {
    constexpr int STAGE = stage_nnz<T>();
    __shared__ T prod[STAGE];
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x);
    // the block's bounds are read while the gate's flag is in flight (reads
    // only: the gate guards the stores)
    const uint2 lo = bnd[bid], hi = bnd[bid + 1];
    const uint32_t r0 = lo.x, r1 = hi.x, k0 = lo.y, k1 = hi.y;
    if (g.closed()) return;
    const uint32_t cnt = k1 - k0;
    const uint32_t t = threadIdx.x;
    if (stage_fits<STAGE>(k0, cnt)) {
        // a block holds at most stage_rows<T>() rows: RPT per thread
        constexpr int RPT = (stage_rows<T>() + kBlock - 1) / kBlock;
        uint32_t rb[RPT], re[RPT];
        EpiIn<T, EPI> ep[RPT];
        T xr[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const uint32_t row = r0 + t + (uint32_t)j * kBlock;
            rb[j] = re[j] = 0;
            xr[j] = (T)0;
            if (row < r1) {
                rb[j] = rp[row] - k0;
                re[j] = rp[row + 1] - k0;
                ep[j].load(row, a);
                if (DOT == 1) xr[j] = x[row];
            }
        }
        if (cnt > 0)
            stage_products_g<T, STAGE, PK>(
                ci, cd, cbase, v, [&](uint32_t c) { return spmv_x<T, EPI>(x, a, c); }, k0, cnt, prod);
        __syncthreads();
        double dsum[1] = {0.0};
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const uint32_t row = r0 + t + (uint32_t)j * kBlock;
            if (row < r1) {
                T s = (T)0;
                for (uint32_t k = rb[j]; k < re[j]; ++k) s += prod[k];
                const T yv = ep[j].apply(s, a);
                y[row] = yv;
                spmv_second<T, EPI>(row, yv, ep[j].d, a);
                if (DOT == 1) dsum[0] += (double)xr[j] * (double)yv;
                if (DOT == 2) dsum[0] += (double)yv * (double)yv;
            }
        }
        if constexpr (DOT != 0) {
            block_sum<1>(dsum);
            if (t == 0) part[bid] = dsum[0];
        }
    } else {
        double s = 0.0;
        for (uint32_t k = k0 + t; k < k1; k += kBlock)
            s += (double)(v[k] * spmv_x<T, EPI>(x, a, col_of<PK>(ci, cd, cbase, k)));
        double tot[1] = {s};
        block_sum<1>(tot);
        if (t == 0) {
            const T yv = spmv_epilogue<T, EPI>(r0, (T)tot[0], a);
            y[r0] = yv;
            if constexpr (EPI == EPI_STORE_J0 || EPI == EPI_STORE_J0_INV) spmv_second<T, EPI>(r0, yv, a.d[r0], a);
            if (DOT == 1) part[bid] = (double)x[r0] * (double)yv;
            if (DOT == 2) part[bid] = (double)yv * (double)yv;
        }
    }
}

// cbase[c] = least column of chunk c (64 entries, one wavefront each);
// *bad |= 1 when the chunk's columns span 65536 or more
__global__ __launch_bounds__(256) void k_pack_base(const uint32_t* __restrict__ ci, uint64_t nnz, uint64_t nch,
                                                   uint32_t* __restrict__ cbase, int* __restrict__ bad) {
    const uint64_t c = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= nch) return;
    const uint64_t k0 = c << kPackShift;
    const uint64_t k = k0 + (threadIdx.x & 63);
    uint32_t mn = ci[k < nnz ? k : k0], mx = mn;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor(mn, off));
        mx = max(mx, (uint32_t)__shfl_xor(mx, off));
    }
    if ((threadIdx.x & 63) == 0) {
        cbase[c] = mn;
        if (mx - mn > 65535u) atomicOr(bad, 1);
    }
}

__global__ void k_pack_delta(const uint32_t* __restrict__ ci, uint64_t nnz, const uint32_t* __restrict__ cbase,
                             uint16_t* __restrict__ cd) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) cd[k] = (uint16_t)(ci[k] - cbase[k >> kPackShift]);
    else if (k < nnz + 4) cd[k] = 0;
}

// Vector-per-row SpMV: LANES consecutive lanes own one row, stride over its
// entries and add their partial sums by a butterfly; lane 0 of the group
// applies the epilogue. For long rows (restriction, coarse Galerkin
// operators) where one thread per row would serialise.
template <typename T, int EPI, int LANES>
__global__ __launch_bounds__(kBlock) void k_spmv_vec(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                                     const T* __restrict__ v, const T* __restrict__ x,
                                                     T* __restrict__ y, uint32_t n, SpmvArgs<T> a, Gate g) {
    const uint32_t row = (blockIdx.x * (uint32_t)kBlock + threadIdx.x) / LANES;
    const int lane = threadIdx.x % LANES;
    uint32_t k0 = 0, k1 = 0;
    if (row < n) {  // read while the gate's flag is in flight
        k0 = rp[row];
        k1 = rp[row + 1];
    }
    if (g.closed()) return;
    T s = (T)0;
    for (uint32_t k = k0 + lane; k < k1; k += LANES) s += v[k] * spmv_x<T, EPI>(x, a, ci[k]);
    s = group_sum<LANES>(s);
    if (row < n && lane == 0) {
        const T yv = spmv_epilogue<T, EPI>(row, s, a);
        y[row] = yv;
        if constexpr (EPI == EPI_STORE_J0 || EPI == EPI_STORE_J0_INV) spmv_second<T, EPI>(row, yv, a.d[row], a);
    }
}

template <typename T, int LANES>
void spmv_vec_launch(const CsrMat& A, const T* vals, const T* x, T* y, int epi, const SpmvArgs<T>& a, hipStream_t s,
                     Gate g) {
    const dim3 grid((unsigned)(((uint64_t)A.n * LANES + kBlock - 1) / kBlock)), block(kBlock);
    switch (epi) {
        case EPI_STORE: hipLaunchKernelGGL((k_spmv_vec<T, EPI_STORE, LANES>), grid, block, 0, s, A.rp.get(), A.ci.get(), vals, x, y, A.n, a, g); break;
        case EPI_RESID: hipLaunchKernelGGL((k_spmv_vec<T, EPI_RESID, LANES>), grid, block, 0, s, A.rp.get(), A.ci.get(), vals, x, y, A.n, a, g); break;
        case EPI_ADD: hipLaunchKernelGGL((k_spmv_vec<T, EPI_ADD, LANES>), grid, block, 0, s, A.rp.get(), A.ci.get(), vals, x, y, A.n, a, g); break;
        case EPI_JACOBI: hipLaunchKernelGGL((k_spmv_vec<T, EPI_JACOBI, LANES>), grid, block, 0, s, A.rp.get(), A.ci.get(), vals, x, y, A.n, a, g); break;
        case EPI_JACOBI_INV: hipLaunchKernelGGL((k_spmv_vec<T, EPI_JACOBI_INV, LANES>), grid, block, 0, s, A.rp.get(), A.ci.get(), vals, x, y, A.n, a, g); break;
        case EPI_JACOBI0: hipLaunchKernelGGL((k_spmv_vec<T, EPI_JACOBI0, LANES>), grid, block, 0, s, A.rp.get(), A.ci.get(), vals, x, y, A.n, a, g); break;
        case EPI_STORE_J0: hipLaunchKernelGGL((k_spmv_vec<T, EPI_STORE_J0, LANES>), grid, block, 0, s, A.rp.get(), A.ci.get(), vals, x, y, A.n, a, g); break;
        case EPI_STORE_J0_INV: hipLaunchKernelGGL((k_spmv_vec<T, EPI_STORE_J0_INV, LANES>), grid, block, 0, s, A.rp.get(), A.ci.get(), vals, x, y, A.n, a, g); break;
        default: hipLaunchKernelGGL((k_spmv_vec<T, EPI_JACOBI0_INV, LANES>), grid, block, 0, s, A.rp.get(), A.ci.get(), vals, x, y, A.n, a, g); break;
    }
}

template <typename T, int EPI>
void spmv_stream_launch(const CsrMat& A, const T* vals, const T* x, T* y, const SpmvArgs<T>& a, hipStream_t s,
                        Gate g) {
    const SpmvPlan& plan = std::is_same<T, double>::value ? A.plan : A.plan32;
    const dim3 grid(plan.nblk), block(kBlock);
    if (A.cd.get())
        hipLaunchKernelGGL((k_spmv<T, EPI, true>), grid, block, 0, s, A.rp.get(), A.ci.get(), A.cd.get(),
                           A.cbase.get(), vals, x, y, plan.bnd.get(), a, g);
    else
        hipLaunchKernelGGL((k_spmv<T, EPI, false>), grid, block, 0, s, A.rp.get(), A.ci.get(), A.cd.get(),
                           A.cbase.get(), vals, x, y, plan.bnd.get(), a, g);
}

template <typename T>
unsigned csr_spmv_dot(const CsrMat& A, const T* x, T* y, double* part, hipStream_t s, Gate g) {
    if (A.lanes > 0 || A.n != A.m || A.n == 0) return 0;
    const T* vals;
    if constexpr (std::is_same<T, double>::value) vals = A.v.get();
    else vals = A.v32.get();
    if (!vals) fail(FEM_ERR_RUNTIME, "float32 copy of the matrix was not built");
    const SpmvPlan& plan = std::is_same<T, double>::value ? A.plan : A.plan32;
    const dim3 grid(plan.nblk), block(kBlock);
    if (A.cd.get())
        hipLaunchKernelGGL((k_spmv<T, EPI_STORE, true, true>), grid, block, 0, s, A.rp.get(), A.ci.get(), A.cd.get(),
                           A.cbase.get(), vals, x, y, plan.bnd.get(), SpmvArgs<T>{}, g, part);
    else
        hipLaunchKernelGGL((k_spmv<T, EPI_STORE, false, true>), grid, block, 0, s, A.rp.get(), A.ci.get(),
                           A.cd.get(), A.cbase.get(), vals, x, y, plan.bnd.get(), SpmvArgs<T>{}, g, part);
    FEM_HIP(hipGetLastError());
    return plan.nblk;
}
template unsigned csr_spmv_dot<double>(const CsrMat&, const double*, double*, double*, hipStream_t, Gate);
template unsigned csr_spmv_dot<float>(const CsrMat&, const float*, float*, double*, hipStream_t, Gate);

template <typename T>
unsigned csr_spmv_selfdot(const CsrMat& A, const T* x, T* y, double* part, hipStream_t s, Gate g) {
    if (A.lanes > 0 || A.n == 0) return 0;
    const T* vals;
    if constexpr (std::is_same<T, double>::value) vals = A.v.get();
    else vals = A.v32.get();
    if (!vals) fail(FEM_ERR_RUNTIME, "float32 copy of the matrix was not built");
    const SpmvPlan& plan = std::is_same<T, double>::value ? A.plan : A.plan32;
    const dim3 grid(plan.nblk), block(kBlock);
    if (A.cd.get())
        hipLaunchKernelGGL((k_spmv<T, EPI_STORE, true, 2>), grid, block, 0, s, A.rp.get(), A.ci.get(), A.cd.get(),
                           A.cbase.get(), vals, x, y, plan.bnd.get(), SpmvArgs<T>{}, g, part);
    else
        hipLaunchKernelGGL((k_spmv<T, EPI_STORE, false, 2>), grid, block, 0, s, A.rp.get(), A.ci.get(), A.cd.get(),
                           A.cbase.get(), vals, x, y, plan.bnd.get(), SpmvArgs<T>{}, g, part);
    FEM_HIP(hipGetLastError());
    return plan.nblk;
}
template unsigned csr_spmv_selfdot<double>(const CsrMat&, const double*, double*, double*, hipStream_t, Gate);
template unsigned csr_spmv_selfdot<float>(const CsrMat&, const float*, float*, double*, hipStream_t, Gate);

unsigned CsrMat::apply_dot(const double* x, double* y, double* part, hipStream_t s, Gate g) const {
    return csr_spmv_dot<double>(*this, x, y, part, s, g);
}
unsigned CsrMat::apply32_dot(const float* x, float* y, double* part, hipStream_t s, Gate g) const {
    return csr_spmv_dot<float>(*this, x, y, part, s, g);
}
unsigned CsrMat::dot_parts() const { return std::max(plan.nblk, plan32.nblk); }

bool pack_columns(const uint32_t* ci, uint64_t nnz, DevBuf<uint16_t>& cd, DevBuf<uint32_t>& cbase, hipStream_t s) {
    cd.release();
    cbase.release();
    if (nnz < kPackMinNnz || nnz > 0xffffffffull) return false;
    const uint64_t nch = (nnz + (1u << kPackShift) - 1) >> kPackShift;
    cbase.alloc(nch);
    DevBuf<int> bad(1);
    FEM_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
    hipLaunchKernelGGL(k_pack_base, dim3((unsigned)((nch + 3) / 4)), dim3(256), 0, s, ci, nnz, nch, cbase.get(),
                       bad.get());
    FEM_HIP(hipGetLastError());
    int h = 0;
    bad.download(&h, 1, s);
    FEM_HIP(hipStreamSynchronize(s));
    if (h) {
        cbase.release();
        return false;
    }
    cd.alloc(nnz + 4);
    hipLaunchKernelGGL(k_pack_delta, dim3((unsigned)((nnz + 4 + 255) / 256)), dim3(256), 0, s, ci, nnz, cbase.get(),
                       cd.get());
    FEM_HIP(hipGetLastError());
    FEM_HIP(hipStreamSynchronize(s));  // bad ends with this call
    return true;
}

void CsrMat::use_vector_rows() {
    const double avg = n ? (double)nnz / n : 0.0;
    lanes = avg < 40 ? 0 : avg < 96 ? 32 : 64;
}

template <typename T>
void csr_spmv(const CsrMat& A, const T* x, T* y, int epi, const SpmvArgs<T>& a, hipStream_t s, Gate g) {
    if (A.n == 0) return;
    const T* vals;
    if constexpr (std::is_same<T, double>::value) vals = A.v.get();
    else vals = A.v32.get();
    if (!vals) fail(FEM_ERR_RUNTIME, "float32 copy of the matrix was not built");
    if (A.lanes > 0) {
        switch (A.lanes) {
            case 4: spmv_vec_launch<T, 4>(A, vals, x, y, epi, a, s, g); break;
            case 8: spmv_vec_launch<T, 8>(A, vals, x, y, epi, a, s, g); break;
            case 16: spmv_vec_launch<T, 16>(A, vals, x, y, epi, a, s, g); break;
            case 32: spmv_vec_launch<T, 32>(A, vals, x, y, epi, a, s, g); break;
            default: spmv_vec_launch<T, 64>(A, vals, x, y, epi, a, s, g); break;
        }
        FEM_HIP(hipGetLastError());
        return;
    }
    switch (epi) {
        case EPI_STORE: spmv_stream_launch<T, EPI_STORE>(A, vals, x, y, a, s, g); break;
        case EPI_RESID: spmv_stream_launch<T, EPI_RESID>(A, vals, x, y, a, s, g); break;
        case EPI_ADD: spmv_stream_launch<T, EPI_ADD>(A, vals, x, y, a, s, g); break;
        case EPI_JACOBI: spmv_stream_launch<T, EPI_JACOBI>(A, vals, x, y, a, s, g); break;
        case EPI_JACOBI_INV: spmv_stream_launch<T, EPI_JACOBI_INV>(A, vals, x, y, a, s, g); break;
        case EPI_JACOBI0: spmv_stream_launch<T, EPI_JACOBI0>(A, vals, x, y, a, s, g); break;
        case EPI_STORE_J0: spmv_stream_launch<T, EPI_STORE_J0>(A, vals, x, y, a, s, g); break;
        case EPI_STORE_J0_INV: spmv_stream_launch<T, EPI_STORE_J0_INV>(A, vals, x, y, a, s, g); break;
        default: spmv_stream_launch<T, EPI_JACOBI0_INV>(A, vals, x, y, a, s, g); break;
    }
    FEM_HIP(hipGetLastError());
}

template void csr_spmv<double>(const CsrMat&, const double*, double*, int, const SpmvArgs<double>&, hipStream_t, Gate);
template void csr_spmv<float>(const CsrMat&, const float*, float*, int, const SpmvArgs<float>&, hipStream_t, Gate);

void CsrMat::apply(const double* x, double* y, hipStream_t s, Gate g) const {
    csr_spmv<double>(*this, x, y, EPI_STORE, {}, s, g);
}
void CsrMat::apply32(const float* x, float* y, hipStream_t s, Gate g) const {
    csr_spmv<float>(*this, x, y, EPI_STORE, {}, s, g);
}

void CsrMat::ensure_fp32(hipStream_t s) {
    if (v32.get() || nnz == 0) return;
    v32.alloc(nnz + 4);
    FEM_HIP(hipMemsetAsync(v32.get() + nnz, 0, 4 * sizeof(float), s));
    vec_convert<float, double>(v32.get(), v.get(), 1.0, nnz, s);
    if (host_valid || rp_host) {
        plan32.build(h_rp.data(), n, s, block_nnz(nnz, stage_nnz<float>(), kPlanMinBlocks), stage_rows<float>());
    } else {
        std::vector<uint32_t> hrp(n + 1);
        download_words(rp.get(), (size_t)n + 1, hrp.data(), s);
        plan32.build(hrp.data(), n, s, block_nnz(nnz, stage_nnz<float>(), kPlanMinBlocks), stage_rows<float>());
    }
}

static uint64_t next_uid() {
    static std::atomic<uint64_t> c{0};
    return ++c;
}

void CsrMat::set_host(uint32_t rows, uint32_t cols, std::vector<uint32_t> rp_, std::vector<uint32_t> ci_,
                      std::vector<double> v_, hipStream_t s) {
    uid = next_uid();
    n = rows;
    m = cols;
    nnz = v_.size();
    stream = s;
    rp.upload(rp_, s);
    // 4 zero entries past the end (room for vector-width reads)
    ci.alloc(nnz + 4);
    v.alloc(nnz + 4);
    FEM_HIP(hipMemsetAsync(ci.get() + nnz, 0, 4 * sizeof(uint32_t), s));
    FEM_HIP(hipMemsetAsync(v.get() + nnz, 0, 4 * sizeof(double), s));
    ci.upload(ci_, s);
    v.upload(v_, s);
    plan.build(rp_.data(), n, s, block_nnz(nnz, stage_nnz<double>(), kPlanMinBlocks), stage_rows<double>());
    pack_columns(ci.get(), nnz, cd, cbase, s);
    h_rp = std::move(rp_);
    h_ci = std::move(ci_);
    h_v = std::move(v_);
    host_valid = true;
    rp_host = true;
    v32.release();
    FEM_HIP(hipStreamSynchronize(s));
}

void download_words(const uint32_t* d, size_t count, uint32_t* h, hipStream_t s) {
    // below this size a pageable copy costs no more than the staging step
    constexpr size_t kStageMin = (size_t)1 << 16;
    if (count < kStageMin) {
        if (count) FEM_HIP(hipMemcpyAsync(h, d, count * 4, hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        return;
    }
    // one staging buffer for the process, grown on demand and kept (pinned
    // host memory is not freed while the runtime may be shutting down)
    static std::mutex mu;
    static PinnedBuf<uint32_t>* stage = new PinnedBuf<uint32_t>();
    std::lock_guard<std::mutex> lock(mu);
    stage->ensure(count);
    FEM_HIP(hipMemcpyAsync(stage->p, d, count * 4, hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    std::memcpy(h, stage->p, count * 4);
}

void CsrMat::adopt(uint32_t rows, uint32_t cols, uint64_t nz, DevBuf<uint32_t>&& rp_, DevBuf<uint32_t>&& ci_,
                   DevBuf<double>&& v_, hipStream_t s, const CsrMat* same) {
    uid = next_uid();
    n = rows;
    m = cols;
    nnz = nz;
    stream = s;
    rp = std::move(rp_);
    ci = std::move(ci_);
    v = std::move(v_);
    const bool copy = same && same->rp_host && same->n == n && same->nnz == nnz && same->h_rp.size() == (size_t)n + 1;
    std::vector<uint32_t> hrp;
    if (copy) {
        hrp = same->h_rp;
    } else {
        hrp.resize((size_t)n + 1);
        download_words(rp.get(), (size_t)n + 1, hrp.data(), s);
    }
    plan.build(hrp.data(), n, s, block_nnz(nnz, stage_nnz<double>(), kPlanMinBlocks), stage_rows<double>());
    if (copy) {
        // packing is a function of the columns alone
        cd.release();
        cbase.release();
        if (same->cd.get() && same->cbase.get()) {
            cd.alloc(same->cd.n);
            cbase.alloc(same->cbase.n);
            FEM_HIP(hipMemcpyAsync(cd.get(), same->cd.get(), same->cd.n * sizeof(uint16_t), hipMemcpyDeviceToDevice, s));
            FEM_HIP(hipMemcpyAsync(cbase.get(), same->cbase.get(), same->cbase.n * sizeof(uint32_t),
                                   hipMemcpyDeviceToDevice, s));
        }
    } else {
        pack_columns(ci.get(), nnz, cd, cbase, s);
    }
    h_rp = std::move(hrp);  // kept for later plans (rp_host)
    h_ci.clear();
    h_v.clear();
    host_valid = false;
    rp_host = true;
    v32.release();
    lanes = 0;
    FEM_HIP(hipStreamSynchronize(s));
}

void CsrMat::sync_host() const {
    if (host_valid) return;
    h_rp.resize(n + 1);
    h_ci.resize(nnz);
    h_v.resize(nnz);
    rp.download(h_rp.data(), n + 1, stream);
    ci.download(h_ci.data(), nnz, stream);
    v.download(h_v.data(), nnz, stream);
    FEM_HIP(hipStreamSynchronize(stream));
    host_valid = true;
}

}  // namespace fem

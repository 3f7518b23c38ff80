// internal.hpp — device data structures of libfemgpu.
#pragma once
#include <chrono>
#include <memory>
#include <vector>

#include "common.hpp"

namespace fem {

// Kernel gate: a kernel runs only while the solve is live (*done == 0) and,
// when `need` is set, only if *need != 0. This is how data-dependent control
// flow (convergence, BiCGSTAB's true-residual check, GMRES early exit) is
// expressed inside captured hipGraphs without host round trips.
//
// The flags are written only by earlier kernels (a kernel that writes one
// reads it first), so a kernel reads them once, through the scalar cache,
// which is invalidated at every kernel launch.
__device__ __forceinline__ int flag_load(const int* p) {
    return *(const __attribute__((address_space(4))) int*)p;
}
// flags read in place of an absent gate pointer
static __device__ const int kFlagClear = 0;
static __device__ const int kFlagSet = 1;
struct Gate {
    const int* done = nullptr;
    const int* need = nullptr;
    __device__ __forceinline__ bool closed() const {
        // both flags are loaded before either is tested (one round trip)
        const int d = flag_load(done ? done : &kFlagClear);
        const int r = flag_load(need ? need : &kFlagSet);
        return (d != 0) | (r == 0);
    }
};

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

// ---------------------------------------------------------------------------
// Linear operators on device vectors
// ---------------------------------------------------------------------------
struct LinOp {
    virtual ~LinOp() {}
    virtual uint32_t rows() const = 0;
    virtual uint32_t cols() const = 0;
    virtual void apply(const double* x, double* y, hipStream_t s, Gate g = {}) const = 0;
    virtual void apply32(const float* x, float* y, hipStream_t s, Gate g = {}) const = 0;
    virtual void ensure_fp32(hipStream_t s) = 0;
    virtual uint64_t bytes_per_apply(bool fp32) const = 0;
    // y = A x and, per workgroup of the product, the float64 partial sum of
    // x[i] * y[i] into part[0..count); returns count, or 0 (nothing launched)
    // when this operator has no fused form
    virtual unsigned apply_dot(const double*, double*, double*, hipStream_t, Gate = {}) const { return 0; }
    virtual unsigned apply32_dot(const float*, float*, double*, hipStream_t, Gate = {}) const { return 0; }
    // most partial sums an apply_dot / apply32_dot call writes
    virtual unsigned dot_parts() const { return 0; }
};

// Workgroups are dispatched to the 8 XCDs round robin (workgroup b runs on
// XCD b % 8, each XCD with its own L2). This permutation hands each XCD a
// contiguous range of the logical block ids, so blocks that run on one XCD
// at about the same time work on neighbouring rows and share the L2 lines of
// the vector they gather from.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// SpMV launch schedule: consecutive row ranges whose nonzeros fit one block's
// LDS staging area, plus single long rows that get a whole block each.
// CSR-stream SpMV blocking: a block stages at most kStageNnz products of at
// most kStageRows consecutive rows in LDS (a longer row gets a block alone)
constexpr int kStageNnz = 2048;
constexpr int kStageRows = 256;
// float32 blocks stage twice the products in the same 16 KB of LDS
template <typename T>
constexpr int stage_nnz() { return sizeof(T) == 4 ? 2 * kStageNnz : kStageNnz; }
template <typename T>
constexpr int stage_rows() { return sizeof(T) == 4 ? 2 * kStageRows : kStageRows; }
// nonzeros per block for an nnz-entry matrix, at most stage, halved while
// fewer than min_blocks blocks would result (csr.hip)
uint32_t block_nnz(uint64_t nnz, uint32_t stage, uint64_t min_blocks);
// append the block boundaries of rows [lo, hi) to b (b starts with lo)
void plan_rows(const uint32_t* h_rp, uint32_t lo, uint32_t hi, std::vector<uint32_t>& b,
               uint32_t max_nnz = kStageNnz, uint32_t max_rows = kStageRows);

// block boundaries (first row, first entry) as one 8-byte load: a block
// reads its bounds and starts streaming entries without first reading the
// row pointers
std::vector<uint2> block_bounds(const std::vector<uint32_t>& b, const uint32_t* h_rp);

struct SpmvPlan {
    DevBuf<uint2> bnd;  // [nblk+1]
    uint32_t nblk = 0;
    void build(const uint32_t* h_rp, uint32_t n, hipStream_t s, uint32_t max_nnz = kStageNnz,
               uint32_t max_rows = kStageRows);
};

// Packed column indices: col[k] = cbase[k >> kPackShift] + cd[k], 2 bytes per
// entry instead of 4, for matrices whose columns span less than 65536 inside
// every chunk of 64 consecutive entries (banded FEM orderings: the
// z-neighbours of a 128^3 grid sit 16384 rows away). The CSR-stream kernels
// read cd and cbase instead of ci; products and summation order are unchanged.
constexpr int kPackShift = 6;
constexpr uint64_t kPackMinNnz = 1u << 16;  // smaller matrices keep ci only
// builds cd (nnz + 4 entries, zero padded) and cbase from the device array ci;
// false (both released) when a chunk spans 65536 columns or more or nnz is small
bool pack_columns(const uint32_t* ci, uint64_t nnz, DevBuf<uint16_t>& cd, DevBuf<uint32_t>& cbase, hipStream_t s);
// column of entry k: from the packed form or from ci
template <bool PK>
__device__ __forceinline__ uint32_t col_of(const uint32_t* __restrict__ ci, const uint16_t* __restrict__ cd,
                                           const uint32_t* __restrict__ cbase, uint32_t k) {
    if constexpr (PK) return cbase[k >> kPackShift] + (uint32_t)cd[k];
    else return ci[k];
}

// CSR-stream staging works on groups of 4 consecutive entries starting at a
// multiple of 4: a block's entries [k0, k0 + cnt) are covered by the groups
// from ka = k0 rounded down to a multiple of 4. A block is staged when those
// groups fit the stage (the plans cut blocks by the same rule).
template <int STAGE>
__device__ __forceinline__ bool stage_fits(uint32_t k0, uint32_t cnt) {
    return cnt + (k0 & 3u) <= (uint32_t)STAGE;
}
__host__ __device__ inline uint32_t stage_span(uint32_t k0, uint32_t k1) { return k1 - (k0 & ~3u); }

// four consecutive values from a 16-byte-aligned address (float32: one
// 16-byte load; float64: two)
__device__ __forceinline__ void load4(const float* p, float (&a)[4]) {
    const float4 w = *reinterpret_cast<const float4*>(p);
    a[0] = w.x; a[1] = w.y; a[2] = w.z; a[3] = w.w;
}
__device__ __forceinline__ void load4(const double* p, double (&a)[4]) {
    const double2 w0 = *reinterpret_cast<const double2*>(p);
    const double2 w1 = *reinterpret_cast<const double2*>(p + 2);
    a[0] = w0.x; a[1] = w0.y; a[2] = w1.x; a[3] = w1.y;
}
// the columns of entries k..k+3 (k a multiple of 4, so one 64-entry chunk):
// one 8-byte load of packed offsets plus the chunk base, or one 16-byte load
template <bool PK>
__device__ __forceinline__ void cols4(const uint32_t* __restrict__ ci, const uint16_t* __restrict__ cd,
                                      const uint32_t* __restrict__ cbase, uint32_t k, uint32_t (&c)[4]) {
    if constexpr (PK) {
        const uint32_t b = cbase[k >> kPackShift];
        const ushort4 d = *reinterpret_cast<const ushort4*>(cd + k);
        c[0] = b + d.x; c[1] = b + d.y; c[2] = b + d.z; c[3] = b + d.w;
    } else {
        const uint4 w = *reinterpret_cast<const uint4*>(ci + k);
        c[0] = w.x; c[1] = w.y; c[2] = w.z; c[3] = w.w;
    }
}

// CSR-stream staging: prod[k] = v[k0+k] * gx(col[k0+k]) for k < cnt
// (cnt >= 1, stage_fits) by a block of kBlock threads. A thread takes two
// groups of 4 entries per batch: it issues their value and column loads (16-
// and 8-byte loads from ka + 4 g; group indices clamped to the last group, so
// every load is in range: the entry arrays hold 4 padding entries past nnz
// and every entry, padding included, names a valid column), then its 8
// gathers gx (x[c], or a value formed from several vectors), then the LDS
// stores of the products inside [k0, k0 + cnt). Four times fewer load
// instructions than one entry per lane and load; the products and their
// order are unchanged.
template <typename T, int STAGE, bool PK, class GX>
__device__ __forceinline__ void stage_products_g(const uint32_t* __restrict__ ci, const uint16_t* __restrict__ cd,
                                                 const uint32_t* __restrict__ cbase, const T* __restrict__ v, GX gx,
                                                 uint32_t k0, uint32_t cnt, T* prod) {
    constexpr int B = 2;  // groups per thread and batch (1 and 4 were slower at 128^3)
    const uint32_t t = threadIdx.x;
    const uint32_t ka = k0 & ~3u, lead = k0 - ka;
    const uint32_t ng = (cnt + lead + 3) >> 2;
    for (uint32_t g0 = 0; g0 < ng; g0 += B * kBlock) {
        T a[B][4], xv[B][4];
        uint32_t c[B][4];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const uint32_t k = ka + 4 * min(g0 + (uint32_t)b * kBlock + t, ng - 1);
            load4(v + k, a[b]);
            cols4<PK>(ci, cd, cbase, k, c[b]);
        }
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) xv[b][j] = gx(c[b][j]);
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                // position in the block's range (wraps below 0 for the lead entries)
                const uint32_t k = 4 * (g0 + (uint32_t)b * kBlock + t) + j - lead;
                if (k < cnt) prod[k] = a[b][j] * xv[b][j];
            }
    }
}
template <typename T, int STAGE, bool PK>
__device__ __forceinline__ void stage_products(const uint32_t* __restrict__ ci, const uint16_t* __restrict__ cd,
                                               const uint32_t* __restrict__ cbase, const T* __restrict__ v,
                                               const T* x, uint32_t k0, uint32_t cnt, T* prod) {
    stage_products_g<T, STAGE, PK>(ci, cd, cbase, v, [x](uint32_t c) { return x[c]; }, k0, cnt, prod);
}

struct CsrMat : LinOp {
    uint32_t n = 0, m = 0;
    uint64_t nnz = 0;
    DevBuf<uint32_t> rp, ci;
    DevBuf<uint16_t> cd;     // packed columns (empty: not packable)
    DevBuf<uint32_t> cbase;  // per chunk of 64 entries
    DevBuf<double> v;
    DevBuf<float> v32;
    SpmvPlan plan, plan32;  // blockings of the float64 and float32 SpMV
    // host mirror (filled on demand for host-side setup)
    mutable std::vector<uint32_t> h_rp, h_ci;
    mutable std::vector<double> h_v;
    mutable bool host_valid = false;
    bool rp_host = false;  // h_rp holds the row pointers (set by set_host and adopt)
    hipStream_t stream = nullptr;  // owning stream
    // a fresh number each time set_host or adopt fills the matrix (user
    // matrices are never changed after that)
    uint64_t uid = 0;
    // 0: CSR-stream with sequential row sums, bit-exact with the reference's
    // spmv; > 0 (operators whose summation order is free: the AMG transfer
    // and coarse operators): that many lanes per row and a butterfly sum
    int lanes = 0;
    void use_vector_rows();  // the free-order variant for this matrix's rows

    uint32_t rows() const override { return n; }
    uint32_t cols() const override { return m; }
    void apply(const double* x, double* y, hipStream_t s, Gate g = {}) const override;
    void apply32(const float* x, float* y, hipStream_t s, Gate g = {}) const override;
    void ensure_fp32(hipStream_t s) override;
    unsigned apply_dot(const double* x, double* y, double* part, hipStream_t s, Gate g = {}) const override;
    unsigned apply32_dot(const float* x, float* y, double* part, hipStream_t s, Gate g = {}) const override;
    unsigned dot_parts() const override;
    uint64_t bytes_per_apply(bool fp32) const override {
        const uint64_t sv = fp32 ? 4 : 8;
        return nnz * (sv + 4) + 4ull * (n + 1) + (uint64_t)m * sv + (uint64_t)n * sv;
    }
    // upload host CSR (takes ownership of the host vectors as mirror)
    void set_host(uint32_t rows, uint32_t cols, std::vector<uint32_t> rp_, std::vector<uint32_t> ci_,
                  std::vector<double> v_, hipStream_t s);
    void sync_host() const;  // download into the host mirror if needed
    // take ownership of device arrays (ci, v with 4 padding entries); the
    // host mirror is filled on demand. `same` (optional): a matrix with the
    // same row pointers and columns, whose host row pointers and packed
    // columns are copied instead of being downloaded and recomputed.
    void adopt(uint32_t rows, uint32_t cols, uint64_t nz, DevBuf<uint32_t>&& rp_, DevBuf<uint32_t>&& ci_,
               DevBuf<double>&& v_, hipStream_t s, const CsrMat* same = nullptr);
};

// count row pointers (or other 32-bit words) from the device into h, through
// a pinned staging buffer for large counts; synchronises s
void download_words(const uint32_t* d, size_t count, uint32_t* h, hipStream_t s);

// the input step on the device (coo_device.hip)
void coo_to_csr_device(uint32_t n, uint32_t m, uint64_t nin, const double* v, const uint32_t* r, const uint32_t* c,
                       CsrMat& out, hipStream_t s);
void condense_device(const CsrMat& A, int bs, uint32_t nbr, uint32_t nbc, std::vector<uint32_t>& rp,
                     std::vector<uint32_t>& ci, std::vector<double>& v, hipStream_t s);

// Epilogue kinds of the CSR SpMV kernel.
enum SpmvEpi : int {
    EPI_STORE = 0,     // y = A x
    EPI_RESID = 1,     // y = f - A x
    EPI_ADD = 2,       // y = u + A x
    EPI_JACOBI = 3,    // y = u + w (f - A x) / d
    EPI_JACOBI_INV = 4, // y = u + w (dinv (f - A x))
    // the first two damped Jacobi sweeps from u = 0 in one launch: EPI_JACOBI
    // (EPI_JACOBI_INV) with x = u = jac0(f, d) formed where it is read (the
    // first sweep's values, bit for bit; the x argument is not read)
    EPI_JACOBI0 = 5,
    EPI_JACOBI0_INV = 6,
    // y = A x, and the first damped Jacobi sweep from zero of the level y is
    // the right-hand side of: o = jac0(y, d) (the restriction hands the
    // next level its first sweep)
    EPI_STORE_J0 = 7,
    EPI_STORE_J0_INV = 8
};

// One damped Jacobi sweep from u = 0 at one row: f - A 0 = f, then the
// update of u = 0 with weight w by the diagonal d (INV: d holds 1 / a_ii).
template <typename T, bool INV>
__device__ __forceinline__ T jac0(T f, T d, T w) {
    const T r = f - (T)0;
    if constexpr (INV) return (T)0 + w * (d * r);
    else return (T)0 + (w * r) / d;
}

template <typename T>
struct SpmvArgs {
    const T* f = nullptr;
    const T* u = nullptr;
    const T* d = nullptr;
    double w = 0.0;
    T* o = nullptr;  // second output (EPI_STORE_J0, EPI_STORE_J0_INV)
};

// Launch CSR SpMV with an epilogue (T = double or float values/vectors).
template <typename T>
void csr_spmv(const CsrMat& A, const T* x, T* y, int epi, const SpmvArgs<T>& a, hipStream_t s, Gate g);
// y = A x with the per-workgroup partial sums of x[i] * y[i] (square A, the
// CSR-stream kernel); returns their count, 0 when A takes another kernel
template <typename T>
unsigned csr_spmv_dot(const CsrMat& A, const T* x, T* y, double* part, hipStream_t s, Gate g);
// y = A x with the per-workgroup partial sums of y[i]^2 (the CSR-stream
// kernel); returns their count, 0 when A takes another kernel
template <typename T>
unsigned csr_spmv_selfdot(const CsrMat& A, const T* x, T* y, double* part, hipStream_t s, Gate g);

// ---------------------------------------------------------------------------
// Preconditioners
// ---------------------------------------------------------------------------
struct Precond {
    virtual ~Precond() {}
    virtual int kind() const = 0;
    virtual uint32_t size() const = 0;
    // z = M^{-1} r on device vectors (z != r)
    virtual void apply(const double* r, double* z, hipStream_t s, Gate g = {}) = 0;
    virtual void apply32(const float* r, float* z, hipStream_t s, Gate g = {}) = 0;
    virtual bool has_fp32() const { return false; }
    virtual void ensure_fp32(hipStream_t) {}
    // a solver will apply this preconditioner to vectors in the matrix's own
    // row order, in float64 (plain64) and/or float32 (plain32): data built on
    // first need are built now, before any graph capture
    virtual void prepare(bool /*plain64*/, bool /*plain32*/, hipStream_t) {}
    virtual bool graph_safe() const { return true; }
    // Non-null for FEM_PC_DIAG: the diagonal d (z = r / d) for fused kernels.
    virtual const double* diag64() const { return nullptr; }
    virtual const float* diag32() const { return nullptr; }
    // Learned preconditioners: the iteration budget of a solve before the
    // fallback preconditioner takes over (0: no fallback), and the fallback
    // (built on first use).
    virtual int budget() const { return 0; }
    virtual Precond* fallback(hipStream_t) { return nullptr; }
    // the fallback needed this many Krylov iterations (recorded as the count
    // a later learned solve must not exceed)
    virtual void fallback_ran(int /*iterations*/) {}
    // z = M^{-1} r with the per-workgroup partial sums of r[i] * z[i], as
    // LinOp::apply_dot (0: no fused form)
    virtual unsigned apply_dot(const double*, double*, double*, hipStream_t, Gate = {}) { return 0; }
    virtual unsigned apply32_dot(const float*, float*, double*, hipStream_t, Gate = {}) { return 0; }
    virtual unsigned dot_parts() const { return 0; }
    // M = G^T G with G and G^T held as CSR matrices (factorized sparse
    // approximate inverses): PCG then forms the products inside its own
    // update kernels. False for every other preconditioner.
    virtual bool factors(const CsrMat*& /*G*/, const CsrMat*& /*Gt*/) const { return false; }
    // Algorithmic bytes of one application with s-byte values: the sum over
    // the kernels it launches, each reading its matrix (uint32 indices) and
    // each vector once and writing its output once. 0: not tallied.
    virtual double apply_bytes(int /*s*/) { return 0.0; }
    // AMG: level l's operator (which = 0), prolongator (1) or restriction (2);
    // null when the level has no such matrix
    virtual const CsrMat* amg_matrix(int /*level*/, int /*which*/) const { return nullptr; }
    // C/F AMG: the split of level l (1 C, 2 F); null past the coarsened levels
    virtual const std::vector<uint8_t>* amg_split(int /*level*/) const { return nullptr; }
    // A float32 form of the system Q A Q^T, Q a row permutation, for inner
    // solves: op applies Q A Q^T, pc applies Q M^{-1} Q^T, and rows[t] is the
    // row of A at position t. True only when this preconditioner was built
    // from A itself (the same matrix contents) and holds such a form.
    virtual bool permuted32(const LinOp* /*A*/, LinOp*& /*op*/, Precond*& /*pc*/,
                            const uint32_t*& /*rows*/) {
        return false;
    }
    double setup_ms = 0.0;
    int levels = 0;
    std::vector<uint32_t> lvl_rows;
    std::vector<uint64_t> lvl_nnz;
};

// Algorithmic bytes of one CSR SpMV y = M x with s-byte values: the matrix
// with uint32 indices, x and y once, and `extra` more vectors of M's rows
// read by the epilogue.
inline double csr_bytes(const CsrMat& M, int s, int extra) {
    return (double)M.nnz * (s + 4) + 4.0 * (M.n + 1) + (double)M.m * s + (double)M.n * s * (1 + extra);
}

std::unique_ptr<Precond> make_precond(CsrMat& A, int kind, const fem_pc_params& p, hipStream_t s);


// ---------------------------------------------------------------------------
// Vector kernels (vec.hip)
// ---------------------------------------------------------------------------
template <typename T>
void vec_fill(T* y, T val, size_t n, hipStream_t s, Gate g = {});
template <typename TD, typename TS>
void vec_convert(TD* y, const TS* x, double scale, size_t n, hipStream_t s, Gate g = {});
// read n doubles (cache eviction without dirty lines); sink: `blocks` doubles
void cache_flush_read(const double* buf, size_t n, double* sink, unsigned blocks, hipStream_t s);
// reads `bytes` (a multiple of 16) with 16-byte loads, `blocks` workgroups
void stream_read(const void* buf, size_t bytes, float* sink, unsigned blocks, hipStream_t s);
// out = a . b, synchronous (host result); partials: kMaxRedBlocks doubles
double dot_sync(const double* a, const double* b, size_t n, double* partials, hipStream_t s);
// out = a . b on device pointers, enqueued on s, with no allocation: the
// partial sums go to a static device array (vec.hip); calls that overlap on
// one device race on it
void dot_noalloc(size_t n, const double* a, const double* b, double* out, hipStream_t s);

}  // namespace fem

// amg_device.hip — on-device setup of the robust smoothed-aggregation
// hierarchy (omega <= 0 and p_omega <= 0), producing exactly the hierarchy
// that the host restatement in amg_setup.cpp builds (same aggregates, same
// prolongator, same Galerkin operators bit for bit):
//   strength      |a_ij| >= theta sqrt(m_i m_j), m_i = max_{k != i} |a_ik|
//   aggregation   distance-2 maximal independent set in synchronous rounds
//                 (keys state<<62 | hash<<32 | index), then the two
//                 attachment phases and singleton leftovers (mis2_aggregate)
//   prolongator   P = (I - w D^{-1} A) P_t, w = (4/3) / rho(D^{-1} A) with
//                 rho the Gershgorin bound; per row the distinct aggregates in
//                 ascending order, each value accumulated in row order
//   Galerkin      A_c = P^T (A P): R = P^T by a counting sort (rows of R in
//                 ascending original row), then two sparse products. The
//                 pattern of row i of C = A B comes from an LDS bitmap over
//                 B's columns (read out in ascending order); the value of
//                 C_ik is sum over p in row i of A (ascending) of a_ip b_{j_p k},
//                 b found by binary search in the sorted row j_p of B — the
//                 order in which the host's row-wise product accumulates.
// The coarsest-level dense solve and the multicolour ordering of the SGS
// smoother stay on the host (they need each level's rows in order).
#include <algorithm>
#include <cmath>

#include <hipcub/hipcub.hpp>

#include "amg.hpp"
#include "reduce.hpp"

namespace fem {
namespace {

inline unsigned g1(size_t n) { return (unsigned)std::max<size_t>(1, (n + 255) / 256); }

__device__ __forceinline__ uint32_t node_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint64_t key_of(uint64_t st, uint32_t i) {
    return (st << 62) | ((uint64_t)(node_hash(i) & 0x3fffffffU) << 32) | i;
}

// out[0..n) -> exclusive prefix sum in out (n + 1 entries, total at out[n]);
// in has n + 1 entries with in[n] = 0
void scan_u32(const uint32_t* in, uint32_t* out, size_t n1, hipStream_t s) {
    size_t tmp = 0;
    FEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n1, s));
    DevBuf<char> t(std::max<size_t>(tmp, 1));
    FEM_HIP(hipcub::DeviceScan::ExclusiveSum(t.get(), tmp, in, out, (int)n1, s));
}

uint32_t read_u32(const uint32_t* p, hipStream_t s) {
    uint32_t h = 0;
    FEM_HIP(hipMemcpyAsync(&h, p, 4, hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    return h;
}

// ---- per-row quantities ---------------------------------------------------
__global__ void k_dinv(const uint32_t* rp, const uint32_t* ci, const double* v, uint32_t n, double* dinv, int* bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int dp = -1;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] == i) { dp = (int)p; break; }
    if (dp < 0) {
        atomicMin(bad, (int)i);
        dinv[i] = 0.0;
        return;
    }
    const double a = v[dp];
    dinv[i] = fabs(a) >= 1e-300 ? 1.0 / a : 0.0;
}

// per-block max of rho_i = sum_j |dinv_i a_ij| (max is order independent)
__global__ __launch_bounds__(kBlock) void k_gersh(const uint32_t* rp, const double* v, const double* dinv, uint32_t n,
                                                  double* partials) {
    double mx = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double rs = 0.0;
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) rs += fabs(0.0 + dinv[i] * v[p]);
        mx = fmax(mx, rs);
    }
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

__global__ void k_rowmax(const uint32_t* rp, const uint32_t* ci, const double* v, uint32_t n, double* m) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double mx = 0.0;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] != i) mx = fmax(mx, sqrt(v[p] * v[p]));
    m[i] = mx;
}

__global__ void k_strong(const uint32_t* rp, const uint32_t* ci, const double* v, const double* m, double theta,
                         uint32_t n, uint8_t* sm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
        const uint32_t j = ci[p];
        const double a = sqrt(v[p] * v[p]);
        sm[p] = (j != i && a != 0.0 && a >= theta * sqrt(m[i] * m[j])) ? 1 : 0;
    }
}

// ---- MIS-2 rounds ------------------------------------------------------------
__global__ void k_mis_max1(const uint32_t* rp, const uint32_t* ci, const uint8_t* sm, const uint8_t* st, uint32_t n,
                           uint64_t* t1) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t m = key_of(st[i], i);
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (sm[p]) m = max(m, key_of(st[ci[p]], ci[p]));
    t1[i] = m;
}

__global__ void k_mis_max2(const uint32_t* rp, const uint32_t* ci, const uint8_t* sm, const uint64_t* t1, uint32_t n,
                           uint8_t* st, uint32_t* undecided) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || st[i] != 1) return;
    uint64_t m = t1[i];
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (sm[p]) m = max(m, t1[ci[p]]);
    // the new state depends on t1 only, which this kernel does not write
    if ((uint32_t)m == i) st[i] = 2;
    else if ((m >> 62) == 2) st[i] = 0;
    else atomicAdd(undecided, 1u);
}

__global__ void k_flag(const uint8_t* st, uint32_t n, uint8_t want, uint32_t* f) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = st[i] == want ? 1u : 0u;
    if (i == n) f[n] = 0u;
}

__global__ void k_phase1(const uint32_t* rp, const uint32_t* ci, const uint8_t* sm, const uint8_t* st,
                         const uint32_t* rootpos, uint32_t n, int* a1) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (st[i] == 2) {
        a1[i] = (int)rootpos[i];
        return;
    }
    bool found = false;
    uint64_t best = 0;
    int g = -1;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
        const uint32_t j = ci[p];
        if (sm[p] && st[j] == 2) {
            const uint64_t k = key_of(0, j);
            if (!found || k > best) { found = true; best = k; g = (int)rootpos[j]; }
        }
    }
    a1[i] = g;
}

__global__ void k_phase2(const uint32_t* rp, const uint32_t* ci, const uint8_t* sm, const int* a1, uint32_t n,
                         int* agg, uint32_t* left) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        if (i == n) left[n] = 0u;
        return;
    }
    int g = a1[i];
    if (g < 0) {
        bool found = false;
        uint64_t best = 0;
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
            const uint32_t j = ci[p];
            if (sm[p] && a1[j] >= 0) {
                const uint64_t k = key_of(0, j);
                if (!found || k > best) { found = true; best = k; g = a1[j]; }
            }
        }
    }
    agg[i] = g;
    left[i] = g < 0 ? 1u : 0u;
}

__global__ void k_phase3(const uint32_t* leftpos, uint32_t nroots, uint32_t n, int* agg) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && agg[i] < 0) agg[i] = (int)(nroots + leftpos[i]);
}

// ---- smoothed prolongator ------------------------------------------------------
constexpr int kMaxPRow = 64;  // distinct aggregates per row of P

// distinct aggregates of row i ({agg[i]} and agg of its columns), ascending
__device__ __forceinline__ int p_row_cols(const uint32_t* rp, const uint32_t* ci, const int* agg, uint32_t i,
                                          int (&c)[kMaxPRow], int* overflow) {
    int k = 0;
    auto add = [&](int g) {
        for (int t = 0; t < k; ++t)
            if (c[t] == g) return;
        if (k == kMaxPRow) { atomicOr(overflow, 1); return; }
        int t = k++;
        while (t > 0 && c[t - 1] > g) { c[t] = c[t - 1]; --t; }
        c[t] = g;
    };
    add(agg[i]);
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) add(agg[ci[p]]);
    return k;
}

__global__ void k_pcount(const uint32_t* rp, const uint32_t* ci, const int* agg, uint32_t n, uint32_t* cnt,
                         int* overflow) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { cnt[n] = 0; return; }
    int c[kMaxPRow];
    cnt[i] = (uint32_t)p_row_cols(rp, ci, agg, i, c, overflow);
}

__global__ void k_pfill(const uint32_t* rp, const uint32_t* ci, const double* v, const int* agg, const double* dinv,
                        double pw, uint32_t n, const uint32_t* prp, uint32_t* pci, double* pv, int* overflow) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c[kMaxPRow];
    const int k = p_row_cols(rp, ci, agg, i, c, overflow);
    const uint32_t o = prp[i];
    for (int t = 0; t < k; ++t) {
        double val = c[t] == agg[i] ? 1.0 : 0.0;
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
            if (agg[ci[p]] == c[t]) val -= pw * (0.0 + dinv[i] * v[p]);
        pci[o + t] = (uint32_t)c[t];
        pv[o + t] = val;
    }
}

// ---- transpose -------------------------------------------------------------------
__global__ void k_colcount(const uint32_t* ci, uint64_t nnz, uint32_t* cnt) {
    const uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (p < nnz) atomicAdd(&cnt[ci[p]], 1u);
}

__global__ void k_tfill(const uint32_t* rp, const uint32_t* ci, const double* v, uint32_t n, uint32_t* cursor,
                        uint32_t* tci, double* tv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
        const uint32_t q = atomicAdd(&cursor[ci[p]], 1u);
        tci[q] = i;
        tv[q] = v[p];
    }
}

// sort each row of the transpose by column (insertion sort, rows are short)
__global__ void k_rowsort(const uint32_t* rp, uint32_t n, uint32_t* ci, double* v) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t b = rp[r], e = rp[r + 1];
    for (uint32_t p = b + 1; p < e; ++p) {
        const uint32_t c = ci[p];
        const double x = v[p];
        uint32_t q = p;
        while (q > b && ci[q - 1] > c) {
            ci[q] = ci[q - 1];
            v[q] = v[q - 1];
            --q;
        }
        ci[q] = c;
        v[q] = x;
    }
}

// ---- sparse product C = A B --------------------------------------------------------
// One wavefront per row of C (block = 64 threads); `words` 32-bit bitmap
// words over B's columns in dynamic LDS, cleared once, then after each row
// only the words that row touched. PASS 0 counts, PASS 1 writes the columns.
template <int PASS>
__global__ __launch_bounds__(64) void k_spgemm_sym(const uint32_t* arp, const uint32_t* aci, const uint32_t* brp,
                                                   const uint32_t* bci, uint32_t n, uint32_t words, uint32_t* cnt,
                                                   const uint32_t* crp, uint32_t* cci) {
    extern __shared__ uint32_t bm[];
    const int l = threadIdx.x;
    for (uint32_t w = l; w < words; w += 64) bm[w] = 0u;
    __syncthreads();
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint32_t a0 = arp[i], a1 = arp[i + 1];
        for (uint32_t p = a0 + l; p < a1; p += 64) {
            const uint32_t j = aci[p];
            for (uint32_t q = brp[j]; q < brp[j + 1]; ++q) {
                const uint32_t k = bci[q];
                atomicOr(&bm[k >> 5], 1u << (k & 31));
            }
        }
        __syncthreads();
        if (PASS == 0) {
            // popcount over the touched words: walk the products again and
            // count each word once (the first product that sees it clears a
            // copy in a second bitmap would need more LDS, so count set bits
            // of words as they are cleared)
            uint32_t c = 0;
            for (uint32_t p = a0 + l; p < a1; p += 64) {
                const uint32_t j = aci[p];
                for (uint32_t q = brp[j]; q < brp[j + 1]; ++q) {
                    const uint32_t w = bci[q] >> 5;
                    const uint32_t bits = atomicExch(&bm[w], 0u);
                    c += __popc(bits);
                }
            }
            for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
            if (l == 0) cnt[i] = c;
        } else {
            // read the set bits out in ascending column order: scan the words
            // between the row's smallest and largest column
            uint32_t lo = 0xffffffffu, hi = 0;
            for (uint32_t p = a0 + l; p < a1; p += 64) {
                const uint32_t j = aci[p];
                const uint32_t b0 = brp[j], b1 = brp[j + 1];
                if (b1 > b0) {
                    lo = min(lo, bci[b0] >> 5);
                    hi = max(hi, bci[b1 - 1] >> 5);
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                lo = min(lo, (uint32_t)__shfl_xor((int)lo, off));
                hi = max(hi, (uint32_t)__shfl_xor((int)hi, off));
            }
            uint32_t out = crp[i];
            if (lo <= hi) {
                for (uint32_t base = lo; base <= hi; base += 64) {
                    const uint32_t w = base + l;
                    const uint32_t bits = w <= hi ? bm[w] : 0u;
                    const uint32_t pc = __popc(bits);
                    // exclusive prefix of popcounts over the wave
                    uint32_t x = pc;
                    for (int off = 1; off < 64; off <<= 1) {
                        const uint32_t y = (uint32_t)__shfl_up((int)x, off);
                        if (l >= off) x += y;
                    }
                    uint32_t pos = out + x - pc;
                    uint32_t b = bits;
                    while (b) {
                        const int t = __ffs(b) - 1;
                        cci[pos++] = (w << 5) + (uint32_t)t;
                        b &= b - 1;
                    }
                    if (w <= hi) bm[w] = 0u;
                    out += (uint32_t)__shfl((int)x, 63);
                }
            }
        }
        __syncthreads();
    }
}

// C_ik = sum over p in row i of A (ascending) of a_ip * B[j_p][k]
__global__ __launch_bounds__(64) void k_spgemm_num(const uint32_t* arp, const uint32_t* aci, const double* av,
                                                   const uint32_t* brp, const uint32_t* bci, const double* bv,
                                                   const uint32_t* crp, const uint32_t* cci, double* cv, uint32_t n) {
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint32_t a0 = arp[i], a1 = arp[i + 1];
        for (uint32_t t = crp[i] + threadIdx.x; t < crp[i + 1]; t += 64) {
            const uint32_t k = cci[t];
            double acc = 0.0;
            for (uint32_t p = a0; p < a1; ++p) {
                const uint32_t j = aci[p];
                uint32_t lo = brp[j], hi = brp[j + 1];
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (bci[mid] < k) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < brp[j + 1] && bci[lo] == k) acc += av[p] * bv[lo];
            }
            cv[t] = acc;
        }
    }
}

struct DCsr {
    uint32_t n = 0, m = 0;
    uint64_t nnz = 0;
    DevBuf<uint32_t> rp, ci;
    DevBuf<double> v;
};

constexpr uint32_t kMaxBitmapWords = 16384;  // 64 KB of LDS: B with up to 524288 columns

DCsr spgemm_dev(const DCsr& A, const DCsr& B, hipStream_t s) {
    if (A.m != B.n) fail(FEM_ERR_RUNTIME, "AMG setup: product shape mismatch");
    const uint32_t words = (B.m + 31) / 32;
    if (words > kMaxBitmapWords) throw DeviceSetupTooLarge{};
    DCsr C;
    C.n = A.n;
    C.m = B.m;
    DevBuf<uint32_t> cnt(A.n + 1);
    C.rp.alloc(A.n + 1);
    const unsigned grid = std::max(1u, std::min(A.n, 8192u));
    const size_t lds = (size_t)words * 4;
    hipLaunchKernelGGL((k_spgemm_sym<0>), dim3(grid), dim3(64), lds, s, A.rp.get(), A.ci.get(), B.rp.get(),
                       B.ci.get(), A.n, words, cnt.get(), nullptr, nullptr);
    FEM_HIP(hipMemsetAsync(cnt.get() + A.n, 0, 4, s));
    scan_u32(cnt.get(), C.rp.get(), (size_t)A.n + 1, s);
    C.nnz = read_u32(C.rp.get() + A.n, s);
    C.ci.alloc(std::max<uint64_t>(C.nnz, 1));
    C.v.alloc(std::max<uint64_t>(C.nnz, 1));
    hipLaunchKernelGGL((k_spgemm_sym<1>), dim3(grid), dim3(64), lds, s, A.rp.get(), A.ci.get(), B.rp.get(),
                       B.ci.get(), A.n, words, nullptr, C.rp.get(), C.ci.get());
    hipLaunchKernelGGL(k_spgemm_num, dim3(grid), dim3(64), 0, s, A.rp.get(), A.ci.get(), A.v.get(), B.rp.get(),
                       B.ci.get(), B.v.get(), C.rp.get(), C.ci.get(), C.v.get(), A.n);
    FEM_HIP(hipGetLastError());
    return C;
}

DCsr transpose_dev(const DCsr& P, hipStream_t s) {
    DCsr R;
    R.n = P.m;
    R.m = P.n;
    R.nnz = P.nnz;
    DevBuf<uint32_t> cnt(P.m + 1), cur(P.m + 1);
    FEM_HIP(hipMemsetAsync(cnt.get(), 0, (size_t)(P.m + 1) * 4, s));
    hipLaunchKernelGGL(k_colcount, dim3(g1(P.nnz)), dim3(256), 0, s, P.ci.get(), P.nnz, cnt.get());
    R.rp.alloc(P.m + 1);
    scan_u32(cnt.get(), R.rp.get(), (size_t)P.m + 1, s);
    FEM_HIP(hipMemcpyAsync(cur.get(), R.rp.get(), (size_t)(P.m + 1) * 4, hipMemcpyDeviceToDevice, s));
    R.ci.alloc(std::max<uint64_t>(R.nnz, 1));
    R.v.alloc(std::max<uint64_t>(R.nnz, 1));
    hipLaunchKernelGGL(k_tfill, dim3(g1(P.n)), dim3(256), 0, s, P.rp.get(), P.ci.get(), P.v.get(), P.n, cur.get(),
                       R.ci.get(), R.v.get());
    hipLaunchKernelGGL(k_rowsort, dim3(g1(R.n)), dim3(256), 0, s, R.rp.get(), R.n, R.ci.get(), R.v.get());
    FEM_HIP(hipGetLastError());
    return R;
}

HCsr to_host(const DCsr& D, hipStream_t s) {
    HCsr H;
    H.n = D.n;
    H.m = D.m;
    H.rp.resize(D.n + 1);
    H.ci.resize(D.nnz);
    H.v.resize(D.nnz);
    D.rp.download(H.rp.data(), D.n + 1, s);
    D.ci.download(H.ci.data(), D.nnz, s);
    D.v.download(H.v.data(), D.nnz, s);
    FEM_HIP(hipStreamSynchronize(s));
    return H;
}

}  // namespace

// The robust SA hierarchy built on the device; the host copies of the level
// operators and prolongators are filled for the consumer (PcAmg uploads and
// colours from them) — a copy per level, made once per setup.
AmgHierarchy build_sa_hierarchy_device(const CsrMat& A0, const fem_pc_params& prm, hipStream_t s) {
    if (!(prm.p_omega <= 0.0)) fail(FEM_ERR_RUNTIME, "device SA setup is the robust variant only");
    AmgHierarchy H;
    const double theta = prm.theta;
    const int maxl = std::max(1, prm.max_levels);
    const uint32_t cmax = (uint32_t)std::max(1, prm.coarse_size);
    const bool auto_s = prm.omega <= 0.0;
    // level 0: a copy of the input (columns sorted and unique, as condense1
    // leaves a CSR built by this library)
    DCsr cur;
    cur.n = A0.n;
    cur.m = A0.m;
    cur.nnz = A0.nnz;
    cur.rp.alloc(A0.n + 1);
    cur.ci.alloc(std::max<uint64_t>(A0.nnz, 1));
    cur.v.alloc(std::max<uint64_t>(A0.nnz, 1));
    FEM_HIP(hipMemcpyAsync(cur.rp.get(), A0.rp.get(), (size_t)(A0.n + 1) * 4, hipMemcpyDeviceToDevice, s));
    FEM_HIP(hipMemcpyAsync(cur.ci.get(), A0.ci.get(), A0.nnz * 4, hipMemcpyDeviceToDevice, s));
    FEM_HIP(hipMemcpyAsync(cur.v.get(), A0.v.get(), A0.nnz * 8, hipMemcpyDeviceToDevice, s));
    DevBuf<int> flag(1);
    DevBuf<double> part(kMaxRedBlocks);
    for (int l = 0; l < maxl; ++l) {
        const uint32_t n = cur.n;
        DevBuf<double> dinv(n);
        const int big = 0x7fffffff;
        FEM_HIP(hipMemcpyAsync(flag.get(), &big, 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_dinv, dim3(g1(n)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), cur.v.get(), n,
                           dinv.get(), flag.get());
        int bad = 0;
        FEM_HIP(hipMemcpyAsync(&bad, flag.get(), 4, hipMemcpyDeviceToHost, s));
        const unsigned G = red_blocks(n);
        hipLaunchKernelGGL(k_gersh, dim3(G), dim3(kBlock), 0, s, cur.rp.get(), cur.v.get(), dinv.get(), n, part.get());
        std::vector<double> hp(G);
        part.download(hp.data(), G, s);
        FEM_HIP(hipStreamSynchronize(s));
        if (bad != big) fail(FEM_ERR_RUNTIME, "Block AMG: block row " + std::to_string(bad) + " has no diagonal block");
        double rho = 0.0;
        for (double x : hp) rho = std::max(rho, x);
        const double wauto = rho > 0.0 ? (4.0 / 3.0) / rho : 2.0 / 3.0;
        H.omega.push_back(auto_s ? wauto : prm.omega);
        const double pw = wauto;
        H.A.push_back(to_host(cur, s));
        if (n <= cmax) break;
        // strength and MIS-2 aggregation
        DevBuf<double> mrow(n);
        DevBuf<uint8_t> sm(std::max<uint64_t>(cur.nnz, 1)), st(n);
        hipLaunchKernelGGL(k_rowmax, dim3(g1(n)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), cur.v.get(), n, mrow.get());
        hipLaunchKernelGGL(k_strong, dim3(g1(n)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), cur.v.get(), mrow.get(),
                           theta, n, sm.get());
        FEM_HIP(hipMemsetAsync(st.get(), 1, n, s));
        DevBuf<uint64_t> t1(n);
        DevBuf<uint32_t> und(1);
        for (int round = 0;; ++round) {
            if (round > 10000) fail(FEM_ERR_RUNTIME, "AMG setup: aggregation did not finish");
            FEM_HIP(hipMemsetAsync(und.get(), 0, 4, s));
            hipLaunchKernelGGL(k_mis_max1, dim3(g1(n)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), sm.get(), st.get(),
                               n, t1.get());
            hipLaunchKernelGGL(k_mis_max2, dim3(g1(n)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), sm.get(), t1.get(),
                               n, st.get(), und.get());
            if (read_u32(und.get(), s) == 0) break;
        }
        DevBuf<uint32_t> f(n + 1), pos(n + 1);
        hipLaunchKernelGGL(k_flag, dim3(g1(n + 1)), dim3(256), 0, s, st.get(), n, (uint8_t)2, f.get());
        scan_u32(f.get(), pos.get(), (size_t)n + 1, s);
        const uint32_t nroots = read_u32(pos.get() + n, s);
        DevBuf<int> a1(n), agg(n);
        hipLaunchKernelGGL(k_phase1, dim3(g1(n)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), sm.get(), st.get(),
                           pos.get(), n, a1.get());
        hipLaunchKernelGGL(k_phase2, dim3(g1(n + 1)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), sm.get(), a1.get(), n,
                           agg.get(), f.get());
        scan_u32(f.get(), pos.get(), (size_t)n + 1, s);
        const uint32_t nleft = read_u32(pos.get() + n, s);
        hipLaunchKernelGGL(k_phase3, dim3(g1(n)), dim3(256), 0, s, pos.get(), nroots, n, agg.get());
        const uint32_t na = nroots + nleft;
        if ((double)na > 0.85 * (double)n) break;  // coarsening stalled
        // smoothed prolongator
        DCsr P;
        P.n = n;
        P.m = na;
        P.rp.alloc(n + 1);
        FEM_HIP(hipMemsetAsync(flag.get(), 0, 4, s));
        hipLaunchKernelGGL(k_pcount, dim3(g1(n + 1)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), agg.get(), n, f.get(),
                           flag.get());
        scan_u32(f.get(), P.rp.get(), (size_t)n + 1, s);
        P.nnz = read_u32(P.rp.get() + n, s);
        P.ci.alloc(P.nnz);
        P.v.alloc(P.nnz);
        hipLaunchKernelGGL(k_pfill, dim3(g1(n)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), cur.v.get(), agg.get(),
                           dinv.get(), pw, n, P.rp.get(), P.ci.get(), P.v.get(), flag.get());
        int ovf = 0;
        FEM_HIP(hipMemcpyAsync(&ovf, flag.get(), 4, hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        if (ovf) fail(FEM_ERR_RUNTIME, "AMG setup: a prolongator row has more than 64 aggregates");
        DCsr R = transpose_dev(P, s);
        DCsr AP = spgemm_dev(cur, P, s);
        DCsr Ac = spgemm_dev(R, AP, s);
        std::vector<double> hd(n);
        dinv.download(hd.data(), n, s);
        H.P.push_back(to_host(P, s));
        H.d.push_back(std::move(hd));
        cur = std::move(Ac);
    }
    H.coarse = sa_coarse_operator_host(H.A.back(), prm.coarse_exact != 0);
    if (prm.smoother == FEM_SMOOTH_SGS) sgs_levels_host(H);
    return H;
}

}  // namespace fem

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

// |a_ij| as sqrt(a_ij^2) (a scalar matrix) or the entry as given (NORM: v
// already holds a norm per entry, e.g. the Frobenius norms of 3x3 blocks)
template <bool NORM>
__device__ __forceinline__ double entry_norm(const double* v, uint32_t p) {
    if constexpr (NORM) return v[p];
    else return sqrt(v[p] * v[p]);
}

template <bool NORM>
__global__ void k_rowmax(const uint32_t* rp, const uint32_t* ci, const double* v, uint32_t n, double* m) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double mx = 0.0;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] != i) mx = fmax(mx, entry_norm<NORM>(v, p));
    m[i] = mx;
}

template <bool NORM>
__global__ void k_strong(const uint32_t* rp, const uint32_t* ci, const double* v, const double* m, double theta,
                         uint32_t n, uint8_t* sm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
        const uint32_t j = ci[p];
        const double a = entry_norm<NORM>(v, p);
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

__global__ void k_iota(uint32_t* x, uint64_t n) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k < n) x[k] = (uint32_t)k;
}

// row index of every stored entry
__global__ void k_entry_rows(const uint32_t* rp, uint32_t n, uint32_t* row) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) row[p] = i;
}

// entries of the transpose in (column, row) order: `idx` lists the entries
// of P stably sorted by column, so each column keeps ascending rows
__global__ void k_tgather(const uint32_t* idx, const uint32_t* row, const double* v, uint64_t nnz, uint32_t* tci,
                          double* tv) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const uint32_t e = idx[k];
    tci[k] = row[e];
    tv[k] = v[e];
}

// ---- sparse product C = A B --------------------------------------------------------
// One wavefront per row of C (block = 64 threads) with a bitmap of `words`
// 32-bit words in dynamic LDS. A row's columns span the words [lo, hi]
// (found from the first and last column of each B row it reads, B's rows
// being sorted); the span is covered in windows of `words` words: each
// window's products set their bits, then the window is read out in
// ascending column order (PASS 1 writes the columns, PASS 0 counts them) and
// cleared. Rows whose span fits one window (banded and stencil operators)
// take one pass; wider rows take more, so the product has no column limit.
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
        uint32_t c = 0;
        uint32_t out = PASS == 1 ? crp[i] : 0u;
        for (uint32_t w0 = lo; lo <= hi && w0 <= hi; w0 += words) {
            const uint32_t span = min(hi - w0, words - 1) + 1;  // words of this window
            for (uint32_t p = a0 + l; p < a1; p += 64) {
                const uint32_t j = aci[p];
                for (uint32_t q = brp[j]; q < brp[j + 1]; ++q) {
                    const uint32_t k = bci[q];
                    const uint32_t w = (k >> 5) - w0;
                    if ((k >> 5) >= w0 && w < span) atomicOr(&bm[w], 1u << (k & 31));
                }
            }
            __syncthreads();
            for (uint32_t base = 0; base < span; base += 64) {
                const uint32_t w = base + l;
                const uint32_t bits = w < span ? bm[w] : 0u;
                const uint32_t pc = __popc(bits);
                if (PASS == 0) {
                    c += pc;
                } else {
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
                        cci[pos++] = ((w0 + w) << 5) + (uint32_t)t;
                        b &= b - 1;
                    }
                    out += (uint32_t)__shfl((int)x, 63);
                }
                if (w < span) bm[w] = 0u;
            }
            __syncthreads();
        }
        if (PASS == 0) {
            c = group_sum<64>(c);
            if (l == 0) cnt[i] = c;
        }
    }
}

// C_ik = sum over p in row i of A (ascending) of a_ip * B[j_p][k]
// C_ik = sum over p in row i of A (ascending) of a_ip * B[j_p][k], C's
// columns given (symbolic pass). Rows of C with at most kNumCap entries:
// the row's columns and accumulators sit in LDS; the products a_ip b_jq are
// laid out p-major, 64 per round (all loads of a round issued together), and
// each lands on its column's accumulator by a binary search in LDS, the
// products of one round added in ascending p (those of one p go to distinct
// columns): per output the same additions in the same order as a search per
// (output, p). Longer rows search B's row for each (output, p).
constexpr uint32_t kNumCap = 512;
__global__ __launch_bounds__(64) void k_spgemm_num(const uint32_t* arp, const uint32_t* aci, const double* av,
                                                   const uint32_t* brp, const uint32_t* bci, const double* bv,
                                                   const uint32_t* crp, const uint32_t* cci, double* cv, uint32_t n) {
    __shared__ uint32_t cols[kNumCap];
    __shared__ double acc[kNumCap];
    __shared__ uint32_t s_off[65], s_bs[64];
    __shared__ double s_av[64];
    const uint32_t l = threadIdx.x;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint32_t a0 = arp[i], a1 = arp[i + 1];
        const uint32_t c0 = crp[i], nc = crp[i + 1] - c0;
        if (nc > kNumCap) {
            for (uint32_t t = c0 + l; t < c0 + nc; t += 64) {
                const uint32_t k = cci[t];
                double sum = 0.0;
                for (uint32_t p = a0; p < a1; ++p) {
                    const uint32_t j = aci[p];
                    uint32_t lo = brp[j], hi = brp[j + 1];
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (bci[mid] < k) lo = mid + 1;
                        else hi = mid;
                    }
                    if (lo < brp[j + 1] && bci[lo] == k) sum += av[p] * bv[lo];
                }
                cv[t] = sum;
            }
            continue;  // uniform: nc is the row's
        }
        for (uint32_t t = l; t < nc; t += 64) {
            cols[t] = cci[c0 + t];
            acc[t] = 0.0;
        }
        for (uint32_t pb = a0; pb < a1; pb += 64) {
            // this chunk's entries of A: B row start, length, value; offsets
            // of their products (inclusive scan of the lengths)
            const uint32_t p = pb + l;
            uint32_t bs = 0, len = 0;
            double a = 0.0;
            if (p < a1) {
                const uint32_t j = aci[p];
                bs = brp[j];
                len = brp[j + 1] - bs;
                a = av[p];
            }
            uint32_t x = len;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t y = (uint32_t)__shfl_up((int)x, off);
                if ((int)l >= off) x += y;
            }
            __syncthreads();  // the previous chunk's readers of s_* are done
            s_off[l + 1] = x;
            if (l == 0) s_off[0] = 0;
            s_bs[l] = bs;
            s_av[l] = a;
            __syncthreads();
            const uint32_t total = s_off[64];
            const uint32_t np = min(64u, a1 - pb);
            for (uint32_t r0 = 0; r0 < total; r0 += 64) {
                const uint32_t f = r0 + l;
                uint32_t o = 0, slot = 0;
                double prod = 0.0;
                if (f < total) {
                    // owner: the last entry whose products start at or before f
                    uint32_t lo = 0, hi = np;  // s_off[lo] <= f < s_off[hi]
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (s_off[mid] <= f) lo = mid;
                        else hi = mid;
                    }
                    o = lo;
                    const uint32_t q = s_bs[o] + (f - s_off[o]);
                    const uint32_t k = bci[q];
                    prod = s_av[o] * bv[q];
                    uint32_t a2 = 0, b2 = nc;  // lower bound of k in cols
                    while (a2 < b2) {
                        const uint32_t mid = (a2 + b2) >> 1;
                        if (cols[mid] < k) a2 = mid + 1;
                        else b2 = mid;
                    }
                    slot = a2 < nc ? a2 : 0u;  // k is among the row's columns (symbolic pass)
                }
                // owners of this round in ascending order
                const uint32_t ofirst = (uint32_t)__shfl((int)o, 0);
                const uint32_t last = min(total - r0, 64u) - 1;
                const uint32_t olast = (uint32_t)__shfl((int)o, (int)last);
                for (uint32_t oo = ofirst; oo <= olast; ++oo) {
                    if (f < total && o == oo) acc[slot] += prod;
                    __syncthreads();
                }
            }
        }
        __syncthreads();
        for (uint32_t t = l; t < nc; t += 64) cv[c0 + t] = acc[t];
        __syncthreads();  // cols and acc are rewritten by the next row
    }
}

struct DCsr {
    uint32_t n = 0, m = 0;
    uint64_t nnz = 0;
    DevBuf<uint32_t> rp, ci;
    DevBuf<double> v;
    // entry arrays with the 4 zero entries past the end that CsrMat keeps
    void alloc_entries(hipStream_t s) {
        ci.alloc(nnz + 4);
        v.alloc(nnz + 4);
        FEM_HIP(hipMemsetAsync(ci.get() + nnz, 0, 4 * sizeof(uint32_t), s));
        FEM_HIP(hipMemsetAsync(v.get() + nnz, 0, 4 * sizeof(double), s));
    }
    void to_mat(CsrMat& M, hipStream_t s) {
        M.adopt(n, m, nnz, std::move(rp), std::move(ci), std::move(v), s);
    }
};

// symbolic product window: 4 KB of LDS per 64-lane workgroup (windows of
// 32768 columns), so LDS does not limit the workgroups resident per CU (the
// 128^3 Poisson level-0 Galerkin products: 9.1 ms; 64 KB windows 17.0 ms,
// 1 KB 9.4 ms, 256 B 13.1 ms)
constexpr uint32_t kMaxBitmapWords = 1024;

// C = A B for A (an x am) and B (am x bm) given by their arrays
DCsr spgemm_arrays(uint32_t an, uint32_t am, const uint32_t* arp, const uint32_t* aci, const double* av, uint32_t bn,
                   uint32_t bm, const uint32_t* brp, const uint32_t* bci, const double* bv, hipStream_t s) {
    if (am != bn) fail(FEM_ERR_RUNTIME, "sparse product: shape mismatch");
    const uint32_t words = std::max(1u, std::min((bm + 31) / 32, kMaxBitmapWords));
    DCsr C;
    C.n = an;
    C.m = bm;
    DevBuf<uint32_t> cnt(an + 1);
    C.rp.alloc(an + 1);
    const unsigned grid = std::max(1u, std::min(an, 8192u));
    const size_t lds = (size_t)words * 4;
    if (an)
        hipLaunchKernelGGL((k_spgemm_sym<0>), dim3(grid), dim3(64), lds, s, arp, aci, brp, bci, an, words, cnt.get(),
                           nullptr, nullptr);
    FEM_HIP(hipMemsetAsync(cnt.get() + an, 0, 4, s));
    scan_u32(cnt.get(), C.rp.get(), (size_t)an + 1, s);
    C.nnz = read_u32(C.rp.get() + an, s);
    C.alloc_entries(s);
    if (an && C.nnz) {
        hipLaunchKernelGGL((k_spgemm_sym<1>), dim3(grid), dim3(64), lds, s, arp, aci, brp, bci, an, words, nullptr,
                           C.rp.get(), C.ci.get());
        hipLaunchKernelGGL(k_spgemm_num, dim3(grid), dim3(64), 0, s, arp, aci, av, brp, bci, bv, C.rp.get(),
                           C.ci.get(), C.v.get(), an);
    }
    FEM_HIP(hipGetLastError());
    FEM_HIP(hipStreamSynchronize(s));  // the temporaries end with this call
    return C;
}

DCsr spgemm_dev(const DCsr& A, const DCsr& B, hipStream_t s) {
    return spgemm_arrays(A.n, A.m, A.rp.get(), A.ci.get(), A.v.get(), B.n, B.m, B.rp.get(), B.ci.get(), B.v.get(), s);
}

// R = P^T for P given by its arrays (n x m, nnz entries)
DCsr transpose_arrays(const uint32_t* prp, const uint32_t* pci, const double* pv, uint32_t n, uint32_t m,
                      uint64_t nnz, hipStream_t s) {
    DCsr R;
    R.n = m;
    R.m = n;
    R.nnz = nnz;
    DevBuf<uint32_t> cnt(m + 1);
    FEM_HIP(hipMemsetAsync(cnt.get(), 0, (size_t)(m + 1) * 4, s));
    if (nnz) hipLaunchKernelGGL(k_colcount, dim3(g1(nnz)), dim3(256), 0, s, pci, nnz, cnt.get());
    R.rp.alloc(m + 1);
    scan_u32(cnt.get(), R.rp.get(), (size_t)m + 1, s);
    R.alloc_entries(s);
    if (nnz == 0) {
        FEM_HIP(hipStreamSynchronize(s));
        return R;
    }
    // stable LSD radix sort of the column indices with the entry position as
    // payload: within a column the entries stay in row order
    DevBuf<uint32_t> keys(nnz), pos(nnz), idx(nnz), row(nnz);
    hipLaunchKernelGGL(k_iota, dim3(g1(nnz)), dim3(256), 0, s, pos.get(), nnz);
    hipLaunchKernelGGL(k_entry_rows, dim3(g1(n)), dim3(256), 0, s, prp, n, row.get());
    int bits = 1;
    while (bits < 32 && (1ull << bits) < (uint64_t)m) ++bits;
    size_t tmp = 0;
    FEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, pci, keys.get(), pos.get(), idx.get(), (int64_t)nnz, 0,
                                               bits, s));
    DevBuf<char> t(std::max<size_t>(tmp, 1));
    FEM_HIP(hipcub::DeviceRadixSort::SortPairs(t.get(), tmp, pci, keys.get(), pos.get(), idx.get(), (int64_t)nnz, 0,
                                               bits, s));
    hipLaunchKernelGGL(k_tgather, dim3(g1(nnz)), dim3(256), 0, s, idx.get(), row.get(), pv, nnz, R.ci.get(),
                       R.v.get());
    FEM_HIP(hipGetLastError());
    FEM_HIP(hipStreamSynchronize(s));  // the temporaries end with this call
    return R;
}

DCsr transpose_dev(const DCsr& P, hipStream_t s) {
    return transpose_arrays(P.rp.get(), P.ci.get(), P.v.get(), P.n, P.m, P.nnz, s);
}

HCsr to_host(const CsrMat& M) {
    M.sync_host();
    HCsr H;
    H.n = M.n;
    H.m = M.m;
    H.rp = M.h_rp;
    H.ci = M.h_ci;
    H.v = M.h_v;
    return H;
}

// ---- SGS smoother data ----------------------------------------------------------
// The lower neighbours of row r in pattern(A + A^T): every off-diagonal entry
// (i, j) lists min(i, j) under max(i, j) (a symmetric pattern lists each
// neighbour twice, which the colouring does not mind).
__global__ void k_low_count(const uint32_t* rp, const uint32_t* ci, uint32_t n, uint32_t* cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
        const uint32_t j = ci[p];
        if (j != i) atomicAdd(&cnt[j > i ? j : i], 1u);
    }
}

__global__ void k_low_fill(const uint32_t* rp, const uint32_t* ci, uint32_t n, uint32_t* cur, uint32_t* lci) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
        const uint32_t j = ci[p];
        if (j == i) continue;
        const uint32_t r = j > i ? j : i;
        lci[atomicAdd(&cur[r], 1u)] = j > i ? i : j;
    }
}

// One round of the greedy colouring: a row whose lower neighbours all hold a
// colour takes the least colour none of them holds (colours from
// kMaxGsColours on are stored as kMaxGsColours). A colour, once written, is
// final, so reading a neighbour coloured earlier in the same round is as
// good as reading it in the next one.
__global__ void k_colour_round(const uint32_t* lrp, const uint32_t* lci, uint32_t n, int8_t* col, int* over) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    volatile int8_t* vc = col;
    if (vc[i] >= 0) return;
    uint32_t mask = 0;
    for (uint32_t p = lrp[i]; p < lrp[i + 1]; ++p) {
        const int c = vc[lci[p]];
        if (c < 0) return;
        mask |= 1u << c;
    }
    const uint32_t avail = ~mask & ((1u << kMaxGsColours) - 1u);
    const int c = avail ? __builtin_ctz(avail) : kMaxGsColours;
    vc[i] = (int8_t)c;
    if (c == kMaxGsColours) *over = 1;
}

__global__ void k_any_uncoloured(const int8_t* col, uint32_t n, int* left) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && col[i] < 0) *left = 1;
}

// rows per colour: a histogram per block, added to the totals
__global__ __launch_bounds__(256) void k_colour_hist(const int8_t* col, uint32_t n, uint32_t* cnt) {
    __shared__ uint32_t h[kMaxGsColours + 1];
    if (threadIdx.x <= kMaxGsColours) h[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicAdd(&h[col[i]], 1u);
    __syncthreads();
    if (threadIdx.x <= kMaxGsColours && h[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], h[threadIdx.x]);
}

// a_ii (the first stored diagonal entry; 0 when there is none)
__global__ void k_diag_val(const uint32_t* rp, const uint32_t* ci, const double* v, uint32_t n, double* d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = 0.0;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] == i) { a = v[p]; break; }
    d[i] = a;
}

// per-block max of sum_j |a_ij / a_ii| (row sums in stored order)
__global__ __launch_bounds__(kBlock) void k_gersh_div(const uint32_t* rp, const double* v, const double* d, uint32_t n,
                                                      double* partials) {
    double mx = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double rs = 0.0;
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) rs += fabs(v[p] / d[i]);
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

__global__ void k_group_len(const uint32_t* rp, const uint32_t* rows, uint32_t n, uint32_t* len) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) len[t] = rp[rows[t] + 1] - rp[rows[t]];
    if (t == n) len[t] = 0;
}

// row t of the grouped copy is row rows[t] of A, entries in stored order
__global__ void k_group_fill(const uint32_t* rp, const uint32_t* ci, const double* v, const double* diag,
                             const uint32_t* rows, uint32_t n, const uint32_t* gp, uint32_t* gci, double* gv,
                             double* gd) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = rows[t];
    uint32_t o = gp[t];
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p, ++o) {
        gci[o] = ci[p];
        gv[o] = v[p];
    }
    gd[t] = diag[i];
}

__global__ void k_colour_diag(const double* diag, const uint32_t* rows, uint32_t n, double* gd) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) gd[t] = diag[rows[t]];
}

__global__ void k_rows_sorted(const uint32_t* rp, const uint32_t* ci, uint32_t n, int* bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (uint32_t p = rp[i] + 1; p < rp[i + 1]; ++p)
        if (ci[p] <= ci[p - 1]) { *bad = 1; return; }
}

int read_i32(const int* p, hipStream_t s) {
    int h = 0;
    FEM_HIP(hipMemcpyAsync(&h, p, 4, hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    return h;
}

}  // namespace

bool rows_sorted_device(const CsrMat& A, hipStream_t s) {
    DevBuf<int> bad(1);
    FEM_HIP(hipMemsetAsync(bad.get(), 0, 4, s));
    hipLaunchKernelGGL(k_rows_sorted, dim3(g1(A.n)), dim3(256), 0, s, A.rp.get(), A.ci.get(), A.n, bad.get());
    FEM_HIP(hipGetLastError());
    return read_i32(bad.get(), s) == 0;
}

void transpose_device(const CsrMat& P, CsrMat& R, hipStream_t s) {
    DCsr T = transpose_arrays(P.rp.get(), P.ci.get(), P.v.get(), P.n, P.m, P.nnz, s);
    T.to_mat(R, s);
}

// Greedy multicolouring in row order, computed in rounds (a row is coloured
// once all its lower neighbours are; the lowest uncoloured row always is, so
// every round makes progress), stopped early once a row needs a colour
// beyond kMaxGsColours. Then the diagonal, the Jacobi weight and, when the
// colours fit, the rows grouped by colour (ascending row order within a
// colour) with a CSR copy of A in that order and its CSR-stream blocking.
void sgs_setup_device(AmgLevelDev& D, hipStream_t s, bool grouped) {
    const CsrMat& A = D.A;
    const uint32_t n = A.n;
    D.diag.alloc(n);
    hipLaunchKernelGGL(k_diag_val, dim3(g1(n)), dim3(256), 0, s, A.rp.get(), A.ci.get(), A.v.get(), n, D.diag.get());
    {
        const unsigned G = red_blocks(n);
        DevBuf<double> part(G);
        hipLaunchKernelGGL(k_gersh_div, dim3(G), dim3(kBlock), 0, s, A.rp.get(), A.v.get(), D.diag.get(), n,
                           part.get());
        std::vector<double> hp(G);
        part.download(hp.data(), G, s);
        FEM_HIP(hipStreamSynchronize(s));
        double rho = 0.0;
        for (double x : hp) rho = std::max(rho, x);
        D.jw = rho > 0.0 ? (4.0 / 3.0) / rho : 2.0 / 3.0;
    }
    // lower-neighbour lists
    DevBuf<uint32_t> cnt(n + 1), lrp(n + 1);
    FEM_HIP(hipMemsetAsync(cnt.get(), 0, (size_t)(n + 1) * 4, s));
    hipLaunchKernelGGL(k_low_count, dim3(g1(n)), dim3(256), 0, s, A.rp.get(), A.ci.get(), n, cnt.get());
    scan_u32(cnt.get(), lrp.get(), (size_t)n + 1, s);
    const uint32_t nlow = read_u32(lrp.get() + n, s);
    DevBuf<uint32_t> lci(std::max<uint32_t>(nlow, 1));
    FEM_HIP(hipMemcpyAsync(cnt.get(), lrp.get(), (size_t)(n + 1) * 4, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_low_fill, dim3(g1(n)), dim3(256), 0, s, A.rp.get(), A.ci.get(), n, cnt.get(), lci.get());
    // colouring rounds, checked every kBatch rounds
    DevBuf<int8_t> col(n);
    DevBuf<int> flags(2);  // [0] a colour beyond the cap, [1] rows left
    FEM_HIP(hipMemsetAsync(col.get(), 0xff, n, s));
    FEM_HIP(hipMemsetAsync(flags.get(), 0, 8, s));
    constexpr int kBatch = 32;
    int rounds = 0;
    bool over = false;
    for (;;) {
        for (int r = 0; r < kBatch; ++r)
            hipLaunchKernelGGL(k_colour_round, dim3(g1(n)), dim3(256), 0, s, lrp.get(), lci.get(), n, col.get(),
                               flags.get());
        rounds += kBatch;
        FEM_HIP(hipMemsetAsync(flags.get() + 1, 0, 4, s));
        hipLaunchKernelGGL(k_any_uncoloured, dim3(g1(n)), dim3(256), 0, s, col.get(), n, flags.get() + 1);
        int h[2];
        FEM_HIP(hipMemcpyAsync(h, flags.get(), 8, hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        if (h[0]) { over = true; break; }
        if (!h[1]) break;
        if ((uint64_t)rounds > (uint64_t)n + kBatch) fail(FEM_ERR_RUNTIME, "AMG setup: colouring did not finish");
    }
    FEM_HIP(hipGetLastError());
    if (over) {
        D.ncol = kMaxGsColours + 1;
        D.use_gs = false;
        return;
    }
    DevBuf<uint32_t> hc(kMaxGsColours + 1);
    FEM_HIP(hipMemsetAsync(hc.get(), 0, (kMaxGsColours + 1) * 4, s));
    hipLaunchKernelGGL(k_colour_hist, dim3(std::min<unsigned>(g1(n), 1024)), dim3(256), 0, s, col.get(), n, hc.get());
    std::vector<uint32_t> counts(kMaxGsColours + 1);
    hc.download(counts.data(), kMaxGsColours + 1, s);
    FEM_HIP(hipStreamSynchronize(s));
    int ncol = 0;
    for (int c = 0; c <= kMaxGsColours; ++c)
        if (counts[c]) ncol = c + 1;
    D.ncol = ncol;
    D.use_gs = ncol <= kMaxGsColours;
    D.h_cptr.assign(ncol + 1, 0);
    for (int c = 0; c < ncol; ++c) D.h_cptr[c + 1] = D.h_cptr[c] + counts[c];
    D.cptr.upload(D.h_cptr, s);
    // rows stably sorted by colour
    D.crows.alloc(std::max<uint32_t>(n, 1));
    {
        DevBuf<uint8_t> ksorted(std::max<uint32_t>(n, 1));
        DevBuf<uint32_t> ids(std::max<uint32_t>(n, 1));
        hipLaunchKernelGGL(k_iota, dim3(g1(n)), dim3(256), 0, s, ids.get(), (uint64_t)n);
        const uint8_t* keys = reinterpret_cast<const uint8_t*>(col.get());
        size_t tmp = 0;
        FEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys, ksorted.get(), ids.get(), D.crows.get(),
                                                   (int64_t)n, 0, 4, s));
        DevBuf<char> t(std::max<size_t>(tmp, 1));
        FEM_HIP(hipcub::DeviceRadixSort::SortPairs(t.get(), tmp, keys, ksorted.get(), ids.get(), D.crows.get(),
                                                   (int64_t)n, 0, 4, s));
        FEM_HIP(hipStreamSynchronize(s));
    }
    if (grouped) {
        sgs_grouped_device(D, s);
        return;
    }
    // the diagonal in colour order only (the grouped copy comes later, if at all)
    D.gd.alloc(std::max<uint32_t>(n, 1));
    hipLaunchKernelGGL(k_colour_diag, dim3(g1(n)), dim3(256), 0, s, D.diag.get(), D.crows.get(), n, D.gd.get());
    FEM_HIP(hipGetLastError());
    FEM_HIP(hipStreamSynchronize(s));  // the temporaries end with this call
}

void sgs_grouped_device(AmgLevelDev& D, hipStream_t s) {
    const CsrMat& A = D.A;
    const uint32_t n = A.n;
    const int ncol = D.ncol;
    if (!D.use_gs || !D.crows.get() || D.h_cptr.size() != (size_t)ncol + 1)
        fail(FEM_ERR_RUNTIME, "AMG: grouped copy of a level without its colouring");
    DevBuf<uint32_t> len(n + 1);
    hipLaunchKernelGGL(k_group_len, dim3(g1(n + 1)), dim3(256), 0, s, A.rp.get(), D.crows.get(), n, len.get());
    D.gp.alloc(n + 1);
    scan_u32(len.get(), D.gp.get(), (size_t)n + 1, s);
    D.gci.alloc(A.nnz + 4);
    D.gv.alloc(A.nnz + 4);
    D.gd.alloc(std::max<uint32_t>(n, 1));
    FEM_HIP(hipMemsetAsync(D.gci.get() + A.nnz, 0, 4 * sizeof(uint32_t), s));
    FEM_HIP(hipMemsetAsync(D.gv.get() + A.nnz, 0, 4 * sizeof(double), s));
    hipLaunchKernelGGL(k_group_fill, dim3(g1(n)), dim3(256), 0, s, A.rp.get(), A.ci.get(), A.v.get(), D.diag.get(),
                       D.crows.get(), n, D.gp.get(), D.gci.get(), D.gv.get(), D.gd.get());
    FEM_HIP(hipGetLastError());
    pack_columns(D.gci.get(), A.nnz, D.gcd, D.gcbase, s);
    D.h_gp.resize(n + 1);
    download_words(D.gp.get(), (size_t)n + 1, D.h_gp.data(), s);
    const std::vector<uint32_t>& gp = D.h_gp;
    std::vector<uint32_t> blk;
    D.h_cblk.assign(1, 0);
    for (int c = 0; c < ncol; ++c) {
        plan_rows(gp.data(), D.h_cptr[c], D.h_cptr[c + 1], blk,
                  block_nnz(gp[D.h_cptr[c + 1]] - gp[D.h_cptr[c]], stage_nnz<double>(), 2048), stage_rows<double>());
        D.h_cblk.push_back((uint32_t)blk.size() - 1);
    }
    D.gblk.upload(block_bounds(blk, gp.data()), s);
    const double avg = n ? (double)A.nnz / n : 0.0;
    D.lanes = avg <= 1.5 ? 1 : avg <= 3 ? 4 : avg <= 12 ? 8 : avg <= 24 ? 16 : avg <= 48 ? 32 : 64;
    FEM_HIP(hipStreamSynchronize(s));  // the host arrays end with this call
}

namespace {
__global__ void k_flag_rows(const uint32_t* rows, uint32_t count, uint8_t* flag) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) flag[rows[t]] = 1;
}

__global__ void k_keep_len(const uint32_t* rp, const uint32_t* ci, const uint8_t* drop, uint32_t n, uint32_t* len) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        uint32_t c = 0;
        for (uint32_t p = rp[t]; p < rp[t + 1]; ++p) c += drop[ci[p]] ? 0u : 1u;
        len[t] = c;
    }
    if (t == n) len[t] = 0;
}

__global__ void k_keep_fill(const uint32_t* rp, const uint32_t* ci, const double* v, const uint8_t* drop, uint32_t n,
                            const uint32_t* op, uint32_t* oci, double* ov) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t q = op[t];
    for (uint32_t p = rp[t]; p < rp[t + 1]; ++p)
        if (!drop[ci[p]]) {
            oci[q] = ci[p];
            ov[q] = v[p];
            ++q;
        }
}
}  // namespace

void drop_columns_device(const CsrMat& R, const uint32_t* rows, uint32_t count, CsrMat& out, hipStream_t s) {
    DevBuf<uint8_t> drop(std::max<uint32_t>(R.m, 1));
    FEM_HIP(hipMemsetAsync(drop.get(), 0, std::max<uint32_t>(R.m, 1), s));
    if (count) hipLaunchKernelGGL(k_flag_rows, dim3(g1(count)), dim3(256), 0, s, rows, count, drop.get());
    DCsr O;
    O.n = R.n;
    O.m = R.m;
    DevBuf<uint32_t> len(R.n + 1);
    hipLaunchKernelGGL(k_keep_len, dim3(g1(R.n + 1)), dim3(256), 0, s, R.rp.get(), R.ci.get(), drop.get(), R.n,
                       len.get());
    O.rp.alloc(R.n + 1);
    scan_u32(len.get(), O.rp.get(), (size_t)R.n + 1, s);
    O.nnz = read_u32(O.rp.get() + R.n, s);
    O.alloc_entries(s);
    hipLaunchKernelGGL(k_keep_fill, dim3(g1(R.n)), dim3(256), 0, s, R.rp.get(), R.ci.get(), R.v.get(), drop.get(),
                       R.n, O.rp.get(), O.ci.get(), O.v.get());
    FEM_HIP(hipGetLastError());
    O.to_mat(out, s);  // synchronises s
}

void sgs_blocks32(AmgLevelDev& D, hipStream_t s) {
    if (!D.gp.get() || D.h_cptr.size() < 2) return;
    const uint32_t n = D.n;
    const std::vector<uint32_t>& gp = D.h_gp;  // kept by sgs_setup_device
    if (gp.size() != (size_t)n + 1) fail(FEM_ERR_RUNTIME, "AMG: grouped row pointers not on the host");
    std::vector<uint32_t> blk;
    D.h_cblk32.assign(1, 0);
    for (size_t c = 0; c + 1 < D.h_cptr.size(); ++c) {
        plan_rows(gp.data(), D.h_cptr[c], D.h_cptr[c + 1], blk,
                  block_nnz(gp[D.h_cptr[c + 1]] - gp[D.h_cptr[c]], stage_nnz<float>(), 2048), stage_rows<float>());
        D.h_cblk32.push_back((uint32_t)blk.size() - 1);
    }
    D.gblk32.upload(block_bounds(blk, gp.data()), s);
    FEM_HIP(hipStreamSynchronize(s));  // blk ends with this call
}

namespace {
__global__ void k_positions(const uint32_t* rows, uint32_t n, uint32_t* pos) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) pos[rows[t]] = t;
}

// entries of row rows[t] other than its first diagonal entry (the one the
// smoother's diagonal holds)
__global__ void k_offdiag_len(const uint32_t* rp, const uint32_t* ci, const uint32_t* rows, uint32_t n,
                              uint32_t* len) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        const uint32_t i = rows[t];
        uint32_t c = rp[i + 1] - rp[i];
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
            if (ci[p] == i) { --c; break; }
        len[t] = c;
    }
    if (t == n) len[t] = 0;
}

__global__ void k_offdiag_fill(const uint32_t* rp, const uint32_t* ci, const double* v, const uint32_t* rows,
                               const uint32_t* pos, uint32_t n, const uint32_t* op, uint32_t* oci, double* ov) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t i = rows[t];
    uint32_t q = op[t];
    bool diag = false;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
        if (!diag && ci[p] == i) { diag = true; continue; }
        oci[q] = pos[ci[p]];
        ov[q] = v[p];
        ++q;
    }
}

// rows [0, nr) of the row-permuted copy: row t is row rows[t]
__global__ void k_prow_len(const uint32_t* rp, const uint32_t* rows, uint32_t nr, uint32_t* len) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nr) len[t] = rp[rows[t] + 1] - rp[rows[t]];
    if (t == nr) len[t] = 0;
}

__global__ void k_prow_fill(const uint32_t* rp, const uint32_t* ci, const double* v, const uint32_t* rows,
                            uint32_t nr, const uint32_t* op, uint32_t* oci, double* ov) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nr) return;
    const uint32_t i = rows[t];
    uint32_t q = op[t];
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p, ++q) {
        oci[q] = ci[p];
        ov[q] = v[p];
    }
}

__global__ void k_map_cols(const uint32_t* ci, uint64_t nnz, const uint32_t* pos, uint32_t* out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) out[k] = pos[ci[k]];
}
}  // namespace

void perm_setup_device(AmgLevelDev& D, hipStream_t s) {
    const uint32_t n = D.n;
    if (!D.crows.get() || !D.gd32.get() || D.ncol < 2 || !D.Rz.n || D.h_cptr.size() != (size_t)D.ncol + 1)
        fail(FEM_ERR_RUNTIME, "AMG: colour-ordered level without the smoother's colouring");
    const CsrMat& A = D.A;
    DevBuf<uint32_t> pos(n);
    hipLaunchKernelGGL(k_positions, dim3(g1(n)), dim3(256), 0, s, D.crows.get(), n, pos.get());
    {
        DCsr O;
        O.n = O.m = n;
        DevBuf<uint32_t> len(n + 1);
        hipLaunchKernelGGL(k_offdiag_len, dim3(g1(n + 1)), dim3(256), 0, s, A.rp.get(), A.ci.get(), D.crows.get(), n,
                           len.get());
        O.rp.alloc(n + 1);
        scan_u32(len.get(), O.rp.get(), (size_t)n + 1, s);
        O.nnz = read_u32(O.rp.get() + n, s);
        O.alloc_entries(s);
        hipLaunchKernelGGL(k_offdiag_fill, dim3(g1(n)), dim3(256), 0, s, A.rp.get(), A.ci.get(), A.v.get(),
                           D.crows.get(), pos.get(), n, O.rp.get(), O.ci.get(), O.v.get());
        FEM_HIP(hipGetLastError());
        O.to_mat(D.O, s);
    }
    {
        const uint32_t nr = D.h_cptr[D.ncol - 1];
        const CsrMat& P = D.P;
        DCsr Q;
        Q.n = nr;
        Q.m = P.m;
        DevBuf<uint32_t> len(nr + 1);
        hipLaunchKernelGGL(k_prow_len, dim3(g1(nr + 1)), dim3(256), 0, s, P.rp.get(), D.crows.get(), nr, len.get());
        Q.rp.alloc(nr + 1);
        scan_u32(len.get(), Q.rp.get(), (size_t)nr + 1, s);
        Q.nnz = read_u32(Q.rp.get() + nr, s);
        Q.alloc_entries(s);
        if (nr)
            hipLaunchKernelGGL(k_prow_fill, dim3(g1(nr)), dim3(256), 0, s, P.rp.get(), P.ci.get(), P.v.get(),
                               D.crows.get(), nr, Q.rp.get(), Q.ci.get(), Q.v.get());
        FEM_HIP(hipGetLastError());
        Q.to_mat(D.Pp, s);
    }
    {
        const CsrMat& R = D.Rz;
        DCsr Q;
        Q.n = R.n;
        Q.m = R.m;
        Q.rp.alloc(R.n + 1);
        FEM_HIP(hipMemcpyAsync(Q.rp.get(), R.rp.get(), (size_t)(R.n + 1) * 4, hipMemcpyDeviceToDevice, s));
        Q.nnz = R.nnz;
        Q.alloc_entries(s);
        if (R.nnz) {
            hipLaunchKernelGGL(k_map_cols, dim3(g1(R.nnz)), dim3(256), 0, s, R.ci.get(), R.nnz, pos.get(),
                               Q.ci.get());
            FEM_HIP(hipMemcpyAsync(Q.v.get(), R.v.get(), (size_t)R.nnz * 8, hipMemcpyDeviceToDevice, s));
        }
        FEM_HIP(hipGetLastError());
        Q.to_mat(D.Rzp, s);
    }
    D.Pp.use_vector_rows();
    D.Rzp.use_vector_rows();
    D.O.ensure_fp32(s);
    D.Pp.ensure_fp32(s);
    D.Rzp.ensure_fp32(s);
    const std::vector<uint32_t>& orp = D.O.h_rp;  // kept by adopt
    if (!D.O.rp_host || orp.size() != (size_t)n + 1) fail(FEM_ERR_RUNTIME, "AMG: row pointers of O not on the host");
    std::vector<uint32_t> blk;
    D.h_oblk32.assign(1, 0);
    D.h_onnz.assign(D.ncol, 0);
    // the blocking of the grouped copy's passes (block sizes 512, 1024,
    // 3072 and 4096 entries and the product's plan at 2048 and 3072 entries
    // per block measured within 1% of this at 128^3)
    for (int c = 0; c < D.ncol; ++c) {
        D.h_onnz[c] = orp[D.h_cptr[c + 1]] - orp[D.h_cptr[c]];
        plan_rows(orp.data(), D.h_cptr[c], D.h_cptr[c + 1], blk,
                  block_nnz(orp[D.h_cptr[c + 1]] - orp[D.h_cptr[c]], stage_nnz<float>(), 2048), stage_rows<float>());
        D.h_oblk32.push_back((uint32_t)blk.size() - 1);
    }
    D.oblk32.upload(block_bounds(blk, orp.data()), s);
    FEM_HIP(hipStreamSynchronize(s));  // blk, pos end with this call
    D.perm = true;
}

AmgDevHierarchy upload_hierarchy(AmgHierarchy&& H, hipStream_t s) {
    AmgDevHierarchy D;
    const size_t L = H.A.size();
    for (size_t k = 0; k < L; ++k) {
        auto lv = std::make_unique<AmgLevelDev>();
        HCsr& A = H.A[k];
        lv->n = A.n;
        lv->omega = H.omega[k];
        lv->A.set_host(A.n, A.m, std::move(A.rp), std::move(A.ci), std::move(A.v), s);
        if (k + 1 < L) {
            HCsr& P = H.P[k];
            lv->nc = P.m;
            lv->P.set_host(P.n, P.m, std::move(P.rp), std::move(P.ci), std::move(P.v), s);
            transpose_device(lv->P, lv->R, s);
            lv->d.upload(H.d[k], s);
        }
        D.lv.push_back(std::move(lv));
    }
    D.coarse = std::move(H.coarse);
    FEM_HIP(hipStreamSynchronize(s));
    return D;
}

// Aggregation by a distance-2 maximal independent set of the strength
// graph (the rounds, attachment phases and singleton leftovers described at
// the top of this file). v holds the matrix entries (norms = false: |a_ij|
// is sqrt(a_ij^2)) or a norm per entry (norms = true). agg receives n
// aggregate indices; returns the number of aggregates.
uint32_t mis2_aggregate_device(const uint32_t* rp, const uint32_t* ci, const double* v, bool norms, uint32_t n,
                               uint64_t nnz, double theta, int* agg, hipStream_t s) {
    DevBuf<double> mrow(n);
    DevBuf<uint8_t> sm(std::max<uint64_t>(nnz, 1)), st(n);
    if (norms) {
        hipLaunchKernelGGL(k_rowmax<true>, dim3(g1(n)), dim3(256), 0, s, rp, ci, v, n, mrow.get());
        hipLaunchKernelGGL(k_strong<true>, dim3(g1(n)), dim3(256), 0, s, rp, ci, v, mrow.get(), theta, n, sm.get());
    } else {
        hipLaunchKernelGGL(k_rowmax<false>, dim3(g1(n)), dim3(256), 0, s, rp, ci, v, n, mrow.get());
        hipLaunchKernelGGL(k_strong<false>, dim3(g1(n)), dim3(256), 0, s, rp, ci, v, mrow.get(), theta, n, sm.get());
    }
    FEM_HIP(hipMemsetAsync(st.get(), 1, n, s));
    DevBuf<uint64_t> t1(n);
    DevBuf<uint32_t> und(1);
    for (int round = 0;; ++round) {
        if (round > 10000) fail(FEM_ERR_RUNTIME, "AMG setup: aggregation did not finish");
        FEM_HIP(hipMemsetAsync(und.get(), 0, 4, s));
        hipLaunchKernelGGL(k_mis_max1, dim3(g1(n)), dim3(256), 0, s, rp, ci, sm.get(), st.get(), n, t1.get());
        hipLaunchKernelGGL(k_mis_max2, dim3(g1(n)), dim3(256), 0, s, rp, ci, sm.get(), t1.get(), n, st.get(), und.get());
        if (read_u32(und.get(), s) == 0) break;
    }
    DevBuf<uint32_t> f(n + 1), pos(n + 1);
    hipLaunchKernelGGL(k_flag, dim3(g1(n + 1)), dim3(256), 0, s, st.get(), n, (uint8_t)2, f.get());
    scan_u32(f.get(), pos.get(), (size_t)n + 1, s);
    const uint32_t nroots = read_u32(pos.get() + n, s);
    DevBuf<int> a1(n);
    hipLaunchKernelGGL(k_phase1, dim3(g1(n)), dim3(256), 0, s, rp, ci, sm.get(), st.get(), pos.get(), n, a1.get());
    hipLaunchKernelGGL(k_phase2, dim3(g1(n + 1)), dim3(256), 0, s, rp, ci, sm.get(), a1.get(), n, agg, f.get());
    scan_u32(f.get(), pos.get(), (size_t)n + 1, s);
    const uint32_t nleft = read_u32(pos.get() + n, s);
    hipLaunchKernelGGL(k_phase3, dim3(g1(n)), dim3(256), 0, s, pos.get(), nroots, n, agg);
    FEM_HIP(hipGetLastError());
    FEM_HIP(hipStreamSynchronize(s));  // the temporaries end with this call
    return nroots + nleft;
}

// The robust SA hierarchy built on the device; every level operator,
// prolongator and restriction stays there (only the coarsest operator is
// copied to the host, for its dense solve).
AmgDevHierarchy build_sa_hierarchy_device(const CsrMat& A0, const fem_pc_params& prm, hipStream_t s) {
    if (!(prm.p_omega <= 0.0)) fail(FEM_ERR_RUNTIME, "device SA setup is the robust variant only");
    AmgDevHierarchy H;
    SetupTrace tr("device SA", s);
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
    cur.alloc_entries(s);
    FEM_HIP(hipMemcpyAsync(cur.rp.get(), A0.rp.get(), (size_t)(A0.n + 1) * 4, hipMemcpyDeviceToDevice, s));
    FEM_HIP(hipMemcpyAsync(cur.ci.get(), A0.ci.get(), A0.nnz * 4, hipMemcpyDeviceToDevice, s));
    FEM_HIP(hipMemcpyAsync(cur.v.get(), A0.v.get(), A0.nnz * 8, hipMemcpyDeviceToDevice, s));
    DevBuf<int> flag(1);
    DevBuf<double> part(kMaxRedBlocks);
    for (int l = 0; l < maxl; ++l) {
        const uint32_t n = cur.n;
        auto lv = std::make_unique<AmgLevelDev>();
        lv->n = n;
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
        lv->omega = auto_s ? wauto : prm.omega;
        const double pw = wauto;
        tr.lap("diagonal, Gershgorin");
        if (n <= cmax) {
            cur.to_mat(lv->A, s);
            H.lv.push_back(std::move(lv));
            break;
        }
        // strength and MIS-2 aggregation
        DevBuf<int> agg(n);
        const uint32_t na = mis2_aggregate_device(cur.rp.get(), cur.ci.get(), cur.v.get(), false, n, cur.nnz, theta,
                                                  agg.get(), s);
        tr.lap("strength, MIS-2 aggregation");
        if ((double)na > 0.85 * (double)n) {  // coarsening stalled
            cur.to_mat(lv->A, s);
            H.lv.push_back(std::move(lv));
            break;
        }
        // smoothed prolongator
        DCsr P;
        P.n = n;
        P.m = na;
        P.rp.alloc(n + 1);
        DevBuf<uint32_t> f(n + 1);
        FEM_HIP(hipMemsetAsync(flag.get(), 0, 4, s));
        hipLaunchKernelGGL(k_pcount, dim3(g1(n + 1)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), agg.get(), n, f.get(),
                           flag.get());
        scan_u32(f.get(), P.rp.get(), (size_t)n + 1, s);
        P.nnz = read_u32(P.rp.get() + n, s);
        P.alloc_entries(s);
        hipLaunchKernelGGL(k_pfill, dim3(g1(n)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), cur.v.get(), agg.get(),
                           dinv.get(), pw, n, P.rp.get(), P.ci.get(), P.v.get(), flag.get());
        int ovf = 0;
        FEM_HIP(hipMemcpyAsync(&ovf, flag.get(), 4, hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        if (ovf) fail(FEM_ERR_RUNTIME, "AMG setup: a prolongator row has more than 64 aggregates");
        tr.lap("prolongator");
        DCsr R = transpose_dev(P, s);
        tr.lap("transpose");
        DCsr AP = spgemm_dev(cur, P, s);
        DCsr Ac = spgemm_dev(R, AP, s);
        tr.lap("Galerkin products");
        lv->nc = na;
        lv->d = std::move(dinv);
        // level 0 is a copy of A0: its host row pointers and packed columns
        // are A0's
        if (l == 0) lv->A.adopt(cur.n, cur.m, cur.nnz, std::move(cur.rp), std::move(cur.ci), std::move(cur.v), s, &A0);
        else cur.to_mat(lv->A, s);
        P.to_mat(lv->P, s);
        R.to_mat(lv->R, s);
        H.lv.push_back(std::move(lv));
        cur = std::move(Ac);
        tr.lap("level operators kept");
    }
    // a hierarchy cut by max_levels ends at the last level built, whose
    // transfer operators go unused
    AmgLevelDev& last = *H.lv.back();
    last.P = CsrMat();
    last.R = CsrMat();
    last.nc = 0;
    H.coarse = sa_coarse_operator_host(to_host(last.A), prm.coarse_exact != 0);
    tr.lap("coarsest operator");
    return H;
}

// ---------------------------------------------------------------------------
// Classical (C/F) AMG setup on the device. The operator it builds is the
// reference's C/F AMG (ode-solver/src/math/array.cpp:3182-3343): strength
// |a_ij| >= theta max_{k != i} |a_ik|, a C/F split in which every F point
// strongly depends on a C point, direct interpolation with the weak
// couplings lumped into the diagonal, Galerkin coarse operators, a dense
// coarsest solve. The split itself is this library's own, parallel, choice
// (the reference pops points from a serial priority queue): C points are
// picked in synchronous rounds, each undecided point whose key beats every
// undecided neighbour in the symmetrised strength graph becoming C, and
// every undecided point that strongly depends on a C point becoming F. This
// is the PMIS selection (De Sterck, Yang & Heys, "Reducing complexity in
// parallel algebraic multigrid preconditioners", SIAM J. Matrix Anal. Appl.
// 27(4), 2006) with colour-augmented keys as in Alber & Olson ("Parallel
// coarse-grid selection", Numer. Linear Algebra Appl. 14, 2007): a point's
// key is (number of points strongly depending on it, its colour in the
// greedy row-order colouring of the symmetrised strength graph, a hash of
// its index, its index). On grid operators the colouring is the red-black
// (or line-wise) one, and the split comes out as the structured coarsening
// the serial priority queue finds; keys with a random tie-break instead
// leave defects that cost about half as many iterations again. Points
// without strong couplings become C, as in the reference. Nothing depends
// on thread scheduling.
namespace {

// strong couplings of row i: per-entry flag and count
__global__ void k_cf_strength(const uint32_t* rp, const uint32_t* ci, const double* v, double theta, uint32_t n,
                              uint8_t* sm, uint32_t* cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { cnt[n] = 0u; return; }
    double mx = 0.0;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] != i) mx = fmax(mx, fabs(v[p]));
    const double thr = theta * mx;
    uint32_t c = 0;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
        const double a = fabs(v[p]);
        const bool st = ci[p] != i && a != 0.0 && a >= thr;
        sm[p] = st ? 1 : 0;
        c += st ? 1u : 0u;
    }
    cnt[i] = c;
}

// strength graph S as CSR (columns in A's stored order)
__global__ void k_cf_sfill(const uint32_t* rp, const uint32_t* ci, const uint8_t* sm, uint32_t n, const uint32_t* srp,
                           uint32_t* sci) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t o = srp[i];
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (sm[p]) sci[o++] = ci[p];
}

// greedy colouring of S + S^T in row order, in rounds: a point whose lower
// neighbours are all coloured takes the least colour none of them holds
// (colours from 63 on stored as 63: the colour only orders keys). A colour,
// once written, is final, so a neighbour coloured earlier in the same round
// reads as well as one coloured in an earlier round.
constexpr uint8_t kNoColour = 0xFF;
__global__ void k_cf_colour(const uint32_t* srp, const uint32_t* sci, const uint32_t* trp, const uint32_t* tci,
                            uint32_t n, uint8_t* col, uint32_t* left) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || col[i] != kNoColour) return;
    uint64_t used = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const uint32_t* rp = pass ? trp : srp;
        const uint32_t* cl = pass ? tci : sci;
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
            const uint32_t j = cl[p];
            if (j >= i) continue;
            const uint8_t c = __hip_atomic_load(&col[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (c == kNoColour) { atomicAdd(left, 1u); return; }
            used |= 1ull << (c < 63 ? c : 63);
        }
    }
    const int c = __builtin_ctzll(~used);
    __hip_atomic_store(&col[i], (uint8_t)(c < 63 ? c : 63), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// key order (lambda, colour, hash(index), index), lambda = |S^T row|
__device__ __forceinline__ bool cf_beats(uint32_t la, uint8_t ca, uint32_t a, uint32_t lb, uint8_t cb, uint32_t b) {
    if (la != lb) return la > lb;
    if (ca != cb) return ca > cb;
    const uint32_t ha = node_hash(a), hb = node_hash(b);
    if (ha != hb) return ha > hb;
    return a > b;
}

enum : uint8_t { CF_U = 0, CF_C = 1, CF_F = 2 };

// one selection round: an undecided point beating every undecided
// neighbour in S + S^T becomes C (reads st_in only, writes st_out)
__global__ void k_cf_select(const uint32_t* srp, const uint32_t* sci, const uint32_t* trp, const uint32_t* tci,
                            const uint8_t* col, const uint8_t* st_in, uint32_t n, uint8_t* st_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t si = st_in[i];
    if (si != CF_U) { st_out[i] = si; return; }
    const uint32_t li = trp[i + 1] - trp[i];
    const uint8_t ci = col[i];
    bool best = true;
    for (uint32_t p = srp[i]; p < srp[i + 1] && best; ++p) {
        const uint32_t j = sci[p];
        if (st_in[j] == CF_U && cf_beats(trp[j + 1] - trp[j], col[j], j, li, ci, i)) best = false;
    }
    for (uint32_t p = trp[i]; p < trp[i + 1] && best; ++p) {
        const uint32_t j = tci[p];
        if (st_in[j] == CF_U && cf_beats(trp[j + 1] - trp[j], col[j], j, li, ci, i)) best = false;
    }
    st_out[i] = best ? CF_C : CF_U;
}

// undecided points that strongly depend on a C point become F; counts the
// points left undecided. C states are final and only undecided -> F is
// written, so the reads of "is C" see the same values in any order.
__global__ void k_cf_fmark(const uint32_t* srp, const uint32_t* sci, uint32_t n, uint8_t* st, uint32_t* left) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || st[i] != CF_U) return;
    for (uint32_t p = srp[i]; p < srp[i + 1]; ++p)
        if (st[sci[p]] == CF_C) { st[i] = CF_F; return; }
    atomicAdd(left, 1u);
}

// rows of P: 1 for a C point, its strong C neighbours for an F point;
// an F point without one counts into *orphan (not produced by the rounds)
__global__ void k_cf_pcount(const uint32_t* srp, const uint32_t* sci, const uint8_t* st, uint32_t n, uint32_t* cnt,
                            uint32_t* orphan) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { cnt[n] = 0u; return; }
    if (st[i] == CF_C) { cnt[i] = 1u; return; }
    uint32_t c = 0;
    for (uint32_t p = srp[i]; p < srp[i + 1]; ++p) c += st[sci[p]] == CF_C ? 1u : 0u;
    if (c == 0) atomicAdd(orphan, 1u);
    cnt[i] = c;
}

// direct interpolation: C row i -> (mapc[i], 1); F row i -> for each strong C
// neighbour j in stored order (mapc[j], -a_ij / den), den = a_ii plus the
// weak off-diagonal couplings of row i in stored order (a_ii if |den| <
// 1e-300). mapc is increasing in the row index, so the columns are sorted.
__global__ void k_cf_pfill(const uint32_t* rp, const uint32_t* ci, const double* v, const uint8_t* sm,
                           const uint8_t* st, const uint32_t* mapc, const double* d, uint32_t n, const uint32_t* prp,
                           uint32_t* pci, double* pv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t o = prp[i];
    if (st[i] == CF_C) {
        pci[o] = mapc[i];
        pv[o] = 1.0;
        return;
    }
    double den = d[i];
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] != i && !sm[p]) den += v[p];
    if (fabs(den) < 1e-300) den = d[i];
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
        const uint32_t j = ci[p];
        if (sm[p] && st[j] == CF_C) {
            pci[o] = mapc[j];
            pv[o] = -v[p] / den;
            ++o;
        }
    }
}

__global__ void k_cf_cflag(const uint8_t* st, uint32_t n, uint32_t* f) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = st[i] == CF_C ? 1u : 0u;
    if (i == n) f[n] = 0u;
}

}  // namespace

AmgDevHierarchy build_cf_hierarchy_device(const CsrMat& A0, const fem_pc_params& prm, hipStream_t s,
                                          std::vector<std::vector<uint8_t>>* splits) {
    AmgDevHierarchy H;
    SetupTrace tr("device C/F", s);
    const double theta = prm.theta;
    const int maxl = std::max(1, prm.max_levels);
    const uint32_t cth = (uint32_t)std::max(1, prm.coarse_size);
    DCsr cur;
    cur.n = A0.n;
    cur.m = A0.m;
    cur.nnz = A0.nnz;
    cur.rp.alloc(A0.n + 1);
    cur.alloc_entries(s);
    FEM_HIP(hipMemcpyAsync(cur.rp.get(), A0.rp.get(), (size_t)(A0.n + 1) * 4, hipMemcpyDeviceToDevice, s));
    FEM_HIP(hipMemcpyAsync(cur.ci.get(), A0.ci.get(), A0.nnz * 4, hipMemcpyDeviceToDevice, s));
    FEM_HIP(hipMemcpyAsync(cur.v.get(), A0.v.get(), A0.nnz * 8, hipMemcpyDeviceToDevice, s));
    DevBuf<uint32_t> cnt1(1);
    for (int k = 0;; ++k) {
        const uint32_t n = cur.n;
        auto lv = std::make_unique<AmgLevelDev>();
        lv->n = n;
        lv->omega = prm.omega;
        DevBuf<double> d(n);
        hipLaunchKernelGGL(k_diag_val, dim3(g1(n)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), cur.v.get(), n,
                           d.get());
        auto keep_as_coarsest = [&] {
            cur.to_mat(lv->A, s);
            H.lv.push_back(std::move(lv));
        };
        if (n <= cth || k + 1 >= maxl) { keep_as_coarsest(); break; }
        // strength graph S and its transpose
        DevBuf<uint8_t> sm(std::max<uint64_t>(cur.nnz, 1));
        DevBuf<uint32_t> scnt(n + 1), srp(n + 1);
        hipLaunchKernelGGL(k_cf_strength, dim3(g1(n + 1)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), cur.v.get(),
                           theta, n, sm.get(), scnt.get());
        scan_u32(scnt.get(), srp.get(), (size_t)n + 1, s);
        const uint32_t snz = read_u32(srp.get() + n, s);
        DevBuf<uint32_t> sci(std::max<uint32_t>(snz, 1));
        hipLaunchKernelGGL(k_cf_sfill, dim3(g1(n)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), sm.get(), n,
                           srp.get(), sci.get());
        DevBuf<double> zeros(std::max<uint32_t>(snz, 1));
        FEM_HIP(hipMemsetAsync(zeros.get(), 0, (size_t)std::max<uint32_t>(snz, 1) * 8, s));
        DCsr St = transpose_arrays(srp.get(), sci.get(), zeros.get(), n, n, snz, s);
        tr.lap("strength");
        // colouring of S + S^T (several rounds per host check)
        DevBuf<uint8_t> colour(n);
        FEM_HIP(hipMemsetAsync(colour.get(), kNoColour, n, s));
        for (uint32_t round = 0;; round += 16) {
            if (round > n + 16) fail(FEM_ERR_RUNTIME, "AMG setup: colouring did not finish");
            for (int r = 0; r < 16; ++r) {
                FEM_HIP(hipMemsetAsync(cnt1.get(), 0, 4, s));
                hipLaunchKernelGGL(k_cf_colour, dim3(g1(n)), dim3(256), 0, s, srp.get(), sci.get(), St.rp.get(),
                                   St.ci.get(), n, colour.get(), cnt1.get());
            }
            if (read_u32(cnt1.get(), s) == 0) break;
        }
        tr.lap("colouring");
        // C/F split in rounds
        DevBuf<uint8_t> sa(n), sb(n);
        FEM_HIP(hipMemsetAsync(sa.get(), CF_U, n, s));
        uint8_t* st = sa.get();
        uint8_t* st2 = sb.get();
        for (uint32_t round = 0;; ++round) {
            if (round > n) fail(FEM_ERR_RUNTIME, "AMG setup: C/F splitting did not finish");
            FEM_HIP(hipMemsetAsync(cnt1.get(), 0, 4, s));
            hipLaunchKernelGGL(k_cf_select, dim3(g1(n)), dim3(256), 0, s, srp.get(), sci.get(), St.rp.get(),
                               St.ci.get(), colour.get(), st, n, st2);
            hipLaunchKernelGGL(k_cf_fmark, dim3(g1(n)), dim3(256), 0, s, srp.get(), sci.get(), n, st2, cnt1.get());
            std::swap(st, st2);
            if (read_u32(cnt1.get(), s) == 0) break;
        }
        DevBuf<uint32_t> f(n + 1), mapc(n + 1);
        hipLaunchKernelGGL(k_cf_cflag, dim3(g1(n + 1)), dim3(256), 0, s, st, n, f.get());
        scan_u32(f.get(), mapc.get(), (size_t)n + 1, s);
        const uint32_t nc = read_u32(mapc.get() + n, s);
        tr.lap("C/F split");
        if (nc == 0 || nc == n) { keep_as_coarsest(); break; }
        // direct interpolation
        DCsr P;
        P.n = n;
        P.m = nc;
        P.rp.alloc(n + 1);
        FEM_HIP(hipMemsetAsync(cnt1.get(), 0, 4, s));
        hipLaunchKernelGGL(k_cf_pcount, dim3(g1(n + 1)), dim3(256), 0, s, srp.get(), sci.get(), st, n, f.get(),
                           cnt1.get());
        if (read_u32(cnt1.get(), s) != 0) fail(FEM_ERR_RUNTIME, "AMG setup: an F point has no strong C neighbour");
        scan_u32(f.get(), P.rp.get(), (size_t)n + 1, s);
        P.nnz = read_u32(P.rp.get() + n, s);
        P.alloc_entries(s);
        hipLaunchKernelGGL(k_cf_pfill, dim3(g1(n)), dim3(256), 0, s, cur.rp.get(), cur.ci.get(), cur.v.get(), sm.get(),
                           st, mapc.get(), d.get(), n, P.rp.get(), P.ci.get(), P.v.get());
        FEM_HIP(hipGetLastError());
        if (splits) {
            splits->emplace_back(n);
            FEM_HIP(hipMemcpyAsync(splits->back().data(), st, n, hipMemcpyDeviceToHost, s));
        }
        tr.lap("interpolation");
        DCsr R = transpose_dev(P, s);
        DCsr AP = spgemm_dev(cur, P, s);
        DCsr Ac = spgemm_dev(R, AP, s);
        tr.lap("Galerkin products");
        lv->nc = nc;
        lv->d = std::move(d);
        cur.to_mat(lv->A, s);
        P.to_mat(lv->P, s);
        R.to_mat(lv->R, s);
        H.lv.push_back(std::move(lv));
        cur = std::move(Ac);
    }
    H.coarse = sa_coarse_operator_host(to_host(H.lv.back()->A), true);
    tr.lap("coarsest operator");
    return H;
}

// C = A B on the device (columns of each row ascending; every structurally
// reached column kept). B's rows must list strictly increasing columns.
void spgemm_device(const CsrMat& A, const CsrMat& B, CsrMat& C, hipStream_t s) {
    DCsr c = spgemm_arrays(A.n, A.m, A.rp.get(), A.ci.get(), A.v.get(), B.n, B.m, B.rp.get(), B.ci.get(), B.v.get(), s);
    c.to_mat(C, s);
}

}  // namespace fem

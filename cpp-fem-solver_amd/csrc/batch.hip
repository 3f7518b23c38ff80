// batch.hip — many independent SPD systems solved in one persistent launch
// (no counterpart in the reference; each system is solved as the reference's
// block-Jacobi preconditioned CG, blockMatrix.hpp:714-735 with the cached
// inverses of test_BlockMatrix_mpi.cpp:132-155).
//
// A workgroup takes the next system from an atomic work queue and solves it
// completely — all right-hand sides together, float64 iterative refinement
// around a float32 block-Jacobi PCG — with workgroup barriers and LDS
// reductions only; then it takes the next one. The queue counter only grows,
// so every workgroup leaves the loop once it passes the number of systems.
//
// Layout (all systems concatenated): BSR arrays (row pointers local to each
// system), float64 and float32 block values, float32 inverse diagonal
// blocks; vectors interleaved by right-hand side, entry (row i, rhs c) at
// [i * NR + c], NR = 4 slots (unused slots stay zero).
//
// Per outer step: r = b - A x in float64 (converted to float32), inner PCG on
// A d = r (float32, stop at ||r_k|| < max(inner_tol, tol ||b|| / (2 ||r||))
// ||r||), x += d in float64; converged when ||b - A x|| < tol ||b|| for every
// right-hand side. Per inner iteration three passes over the rows:
//   A: q = A p, partial p.q;  B: d += a p, r -= a q, partial r.r, r.z;
//   C: p = z + beta p (z = D^{-1} r recomputed).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <numeric>

#include "handles.hpp"
#include "reduce.hpp"

namespace fem {
namespace {

constexpr int NR = 4;  // right-hand-side slots per system

struct BSys {
    int bs;
    int nrhs;
    uint32_t nb;      // block rows
    uint64_t rp;      // offset into rp (nb + 1 entries, local values)
    uint64_t ci;      // offset into ci (blocks)
    uint64_t val;     // offset into values (scalars)
    uint64_t dinv;    // offset into dinv (scalars)
    uint64_t vec;     // offset into vectors (scalars * NR)
    int cap;          // > 0: at most this many inner iterations (autotuner trials)
    int nr;           // vector layout: 4 interleaved right-hand-side slots, or 1
    int maxrow;       // most blocks in a block row
};

// scalar systems whose rows hold at most this many entries take the product
// path that issues all of a step's loads before any use
constexpr int kShortRow = 5;

struct BOut {
    int iterations;   // inner iterations (all right-hand sides advance together)
    int refinements;
    int converged;    // number of converged right-hand sides
    int pad;
    double relres[NR];
    long long t0, t1;  // device wall clock when the item was taken and finished
};

struct BArgs {
    const BSys* sys;
    const int* order;
    int nsys;
    int* queue;
    const uint32_t* rp;
    const uint32_t* ci;
    const double* v64;
    const float* v32;
    const float* dinv;
    const double* b;
    double* x;
    float *r, *d, *p, *q;
    BOut* out;
    double tol, inner_tol;
    int max_refine, max_inner;
};

// deterministic workgroup sum of K doubles per thread: wave butterflies, one
// LDS row per wave, then threads 0..K-1 add the rows in wave order and every
// thread reads the K totals. `red` holds (BLOCK/64 + 1) * K doubles and
// alternates between calls, so a buffer is not rewritten before all threads
// have read it.
template <int BLOCK, int K>
__device__ __forceinline__ void wg_sum(double (&v)[K], double* red) {
    constexpr int NW = BLOCK / 64;
#pragma unroll
    for (int k = 0; k < K; ++k)
        v[k] = wave_sum(v[k]);
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) red[w * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0.0;
#pragma unroll 1
        for (int ww = 0; ww < NW; ++ww) s += red[ww * K + threadIdx.x];
        red[NW * K + threadIdx.x] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = red[NW * K + k];
}

template <int BS, int NRT>
__device__ __forceinline__ void dinv_apply(const float* Di, const float (&r)[BS][NRT], float (&z)[BS][NRT]) {
#pragma unroll
    for (int a = 0; a < BS; ++a)
#pragma unroll
        for (int c = 0; c < NRT; ++c) {
            float s = 0.f;
#pragma unroll
            for (int b = 0; b < BS; ++b) s += Di[a * BS + b] * r[b][c];
            z[a][c] = s;
        }
}

// block row I of a vector: BS rows of NRT interleaved right-hand sides
template <int BS, int NRT>
__device__ __forceinline__ void load_blk(const float* v, uint32_t I, float (&o)[BS][NRT]) {
    if constexpr (NRT == 4) {
        const float4* p = reinterpret_cast<const float4*>(v + (size_t)I * BS * 4);
#pragma unroll
        for (int a = 0; a < BS; ++a) {
            const float4 t = p[a];
            o[a][0] = t.x; o[a][1] = t.y; o[a][2] = t.z; o[a][3] = t.w;
        }
    } else {
#pragma unroll
        for (int a = 0; a < BS; ++a) o[a][0] = v[(size_t)I * BS + a];
    }
}

template <int BS, int NRT>
__device__ __forceinline__ void store_blk(float* v, uint32_t I, const float (&o)[BS][NRT]) {
    if constexpr (NRT == 4) {
        float4* p = reinterpret_cast<float4*>(v + (size_t)I * BS * 4);
#pragma unroll
        for (int a = 0; a < BS; ++a) p[a] = make_float4(o[a][0], o[a][1], o[a][2], o[a][3]);
    } else {
#pragma unroll
        for (int a = 0; a < BS; ++a) v[(size_t)I * BS + a] = o[a][0];
    }
}

// per-system scalars that stay fixed through an inner solve, kept in LDS
// (written by thread 0 between barriers) rather than in every thread's
// registers
struct SysShared {
    double normb[NR], relres[NR], rr0[NR];
    float want[NR];
};

template <int BLOCK, int BS, int NRT>
__device__ void solve_system(const BArgs& A, const BSys& S, double* red, SysShared& sh, BOut* o) {
    constexpr int BV = BS * BS;
    constexpr int U = BS == 1 ? (NRT == 1 ? 4 : 2) : 1;  // rows per thread per step
    const uint32_t nb = S.nb;
    const uint32_t* rp = A.rp + S.rp;
    const uint32_t* ci = A.ci + S.ci;
    const double* v64 = A.v64 + S.val;
    const float* v32 = A.v32 + S.val;
    const float* dv = A.dinv + S.dinv;
    const double* b = A.b + S.vec;
    double* x = A.x + S.vec;
    float* __restrict__ r = A.r + S.vec;
    float* __restrict__ d = A.d + S.vec;
    float* __restrict__ p = A.p + S.vec;
    float* __restrict__ q = A.q + S.vec;
    const uint32_t tid = threadIdx.x;
    int rb = 0;  // alternating reduction buffer
    auto RED = [&]() { double* p_ = red + (rb & 1) * (BLOCK / 64 + 1) * 8; ++rb; return p_; };

    // ||b||, x = 0
    double nb2[NRT];
    for (int c = 0; c < NRT; ++c) nb2[c] = 0.0;
    for (uint32_t i = tid; i < nb * BS; i += BLOCK)
        for (int c = 0; c < NRT; ++c) {
            const double t = b[(size_t)i * NRT + c];
            nb2[c] += t * t;
            x[(size_t)i * NRT + c] = 0.0;
        }
    wg_sum<BLOCK, NRT>(nb2, RED());
    if (tid == 0)
        for (int c = 0; c < NRT; ++c) {
            sh.normb[c] = sqrt(nb2[c]);
            sh.relres[c] = 0.0;
        }
    int iters = 0, refs = 0;
    bool done = false;
    for (int ref = 0; !done; ++ref) {
        // r = b - A x (float64), stored as float32; ||r||^2
        __syncthreads();
        double rn2[NRT] = {};
        for (uint32_t I = tid; I < nb; I += BLOCK) {
            double acc[BS][NRT];
            for (int a = 0; a < BS; ++a)
                for (int c = 0; c < NRT; ++c) acc[a][c] = b[((size_t)I * BS + a) * NRT + c];
            for (uint32_t k = rp[I]; k < rp[I + 1]; ++k) {
                const uint32_t J = ci[k];
                const double* blk = v64 + (size_t)k * BV;
                for (int bb = 0; bb < BS; ++bb) {
                    double xv[NRT];
                    for (int c = 0; c < NRT; ++c) xv[c] = x[((size_t)J * BS + bb) * NRT + c];
                    for (int a = 0; a < BS; ++a)
                        for (int c = 0; c < NRT; ++c) acc[a][c] -= blk[a * BS + bb] * xv[c];
                }
            }
            float rf[BS][NRT];
            for (int a = 0; a < BS; ++a)
                for (int c = 0; c < NRT; ++c) {
                    rn2[c] += acc[a][c] * acc[a][c];
                    rf[a][c] = (float)acc[a][c];
                }
            store_blk<BS, NRT>(r, I, rf);
        }
        wg_sum<BLOCK, NRT>(rn2, RED());  // ends with a barrier after sh was last written
        bool all = true;
        bool act[NRT];
        for (int c = 0; c < NRT; ++c) {
            const double rn = sqrt(rn2[c]);
            const double nbc = sh.normb[c];
            act[c] = c < S.nrhs && nbc > 0 && rn >= A.tol * nbc;
            if (act[c]) all = false;
        }
        __syncthreads();  // every thread has read sh before thread 0 rewrites it
        if (tid == 0)
            for (int c = 0; c < NRT; ++c) {
                const double rn = sqrt(rn2[c]);
                const double nbc = sh.normb[c];
                sh.relres[c] = nbc > 0 ? rn / nbc : 0.0;
                sh.want[c] = act[c] ? (float)fmax(A.inner_tol, 0.5 * A.tol * nbc / rn) : 0.f;
            }
        if (all || ref >= A.max_refine) {
            done = true;
            break;
        }
        ++refs;
        // inner PCG: d = 0, z = D^-1 r, p = z, rz = r.z, rr0 = ||r||
        double t2[2 * NRT];
        for (int k = 0; k < 2 * NRT; ++k) t2[k] = 0.0;
        for (uint32_t I = tid; I < nb; I += BLOCK) {
            float rr[BS][NRT], zz[BS][NRT], zero[BS][NRT];
            load_blk<BS, NRT>(r, I, rr);
            dinv_apply<BS, NRT>(dv + (size_t)I * BV, rr, zz);
            for (int a = 0; a < BS; ++a)
                for (int c = 0; c < NRT; ++c) {
                    zero[a][c] = 0.f;
                    t2[c] += (double)rr[a][c] * zz[a][c];
                    t2[NRT + c] += (double)rr[a][c] * rr[a][c];
                }
            store_blk<BS, NRT>(p, I, zz);
            store_blk<BS, NRT>(d, I, zero);
        }
        wg_sum<BLOCK, 2 * NRT>(t2, RED());
        double rz[NRT];
        for (int c = 0; c < NRT; ++c) {
            rz[c] = t2[c];
            if (!(rz[c] > 0.0)) act[c] = false;
        }
        if (tid == 0)
            for (int c = 0; c < NRT; ++c) sh.rr0[c] = sqrt(t2[NRT + c]);
        __syncthreads();
        const int max_inner = S.cap > 0 ? min(A.max_inner, S.cap) : A.max_inner;
        for (int it = 0; it < max_inner; ++it) {
            bool any = false;
            for (int c = 0; c < NRT; ++c) any |= act[c];
            if (!any) break;
            ++iters;
            __syncthreads();  // p complete
            // pass A: q = A p, p.q. U rows per thread at a time: their row
            // pointers, entries and gathers are all issued before any use,
            // so a thread keeps U dependent load chains in flight.
            double pq[NRT] = {};
            for (uint32_t base = tid; base < nb; base += U * BLOCK) {
                uint32_t k0[U], k1[U], len = 0;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t I = base + u * BLOCK;
                    k0[u] = I < nb ? rp[I] : 0u;
                    k1[u] = I < nb ? rp[I + 1] : 0u;
                    len = max(len, k1[u] - k0[u]);
                }
                float acc[U][BS][NRT];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    for (int a = 0; a < BS; ++a)
                        for (int c = 0; c < NRT; ++c) acc[u][a][c] = 0.f;
                bool short_rows = false;
                if constexpr (BS == 1 && NRT == 1) short_rows = S.maxrow <= kShortRow;
                if (short_rows) {
                    // three dependent rounds per step (row pointers, then every
                    // entry and column, then every gather) instead of two per
                    // entry; an index past a row's end re-reads its first entry
                    // (rows past nb read entry 0) and its term is not added, so
                    // the sums are the generic loop's
                    uint32_t J[U][kShortRow];
                    float vv[U][kShortRow], pj[U][kShortRow];
#pragma unroll
                    for (int u = 0; u < U; ++u)
#pragma unroll
                        for (int t = 0; t < kShortRow; ++t) {
                            const uint32_t k = k0[u] + ((uint32_t)t < k1[u] - k0[u] ? (uint32_t)t : 0u);
                            J[u][t] = ci[k];
                            vv[u][t] = v32[k];
                        }
#pragma unroll
                    for (int u = 0; u < U; ++u)
#pragma unroll
                        for (int t = 0; t < kShortRow; ++t) pj[u][t] = p[J[u][t]];
#pragma unroll
                    for (int u = 0; u < U; ++u)
#pragma unroll
                        for (int t = 0; t < kShortRow; ++t)
                            if ((uint32_t)t < k1[u] - k0[u]) acc[u][0][0] += vv[u][t] * pj[u][t];
                    len = 0;
                }
                for (uint32_t t = 0; t < len; ++t) {
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        if (k0[u] + t >= k1[u]) continue;
                        const uint32_t k = k0[u] + t;
                        const uint32_t J = ci[k];
                        const float* blk = v32 + (size_t)k * BV;
                        float pj[BS][NRT];
                        load_blk<BS, NRT>(p, J, pj);
                        for (int a = 0; a < BS; ++a)
                            for (int bb = 0; bb < BS; ++bb)
                                for (int c = 0; c < NRT; ++c) acc[u][a][c] += blk[a * BS + bb] * pj[bb][c];
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t I = base + u * BLOCK;
                    if (I >= nb) continue;
                    float pi[BS][NRT];
                    load_blk<BS, NRT>(p, I, pi);
                    for (int a = 0; a < BS; ++a)
                        for (int c = 0; c < NRT; ++c) pq[c] += (double)pi[a][c] * acc[u][a][c];
                    store_blk<BS, NRT>(q, I, acc[u]);
                }
            }
            wg_sum<BLOCK, NRT>(pq, RED());
            float alpha[NRT];
            for (int c = 0; c < NRT; ++c) {
                if (act[c] && !(pq[c] > 0.0)) act[c] = false;  // breakdown: keep the iterate
                alpha[c] = act[c] ? (float)(rz[c] / pq[c]) : 0.f;
            }
            // pass B: d += a p, r -= a q, r.r, r.z (U rows per step, loads first)
            double t[2 * NRT];
            for (int k = 0; k < 2 * NRT; ++k) t[k] = 0.0;
            for (uint32_t base = tid; base < nb; base += U * BLOCK) {
                float pi[U][BS][NRT], qi[U][BS][NRT], ri[U][BS][NRT], di[U][BS][NRT];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t I = base + u * BLOCK;
                    if (I >= nb) continue;
                    load_blk<BS, NRT>(p, I, pi[u]);
                    load_blk<BS, NRT>(q, I, qi[u]);
                    load_blk<BS, NRT>(r, I, ri[u]);
                    load_blk<BS, NRT>(d, I, di[u]);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t I = base + u * BLOCK;
                    if (I >= nb) continue;
                    float zz[BS][NRT];
                    for (int a = 0; a < BS; ++a)
                        for (int c = 0; c < NRT; ++c) {
                            di[u][a][c] += alpha[c] * pi[u][a][c];
                            ri[u][a][c] -= alpha[c] * qi[u][a][c];
                        }
                    dinv_apply<BS, NRT>(dv + (size_t)I * BV, ri[u], zz);
                    for (int a = 0; a < BS; ++a)
                        for (int c = 0; c < NRT; ++c) {
                            t[c] += (double)ri[u][a][c] * ri[u][a][c];
                            t[NRT + c] += (double)ri[u][a][c] * zz[a][c];
                        }
                    store_blk<BS, NRT>(d, I, di[u]);
                    store_blk<BS, NRT>(r, I, ri[u]);
                }
            }
            wg_sum<BLOCK, 2 * NRT>(t, RED());
            float beta[NRT];
            for (int c = 0; c < NRT; ++c) {
                if (act[c] && sqrt(t[c]) < sh.want[c] * sh.rr0[c]) act[c] = false;
                beta[c] = (act[c] && rz[c] != 0.0) ? (float)(t[NRT + c] / rz[c]) : 0.f;
                if (act[c]) rz[c] = t[NRT + c];
                if (act[c] && !(rz[c] > 0.0)) act[c] = false;
            }
            // pass C: p = z + beta p (rows of finished right-hand sides keep p)
            for (uint32_t base = tid; base < nb; base += U * BLOCK) {
                float ri[U][BS][NRT], pi[U][BS][NRT];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t I = base + u * BLOCK;
                    if (I >= nb) continue;
                    load_blk<BS, NRT>(r, I, ri[u]);
                    load_blk<BS, NRT>(p, I, pi[u]);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t I = base + u * BLOCK;
                    if (I >= nb) continue;
                    float zz[BS][NRT];
                    dinv_apply<BS, NRT>(dv + (size_t)I * BV, ri[u], zz);
                    for (int a = 0; a < BS; ++a)
                        for (int c = 0; c < NRT; ++c) pi[u][a][c] = act[c] ? zz[a][c] + beta[c] * pi[u][a][c] : pi[u][a][c];
                    store_blk<BS, NRT>(p, I, pi[u]);
                }
            }
        }
        __syncthreads();
        // x += d (float64)
        for (uint32_t i = tid; i < nb * BS; i += BLOCK)
            for (int c = 0; c < NRT; ++c) x[(size_t)i * NRT + c] += (double)d[(size_t)i * NRT + c];
    }
    if (tid == 0) {
        o->iterations = iters;
        o->refinements = refs;
        int conv = 0;
        for (int c = 0; c < S.nrhs; ++c) conv += (sh.relres[c] < A.tol) ? 1 : 0;
        o->converged = conv;
        for (int c = 0; c < NR; ++c) o->relres[c] = c < NRT ? sh.relres[c] : 0.0;
    }
    __syncthreads();
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch_pcg(BArgs a) {
    __shared__ double red[2 * (BLOCK / 64 + 1) * 8];
    __shared__ SysShared sh;
    __shared__ int slot;
    while (true) {
        __syncthreads();
        if (threadIdx.x == 0) slot = atomicAdd(a.queue, 1);
        __syncthreads();
        const int k = slot;
        if (k >= a.nsys) break;
        const int s = a.order[k];
        const BSys S = a.sys[s];
        const long long t0 = wall_clock64();
        if (S.bs == 3) {
            if (S.nr == 1) solve_system<BLOCK, 3, 1>(a, S, red, sh, a.out + s);
            else solve_system<BLOCK, 3, NR>(a, S, red, sh, a.out + s);
        } else {
            if (S.nr == 1) solve_system<BLOCK, 1, 1>(a, S, red, sh, a.out + s);
            else solve_system<BLOCK, 1, NR>(a, S, red, sh, a.out + s);
        }
        if (threadIdx.x == 0) {
            a.out[s].t0 = t0;
            a.out[s].t1 = wall_clock64();
        }
    }
}

// host-side Gauss-Jordan inverse of a BS x BS block (blockMatrix.hpp:70-100)
template <int BS>
bool blk_inv(const double* M, double* Mi) {
    double a[BS][2 * BS];
    for (int i = 0; i < BS; ++i)
        for (int j = 0; j < 2 * BS; ++j) a[i][j] = j < BS ? M[i * BS + j] : (j - BS == i ? 1.0 : 0.0);
    for (int c = 0; c < BS; ++c) {
        int piv = c;
        for (int i = c + 1; i < BS; ++i)
            if (std::fabs(a[i][c]) > std::fabs(a[piv][c])) piv = i;
        if (std::fabs(a[piv][c]) < 1e-300) return false;
        if (piv != c)
            for (int j = 0; j < 2 * BS; ++j) std::swap(a[c][j], a[piv][j]);
        const double inv = 1.0 / a[c][c];
        for (int j = 0; j < 2 * BS; ++j) a[c][j] *= inv;
        for (int i = 0; i < BS; ++i) {
            if (i == c) continue;
            const double f = a[i][c];
            for (int j = 0; j < 2 * BS; ++j) a[i][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < BS; ++i)
        for (int j = 0; j < BS; ++j) Mi[i * BS + j] = a[i][BS + j];
    return true;
}

}  // namespace
}  // namespace fem

using namespace fem;

struct fem_batch_s {
    hipStream_t stream = nullptr;
    int nsys = 0;
    std::vector<BSys> h_sys;            // the systems (all right-hand sides interleaved)
    std::vector<double> work;           // work estimate per system (static, then from the profile)
    // measured device time of a system's item (unsplit) and of its
    // right-hand-side items (split) under the tuned schedule; 0 = not
    // measured. They order the queue when present.
    std::vector<double> time_whole, time_rhs;
    std::vector<uint64_t> nnzb;         // stored blocks per system
    std::vector<char> split;            // system solved as one work item per right-hand side
    // work items: a whole system, or one right-hand side of a split system
    // (vectors of that right-hand side stored contiguously in the system's
    // vector range)
    std::vector<BSys> items;
    std::vector<int> item_sys, item_rhs;
    std::vector<double> item_work;
    std::vector<double> hb;             // right-hand sides as given (for re-layout)
    DevBuf<BSys> sys;                   // items on the device
    DevBuf<int> order, queue;
    DevBuf<uint32_t> rp, ci;
    DevBuf<double> v64, b, x;
    DevBuf<float> v32, dinv, r, d, p, q;
    DevBuf<BOut> out;
    uint64_t nvec = 0;                  // scalars * NR
    // schedule: block size, workgroups per CU, queue order (0 largest work
    // first, 1 input order), number of largest systems split per right-hand side
    int block = 256, wg_per_cu = 0, order_kind = 0, nsplit = 0;
    ~fem_batch_s() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void set_order(fem_batch_s* h, int kind) {
    std::vector<int> ord(h->items.size());
    std::iota(ord.begin(), ord.end(), 0);
    if (kind == 0)
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return h->item_work[a] > h->item_work[b]; });
    h->order.upload(ord, h->stream);
    h->order_kind = kind;
}

// the right-hand sides in the current layout
void upload_b(fem_batch_s* h) {
    std::vector<double> v(h->nvec, 0.0);
    if (!h->hb.empty()) {
        const double* src = h->hb.data();
        for (int k = 0; k < h->nsys; ++k) {
            const BSys& S = h->h_sys[k];
            const uint64_t n = (uint64_t)S.nb * S.bs;
            for (int c = 0; c < S.nrhs; ++c)
                for (uint64_t i = 0; i < n; ++i) {
                    const uint64_t at = h->split[k] ? S.vec + (uint64_t)c * n + i : S.vec + i * NR + c;
                    v[at] = src[(uint64_t)c * n + i];
                }
            src += (uint64_t)S.nrhs * n;
        }
    }
    h->b.upload(v, h->stream);
    FEM_HIP(hipStreamSynchronize(h->stream));
}

// split the `count` systems of largest work (ties: lower index), rebuild the
// work items, re-lay the right-hand sides and the queue order
void apply_split(fem_batch_s* h, int count) {
    count = std::max(0, std::min(count, h->nsys));
    std::vector<int> rank(h->nsys);
    std::iota(rank.begin(), rank.end(), 0);
    std::stable_sort(rank.begin(), rank.end(), [&](int a, int b) { return h->work[a] > h->work[b]; });
    h->split.assign(h->nsys, 0);
    for (int t = 0; t < count; ++t)
        if (h->h_sys[rank[t]].nrhs > 1) h->split[rank[t]] = 1;
    h->nsplit = count;
    h->items.clear();
    h->item_sys.clear();
    h->item_rhs.clear();
    h->item_work.clear();
    // measured item times order the queue only when every item of this
    // layout has one (they are device clock ticks, not comparable with the
    // work model)
    bool measured = h->time_whole.size() == (size_t)h->nsys && h->time_rhs.size() == (size_t)h->nsys;
    for (int k = 0; measured && k < h->nsys; ++k)
        measured = h->split[k] ? h->time_rhs[k] > 0 : h->time_whole[k] > 0;
    for (int k = 0; k < h->nsys; ++k) {
        const BSys& S = h->h_sys[k];
        if (!h->split[k]) {
            BSys it = S;
            it.nr = NR;
            h->items.push_back(it);
            h->item_sys.push_back(k);
            h->item_rhs.push_back(-1);
            h->item_work.push_back(measured ? h->time_whole[k] : h->work[k]);
            continue;
        }
        const uint64_t n = (uint64_t)S.nb * S.bs;
        // one right-hand side reads the matrix as often as all four do
        const double bv = (double)S.bs * S.bs;
        const double frac = (bv + 8.0) / (bv + 8.0 * NR);
        for (int c = 0; c < S.nrhs; ++c) {
            BSys it = S;
            it.nrhs = 1;
            it.nr = 1;
            it.vec = S.vec + (uint64_t)c * n;
            h->items.push_back(it);
            h->item_sys.push_back(k);
            h->item_rhs.push_back(c);
            h->item_work.push_back(measured ? h->time_rhs[k] : h->work[k] * frac);
        }
    }
    h->sys.upload(h->items, h->stream);
    set_order(h, h->order_kind);
    upload_b(h);
}

template <int BLOCK>
int max_wg_per_cu() {
    int nb = 0;
    FEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k_batch_pcg<BLOCK>, BLOCK, 0));
    return std::max(1, nb);
}

int cu_count() {
    int dev = 0, cu = 0;
    FEM_HIP(hipGetDevice(&dev));
    FEM_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
    return cu;
}

// one persistent launch over work items sys[0..nitems) (queue reset first);
// returns ms
double launch(fem_batch_s* h, const BSys* sys, const int* order, int nitems, BOut* out, double tol, double inner_tol,
              int max_refine, int max_inner) {
    hipStream_t s = h->stream;
    vec_fill<uint32_t>((uint32_t*)h->queue.get(), 0u, 1, s);
    BArgs a{sys, order, nitems, h->queue.get(), h->rp.get(), h->ci.get(), h->v64.get(), h->v32.get(), h->dinv.get(),
            h->b.get(), h->x.get(), h->r.get(), h->d.get(), h->p.get(), h->q.get(), out, tol, inner_tol, max_refine,
            max_inner};
    int per = h->wg_per_cu;
    int cap;
    switch (h->block) {
        case 512: cap = max_wg_per_cu<512>(); break;
        case 1024: cap = max_wg_per_cu<1024>(); break;
        default: cap = max_wg_per_cu<256>(); break;
    }
    if (per <= 0 || per > cap) per = cap;
    const unsigned grid = (unsigned)std::max(1, std::min(nitems, per * cu_count()));
    hipEvent_t e0, e1;
    FEM_HIP(hipEventCreate(&e0));
    FEM_HIP(hipEventCreate(&e1));
    FEM_HIP(hipEventRecord(e0, s));
    switch (h->block) {
        case 512: hipLaunchKernelGGL(k_batch_pcg<512>, dim3(grid), dim3(512), 0, s, a); break;
        case 1024: hipLaunchKernelGGL(k_batch_pcg<1024>, dim3(grid), dim3(1024), 0, s, a); break;
        default: hipLaunchKernelGGL(k_batch_pcg<256>, dim3(grid), dim3(256), 0, s, a); break;
    }
    FEM_HIP(hipGetLastError());
    FEM_HIP(hipEventRecord(e1, s));
    FEM_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FEM_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

double run_items(fem_batch_s* h, const fem_batch_opts* o, int max_refine) {
    return launch(h, h->sys.get(), h->order.get(), (int)h->items.size(), h->out.get(), o->tol, o->inner_tol,
                  max_refine, o->max_inner);
}

// FEM_TRACE_BATCH: the ten items that finished last and the ten that ran
// longest (device wall clock), to stderr
void trace_items(fem_batch_s* h, const std::vector<BOut>& o, double ms) {
    int dev = 0, khz = 0;
    FEM_HIP(hipGetDevice(&dev));
    FEM_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    if (khz <= 0 || o.empty()) return;
    long long start = o[0].t0;
    for (const BOut& b : o) start = std::min(start, b.t0);
    auto msof = [&](long long t) { return (double)t / khz; };
    std::vector<size_t> id(o.size());
    for (size_t t = 0; t < id.size(); ++t) id[t] = t;
    auto show = [&](const char* what) {
        std::fprintf(stderr, "[femgpu] batch %.1f ms, %s:\n", ms, what);
        for (size_t j = 0; j < std::min<size_t>(10, id.size()); ++j) {
            const size_t t = id[j];
            const BSys& S = h->items[t];
            std::fprintf(stderr, "[femgpu]   item %zu system %d rhs %d bs %d rows %u it %d: %.2f .. %.2f ms\n", t,
                         h->item_sys[t], h->item_rhs[t], S.bs, S.nb, o[t].iterations, msof(o[t].t0 - start),
                         msof(o[t].t1 - start));
        }
    };
    std::sort(id.begin(), id.end(), [&](size_t a, size_t b) { return o[a].t1 > o[b].t1; });
    show("last to finish");
    std::sort(id.begin(), id.end(), [&](size_t a, size_t b) { return o[a].t1 - o[a].t0 > o[b].t1 - o[b].t0; });
    show("longest");
}

// per-system results from the items' (a split system: the largest counts of
// its right-hand sides, each right-hand side's own residual)
std::vector<fem_batch_stats> merge_out(fem_batch_s* h, const std::vector<BOut>& o) {
    std::vector<fem_batch_stats> st(h->nsys);
    for (int k = 0; k < h->nsys; ++k) st[k] = fem_batch_stats{0, 0, 0, 0.0};
    for (size_t t = 0; t < h->items.size(); ++t) {
        fem_batch_stats& S = st[h->item_sys[t]];
        S.iterations = std::max(S.iterations, o[t].iterations);
        S.refinements = std::max(S.refinements, o[t].refinements);
        S.converged += o[t].converged;
        const int nr = h->items[t].nrhs;
        for (int c = 0; c < nr; ++c) S.max_relres = std::max(S.max_relres, o[t].relres[c]);
    }
    return st;
}

std::vector<BOut> download_out(fem_batch_s* h) {
    std::vector<BOut> o(h->items.size());
    h->out.download(o.data(), o.size(), h->stream);
    FEM_HIP(hipStreamSynchronize(h->stream));
    return o;
}

}  // namespace

extern "C" {

int fem_batch_create(int nsys, const fem_batch_system* systems, int nrhs, fem_batch* out) {
    return api_guard([&] {
        if (!systems || !out || nsys < 1) fail(FEM_ERR_INVALID_ARGUMENT, "need at least one system");
        if (nrhs < 1 || nrhs > NR) fail(FEM_ERR_INVALID_ARGUMENT, "nrhs must be 1..4");
        ensure_device();
        auto h = std::make_unique<fem_batch_s>();
        h->stream = new_stream();
        h->nsys = nsys;
        std::vector<uint32_t> rp, ci;
        std::vector<double> v64, dv;
        uint64_t vec = 0;
        for (int k = 0; k < nsys; ++k) {
            const fem_batch_system& in = systems[k];
            const int bs = in.bs;
            if (bs != 1 && bs != 3) fail(FEM_ERR_INVALID_ARGUMENT, "batch systems need block size 1 or 3");
            if (!in.row_ptr || !in.col_idx || !in.values || in.nb == 0)
                fail(FEM_ERR_INVALID_ARGUMENT, "system " + std::to_string(k) + ": empty or null arrays");
            if (in.row_ptr[0] != 0) fail(FEM_ERR_INVALID_ARGUMENT, "row_ptr must start at 0");
            const uint32_t nb = in.nb;
            const uint64_t nnzb = in.row_ptr[nb];
            BSys S{bs, nrhs, nb, rp.size(), ci.size(), v64.size(), dv.size(), vec, 0, NR};
            const int BV = bs * bs;
            for (uint32_t I = 0; I <= nb; ++I) rp.push_back(in.row_ptr[I]);
            std::vector<double> inv((size_t)nb * BV, 0.0);
            for (uint32_t I = 0; I < nb; ++I) {
                if (in.row_ptr[I + 1] < in.row_ptr[I]) fail(FEM_ERR_INVALID_ARGUMENT, "row_ptr must be non-decreasing");
                int64_t dp = -1;
                for (uint32_t t = in.row_ptr[I]; t < in.row_ptr[I + 1]; ++t) {
                    if (in.col_idx[t] >= nb)
                        fail(FEM_ERR_OUT_OF_RANGE, "system " + std::to_string(k) + ": column index out of range");
                    if (in.col_idx[t] == I) dp = t;
                }
                if (dp < 0) fail(FEM_ERR_RUNTIME, "system " + std::to_string(k) + ": block row " + std::to_string(I) + " has no diagonal block");
                const bool ok = bs == 3 ? blk_inv<3>(in.values + (size_t)dp * BV, &inv[(size_t)I * BV])
                                        : blk_inv<1>(in.values + (size_t)dp * BV, &inv[(size_t)I * BV]);
                if (!ok) fail(FEM_ERR_RUNTIME, "system " + std::to_string(k) + ": singular diagonal block " + std::to_string(I));
            }
            ci.insert(ci.end(), in.col_idx, in.col_idx + nnzb);
            v64.insert(v64.end(), in.values, in.values + nnzb * BV);
            dv.insert(dv.end(), inv.begin(), inv.end());
            vec += (uint64_t)nb * bs * NR;
            for (uint32_t I = 0; I < nb; ++I)
                S.maxrow = std::max(S.maxrow, (int)(in.row_ptr[I + 1] - in.row_ptr[I]));
            h->h_sys.push_back(S);
            h->nnzb.push_back(nnzb);
            // work estimate: matrix entries x sqrt(rows) (CG iterations grow
            // with the square root of the condition number)
            h->work.push_back((double)nnzb * BV * std::sqrt((double)nb * bs));
        }
        hipStream_t s = h->stream;
        h->nvec = vec;
        h->rp.upload(rp, s);
        h->ci.upload(ci, s);
        h->v64.upload(v64, s);
        std::vector<float> f32(v64.begin(), v64.end()), d32(dv.begin(), dv.end());
        h->v32.upload(f32, s);
        h->dinv.upload(d32, s);
        h->b.alloc(vec);
        h->x.alloc(vec);
        h->r.alloc(vec);
        h->d.alloc(vec);
        h->p.alloc(vec);
        h->q.alloc(vec);
        vec_fill<double>(h->x.get(), 0.0, vec, s);
        vec_fill<float>(h->r.get(), 0.f, vec, s);
        vec_fill<float>(h->d.get(), 0.f, vec, s);
        vec_fill<float>(h->p.get(), 0.f, vec, s);
        vec_fill<float>(h->q.get(), 0.f, vec, s);
        h->out.alloc((size_t)nsys * NR);
        h->queue.alloc(1);
        apply_split(h.get(), 0);
        FEM_HIP(hipStreamSynchronize(s));
        *out = h.release();
    });
}

int fem_batch_destroy(fem_batch h) {
    return api_guard([&] { delete h; });
}

int fem_batch_set_rhs(fem_batch h, const double* B) {
    return api_guard([&] {
        if (!h || !B) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        uint64_t total = 0;
        for (const BSys& S : h->h_sys) total += (uint64_t)S.nrhs * S.nb * S.bs;
        h->hb.assign(B, B + total);
        upload_b(h);
    });
}

int fem_batch_get_solution(fem_batch h, double* X) {
    return api_guard([&] {
        if (!h || !X) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        std::vector<double> v(h->nvec);
        h->x.download(v.data(), h->nvec, h->stream);
        FEM_HIP(hipStreamSynchronize(h->stream));
        double* dst = X;
        for (int k = 0; k < h->nsys; ++k) {
            const BSys& S = h->h_sys[k];
            const uint64_t n = (uint64_t)S.nb * S.bs;
            for (int c = 0; c < S.nrhs; ++c)
                for (uint64_t i = 0; i < n; ++i)
                    dst[(uint64_t)c * n + i] = h->split[k] ? v[S.vec + (uint64_t)c * n + i] : v[S.vec + i * NR + c];
            dst += (uint64_t)S.nrhs * n;
        }
    });
}

int fem_batch_solve(fem_batch h, const fem_batch_opts* opts, fem_batch_stats* per_system, double* ms) {
    return api_guard([&] {
        if (!h || !opts) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        if (!(opts->tol > 0) || opts->max_refine < 1 || opts->max_inner < 1)
            fail(FEM_ERR_INVALID_ARGUMENT, "tol > 0, max_refine >= 1 and max_inner >= 1 are required");
        const double t = run_items(h, opts, opts->max_refine);
        if (ms) *ms = t;
        if (std::getenv("FEM_TRACE_BATCH")) trace_items(h, download_out(h), t);
        if (per_system) {
            const std::vector<fem_batch_stats> st = merge_out(h, download_out(h));
            std::copy(st.begin(), st.end(), per_system);
        }
    });
}

// Autotuner (profile-guided): one solve with the current schedule records
// every system's inner iterations, which become the work estimates (inner
// iterations x matrix entries). Each schedule (block size 256/512/1024 x
// workgroups per CU 1/2/occupancy limit x queue order) then runs the whole
// batch, unsplit, with one refinement step of iterations_k / divisor inner
// iterations per system, which keeps the real mix of long and short systems
// at 1/divisor of the cost. The best capped schedule of each block size is
// timed on full solves, unsplit and with the systems whose work exceeds an
// even share per resident workgroup split into one item per right-hand side
// (their chains then run side by side); the fastest is returned.
int fem_batch_tune(fem_batch h, const fem_batch_opts* opts, int divisor, fem_batch_schedule* best, double* trial_ms,
                   int cap) {
    return api_guard([&] {
        if (!h || !opts) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        if (divisor < 1) divisor = 8;
        const int keep_block = h->block, keep_per = h->wg_per_cu, keep_order = h->order_kind;
        apply_split(h, 0);
        run_items(h, opts, opts->max_refine);
        const std::vector<BOut> o = download_out(h);
        // profiled work: inner iterations x bytes of one iteration (matrix in
        // float32 with its indices, about nine vector passes); it ranks the
        // systems for splitting and orders the queue until measured item
        // times (below) replace it
        for (int k = 0; k < h->nsys; ++k) {
            const BSys& S = h->h_sys[k];
            const double bv = (double)S.bs * S.bs;
            h->work[k] = (double)std::max(1, o[k].iterations) *
                         ((double)h->nnzb[k] * (4.0 * bv + 4.0) + (double)S.nb * S.bs * NR * 4.0 * 9.0);
        }
        std::vector<BSys> tsys = h->items;  // unsplit: item k is system k
        for (int k = 0; k < h->nsys; ++k) tsys[k].cap = std::max(1, o[k].iterations / divisor);
        DevBuf<BSys> dsys;
        dsys.upload(tsys, h->stream);
        std::vector<int> ord0(h->nsys), ord1(h->nsys);
        std::iota(ord1.begin(), ord1.end(), 0);
        ord0 = ord1;
        std::stable_sort(ord0.begin(), ord0.end(), [&](int a, int b) { return h->work[a] > h->work[b]; });
        DevBuf<int> dord0, dord1;
        dord0.upload(ord0, h->stream);
        dord1.upload(ord1, h->stream);
        DevBuf<BOut> dout(h->nsys);
        const int blocks[3] = {256, 512, 1024};
        const int pers[3] = {1, 2, 0};
        std::vector<std::pair<double, fem_batch_schedule>> cand;
        int t = 0;
        for (int bi = 0; bi < 3; ++bi)
            for (int pi = 0; pi < 3; ++pi)
                for (int ok = 0; ok < 2; ++ok) {
                    h->block = blocks[bi];
                    h->wg_per_cu = pers[pi];
                    const double ms = launch(h, dsys.get(), ok == 0 ? dord0.get() : dord1.get(), h->nsys, dout.get(),
                                             opts->tol, opts->inner_tol, 1, opts->max_inner);
                    if (trial_ms && t < cap) trial_ms[t] = ms;
                    ++t;
                    cand.push_back({ms, fem_batch_schedule{blocks[bi], pers[pi], ok, 0}});
                }
        // the capped runs rank the schedules by throughput; the makespan of a
        // full run also depends on the longest systems' chains, so the best
        // capped schedule of each workgroup size is timed again on the whole
        // batch, with and without splitting the largest systems
        double total = 0.0;
        for (double w : h->work) total += w;
        const double share = total / std::max(1, 2 * cu_count());
        int nbig = 0;
        for (int k = 0; k < h->nsys; ++k)
            if (h->work[k] > share && h->h_sys[k].nrhs > 1) ++nbig;
        std::vector<fem_batch_schedule> finalists;
        for (int bi = 0; bi < 3; ++bi) {
            int kbest = -1;
            for (int k = 0; k < (int)cand.size(); ++k)
                if (cand[k].second.block == blocks[bi] && (kbest < 0 || cand[k].first < cand[kbest].first)) kbest = k;
            finalists.push_back(cand[kbest].second);
        }
        double best_ms = 1e300;
        fem_batch_schedule bs = finalists[0];
        std::vector<int> splits{0};
        if (nbig > 0) splits.push_back(nbig);
        for (int sp : splits) {
            for (const fem_batch_schedule& c0 : finalists) {
                fem_batch_schedule c = c0;
                c.split = sp;
                h->block = c.block;
                h->wg_per_cu = c.wg_per_cu;
                h->order_kind = c.order;
                apply_split(h, sp);
                const double ms = run_items(h, opts, opts->max_refine);
                if (trial_ms && t < cap) trial_ms[t] = ms;
                ++t;
                if (ms < best_ms) {
                    best_ms = ms;
                    bs = c;
                }
            }
        }
        // queue order from measured item times: run the chosen schedule,
        // record every item's device time (a split system: its slowest
        // right-hand side), order the queue longest first by those times and
        // keep them if that run is faster
        if (bs.order == 0) {
            h->block = bs.block;
            h->wg_per_cu = bs.wg_per_cu;
            h->order_kind = 0;
            apply_split(h, bs.split);
            run_items(h, opts, opts->max_refine);
            const std::vector<BOut> ob = download_out(h);
            std::vector<double> tw(h->nsys, 0.0), tr(h->nsys, 0.0);
            bool ok = true;
            for (size_t it = 0; it < ob.size(); ++it) {
                const double d = (double)(ob[it].t1 - ob[it].t0);
                ok = ok && d > 0;
                const int k = h->item_sys[it];
                if (h->item_rhs[it] < 0) tw[k] = d;
                else tr[k] = std::max(tr[k], d);
            }
            if (ok) {
                std::swap(h->time_whole, tw);
                std::swap(h->time_rhs, tr);
                apply_split(h, bs.split);
                const double ms = run_items(h, opts, opts->max_refine);
                if (trial_ms && t < cap) trial_ms[t] = ms;
                ++t;
                if (!(ms < best_ms)) {  // keep the previous order
                    std::swap(h->time_whole, tw);
                    std::swap(h->time_rhs, tr);
                }
            }
        }
        h->block = keep_block;
        h->wg_per_cu = keep_per;
        h->order_kind = keep_order;
        apply_split(h, 0);
        if (best) *best = bs;
    });
}

int fem_batch_set_schedule(fem_batch h, const fem_batch_schedule* sch) {
    return api_guard([&] {
        if (!h || !sch) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        if (sch->block != 256 && sch->block != 512 && sch->block != 1024)
            fail(FEM_ERR_INVALID_ARGUMENT, "block must be 256, 512 or 1024");
        if (sch->order != 0 && sch->order != 1) fail(FEM_ERR_INVALID_ARGUMENT, "order must be 0 or 1");
        if (sch->split < 0) fail(FEM_ERR_INVALID_ARGUMENT, "split must be >= 0");
        h->block = sch->block;
        h->wg_per_cu = std::max(0, sch->wg_per_cu);
        h->order_kind = sch->order;
        apply_split(h, sch->split);
    });
}

}  // extern "C"

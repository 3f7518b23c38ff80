// precond_tri.hip — IC(0) and ILU(0) preconditioners with level-scheduled
// factorization and triangular solves on the device.
//   make_ic0_preconditioner   ode-solver/src/math/array.cpp:2509-2732
//   make_ilu0_preconditioner  ode-solver/src/math/array.cpp:2733-2934
//
// Setup: the host derives the sparsity patterns (lower part with diagonal for
// IC0; strict-lower L and upper-with-diagonal U for ILU0) and the level sets
// of the triangular dependency graphs (level(i) = 1 + max level of the rows
// row i reads). The numeric factorization runs on the device one level per
// launch, one thread per row, with exactly the reference's per-row
// arithmetic (left-looking two-pointer merges for IC0; the "upper-only" IKJ
// update for ILU0), so the factors are bit-identical to the reference's.
//
// Apply: each sweep stores its rows in level order as SELL-64 (groups of 64
// consecutive level-ordered rows, entries column-major within a group) so
// every load of the sweep is coalesced; each row keeps the reference's
// summation order (ascending column for the forward sweep and for U,
// ascending row for the L^T sweep of IC0). Levels with many rows get one
// multi-block launch; runs of consecutive small levels are processed by a
// single workgroup with a barrier between levels.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "internal.hpp"
#include "precond.hpp"

namespace fem {

constexpr uint32_t kChainRows = 1024;  // levels this small run inside one workgroup
constexpr int kChainThreads = 1024;
constexpr uint32_t kChainLevels = 4096;  // levels per chain launch (LDS level table)
constexpr uint32_t kRing = 4096;         // LDS ring of the most recent sweep results
// wide chains: levels of up to wide_rows(kc) rows, wide_r(kc) rows per thread
// (two levels of rows are held in registers: 4 rows for rows of at most 2
// entries, 2 for at most 4, none above), an LDS ring of kRingW results
// (128 KB of float64)
constexpr uint32_t kRingW = 16384;
constexpr int wide_r(uint32_t kc) { return kc <= 2 ? 4 : kc <= 4 ? 2 : 0; }
constexpr uint32_t wide_rows(uint32_t kc) { return (uint32_t)wide_r(kc) * 1024u; }

// ---------------------------------------------------------------------------
// Level schedule + SELL-64 sweep layout
// ---------------------------------------------------------------------------
struct Segment {
    uint32_t s0, s1;       // slot range
    uint32_t lvl0, lvl1;   // level range (chain when lvl1 - lvl0 > 1)
    uint32_t kc = 0;       // chain: entries per slot in the dense layout
    uint64_t cbase = 0;    // chain: offset of slot s0 in the dense arrays
    bool wide = false;     // chain of levels wider than kChainRows (wide-chain kernel)
};

struct Sweep {
    uint32_t n = 0, nlev = 0;
    std::vector<uint32_t> h_lvl_ptr;    // [nlev+1]
    std::vector<Segment> segs;
    DevBuf<uint32_t> lvl_ptr, slot_row, gptr, col, dslot, vmap, dmap;
    DevBuf<uint16_t> len;
    DevBuf<double> val, diag;
    DevBuf<float> val32, diag32;
    uint64_t nent = 0;  // SELL slots (rows padded to the widest of their group of 64)
    uint64_t ndep = 0;  // dependency entries
    // chains: dense per-slot copies (entry e of slot s at cbase + (s - s0) kc + e)
    // of the columns, dependency slots and values, the SELL index each value
    // comes from, and right-hand-side staging in slot order
    DevBuf<uint32_t> cds, cmap;
    DevBuf<double> sink;     // chain kernels: target of the stores of threads without a row
    DevBuf<float> sink32;
    DevBuf<double> cval, rslot;
    DevBuf<float> cval32, rslot32;
    uint64_t cent = 0;

    // rows: dependency lists per row (entries in summation order), as
    // (col, position-in-factor-values). lev: level per row.
    void build(uint32_t n_, const std::vector<uint32_t>& rp, const std::vector<uint32_t>& cols,
               const std::vector<uint32_t>& pos, const std::vector<uint32_t>& lev, const std::vector<uint32_t>* diagpos,
               hipStream_t s) {
        n = n_;
        nlev = n ? 1 + *std::max_element(lev.begin(), lev.end()) : 0;
        h_lvl_ptr.assign(nlev + 1, 0);
        for (uint32_t i = 0; i < n; ++i) h_lvl_ptr[lev[i] + 1]++;
        for (uint32_t l = 0; l < nlev; ++l) h_lvl_ptr[l + 1] += h_lvl_ptr[l];
        std::vector<uint32_t> order(n), off(h_lvl_ptr.begin(), h_lvl_ptr.end() - 1);
        for (uint32_t i = 0; i < n; ++i) order[off[lev[i]]++] = i;  // stable: ascending row within a level
        const uint32_t ngrp = (n + 63) / 64;
        std::vector<uint32_t> g(ngrp + 1, 0);
        std::vector<uint16_t> ln(n);
        for (uint32_t s = 0; s < n; ++s) {
            const uint32_t i = order[s];
            const uint32_t L = rp[i + 1] - rp[i];
            if (L > 65535) fail(FEM_ERR_RUNTIME, "triangular factor row too long for the sweep layout");
            ln[s] = (uint16_t)L;
        }
        for (uint32_t gi = 0; gi < ngrp; ++gi) {
            uint32_t w = 0;
            for (uint32_t s = gi * 64; s < std::min(n, gi * 64 + 64); ++s) w = std::max<uint32_t>(w, ln[s]);
            g[gi + 1] = g[gi] + w * 64;
        }
        nent = g[ngrp];
        ndep = n ? (uint64_t)rp[n] - rp[0] : 0;
        std::vector<uint32_t> slot_of(n);
        for (uint32_t s = 0; s < n; ++s) slot_of[order[s]] = s;
        std::vector<uint32_t> hc(nent, 0), hm(nent, UINT32_MAX), hd(nent, 0);
        for (uint32_t s = 0; s < n; ++s) {
            const uint32_t i = order[s];
            const uint32_t base = g[s >> 6] + (s & 63);
            for (uint32_t e = 0; e < ln[s]; ++e) {
                hc[base + e * 64] = cols[rp[i] + e];
                hd[base + e * 64] = slot_of[cols[rp[i] + e]];
                hm[base + e * 64] = pos[rp[i] + e];
            }
        }
        // segments: chains of consecutive small levels, big levels alone
        segs.clear();
        // a chain holds only levels whose rows depend on slots at most
        // `ring` behind the level's end (the LDS ring window; the chain
        // kernel preloads the ring with the slots before the chain), so every
        // dependency of a chain is read from LDS
        auto in_window = [&](uint32_t lev, uint32_t ring) {
            const uint32_t end = h_lvl_ptr[lev + 1];
            for (uint32_t q = h_lvl_ptr[lev]; q < end; ++q) {
                const uint32_t i = order[q];
                for (uint32_t t = rp[i]; t < rp[i + 1]; ++t)
                    if (end - slot_of[cols[t]] > ring) return false;
            }
            return true;
        };
        auto width = [&](uint32_t lev) { return h_lvl_ptr[lev + 1] - h_lvl_ptr[lev]; };
        // the wide chains' row limit follows the longest row of the sweep
        uint32_t kcmax = 1;
        for (uint32_t q = 0; q < n; ++q) kcmax = std::max<uint32_t>(kcmax, ln[q]);
        const uint32_t kWideRows = std::max(kChainRows, wide_rows(kcmax <= 2 ? 2u : kcmax <= 4 ? 4u : 8u));
        uint32_t l = 0;
        while (l < nlev) {
            const uint32_t cnt = width(l);
            if (cnt > kWideRows) {
                segs.push_back({h_lvl_ptr[l], h_lvl_ptr[l + 1], l, l + 1});
                ++l;
                continue;
            }
            const bool wide = cnt > kChainRows;
            uint32_t e = l;
            if (wide) {
                while (e < nlev && e - l < kChainLevels && width(e) > kChainRows && width(e) <= kWideRows &&
                       in_window(e, kRingW))
                    ++e;
            } else {
                while (e < nlev && e - l < kChainLevels && width(e) <= kChainRows && in_window(e, kRing)) ++e;
            }
            if (e == l) e = l + 1;  // a level alone (one launch)
            Segment sg{h_lvl_ptr[l], h_lvl_ptr[e], l, e};
            sg.wide = wide && e - l > 1;
            segs.push_back(sg);
            l = e;
        }
        // dense layout of the chains
        std::vector<uint32_t> dd, dm;
        for (Segment& sg : segs) {
            if (sg.lvl1 - sg.lvl0 <= 1) continue;
            uint32_t kc = 1;
            for (uint32_t q = sg.s0; q < sg.s1; ++q) kc = std::max<uint32_t>(kc, ln[q]);
            // the chain kernel's stride is its template width: pad to 2, 4 or 8
            // (wider chains are swept one level per launch, no dense copy)
            if (kc > 8) continue;
            kc = kc <= 2 ? 2u : kc <= 4 ? 4u : 8u;
            sg.kc = kc;
            sg.cbase = dd.size();
            for (uint32_t q = sg.s0; q < sg.s1; ++q) {
                const uint32_t base = g[q >> 6] + (q & 63);
                for (uint32_t e = 0; e < kc; ++e) {
                    const bool real = e < ln[q];
                    dd.push_back(real ? hd[base + e * 64] : 0u);
                    dm.push_back(real ? (uint32_t)(base + e * 64) : UINT32_MAX);
                }
            }
        }
        cent = dd.size();
        if (cent) {
            sink.alloc(1);
            sink32.alloc(1);
            cds.upload(dd, s);
            cmap.upload(dm, s);
            cval.alloc(cent);
            rslot.alloc(n);
        }
        lvl_ptr.upload(h_lvl_ptr, s);
        slot_row.upload(order, s);
        gptr.upload(g, s);
        col.upload(hc, s);
        dslot.upload(hd, s);
        vmap.upload(hm, s);
        len.upload(ln, s);
        val.alloc(std::max<uint64_t>(nent, 1));
        if (diagpos) {
            std::vector<uint32_t> dm(n);
            for (uint32_t s2 = 0; s2 < n; ++s2) dm[s2] = (*diagpos)[order[s2]];
            dmap.upload(dm, s);
            diag.alloc(n);
        }
    }
};

// gather factor values into the sweep layout (padding entries get 0)
__global__ void k_gather_vals(const double* src, const uint32_t* map, double* dst, uint64_t n) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t m = map[i];
        dst[i] = m == UINT32_MAX ? 0.0 : src[m];
    }
}

template <typename T>
__global__ void k_gather_dense(const T* src, const uint32_t* map, T* dst, uint64_t n) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t m = map[i];
        dst[i] = m == UINT32_MAX ? (T)0 : src[m];
    }
}

// rhs in slot order for the slots [s0, s1) of a chain
template <typename T>
__global__ void k_gather_rhs(const uint32_t* __restrict__ slot_row, uint32_t s0, uint32_t s1, const T* __restrict__ rhs,
                             T* __restrict__ out, Gate g) {
    if (g.closed()) return;
    const uint32_t s = s0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (s < s1) out[s] = rhs[slot_row[s]];
}

// y[slot_row[s]] = ys[s] for the slots of a chain (its results in slot order)
template <typename T>
__global__ void k_scatter_slots(const uint32_t* __restrict__ slot_row, uint32_t s0, uint32_t s1,
                                const T* __restrict__ ys, T* __restrict__ y, Gate g) {
    if (g.closed()) return;
    const uint32_t s = s0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (s < s1) y[slot_row[s]] = ys[s];
}

// copy the sweep's values (and float32 copy when present) into the chain layout
static void densify(Sweep& S, hipStream_t s) {
    if (!S.cent) return;
    const unsigned G = (unsigned)((S.cent + 255) / 256);
    hipLaunchKernelGGL((k_gather_dense<double>), dim3(G), dim3(256), 0, s, S.val.get(), S.cmap.get(), S.cval.get(), S.cent);
    if (S.val32.get()) {
        S.cval32.alloc(S.cent);
        S.rslot32.alloc(S.n);
        hipLaunchKernelGGL((k_gather_dense<float>), dim3(G), dim3(256), 0, s, S.val32.get(), S.cmap.get(),
                           S.cval32.get(), S.cent);
    }
    FEM_HIP(hipGetLastError());
}

static void gather(const double* src, const uint32_t* map, double* dst, uint64_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_gather_vals, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, map, dst, n);
    FEM_HIP(hipGetLastError());
}

// One level, many blocks: y[i] = (rhs[i] - sum_e val*y[col]) (/ diag)
template <typename T, bool DIV>
__global__ __launch_bounds__(256) void k_sweep_level(uint32_t s0, uint32_t s1, const uint32_t* __restrict__ slot_row,
                                                     const uint16_t* __restrict__ len, const uint32_t* __restrict__ gptr,
                                                     const uint32_t* __restrict__ col, const T* __restrict__ val,
                                                     const T* __restrict__ diag, const T* __restrict__ rhs, T* y,
                                                     Gate g) {
    if (g.closed()) return;
    const uint32_t s = s0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= s1) return;
    const uint32_t i = slot_row[s];
    const uint32_t base = gptr[s >> 6] + (s & 63);
    const uint32_t L = len[s];
    T acc = rhs[i];
    if (L <= 8) {
        // every entry's column and value, then every gather, are issued
        // before any use (three dependent rounds instead of two per entry);
        // entries past L repeat entry 0 and are not subtracted, so the
        // arithmetic is the loop's below
        uint32_t cl[8];
        T vl[8], yl[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t k = base + 64u * ((uint32_t)e < L ? (uint32_t)e : 0u);
            cl[e] = L ? col[k] : 0u;
            vl[e] = L ? val[k] : (T)0;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) yl[e] = L ? y[cl[e]] : (T)0;
#pragma unroll
        for (int e = 0; e < 8; ++e)
            if ((uint32_t)e < L) acc -= vl[e] * yl[e];
    } else {
        for (uint32_t e = 0; e < L; ++e) acc -= val[base + e * 64] * y[col[base + e * 64]];
    }
    if constexpr (DIV) acc = acc / diag[s];
    y[i] = acc;
}

// A run of consecutive small levels (<= kChainRows rows each) in one
// workgroup, one row per thread per level. Results go to the slot-ordered
// right-hand-side buffer (in place) and to an LDS ring
// indexed by sweep slot (position slot % kRing); the kernel first loads the
// kRing slots before the chain into it, and the host builds chains whose
// rows depend only on slots at most kRing behind their level's end, so every
// dependency is read from the ring. For level [a, b) the ring then holds
// slots b - kRing .. b - 1 at distinct positions: this level's writes (slots
// a..b-1) never clobber an entry still being read. The chain's entries sit
// in a dense per-slot layout and its right-hand side in slot order, so every
// load of a row is addressed by the slot alone; each thread loads the row it
// will handle D levels ahead, unconditionally (a slot past the level reloads
// slot s0), so no wait on those loads precedes their use. Per-row arithmetic
// order is unchanged (rhs first, then the entries in ascending order).
template <typename T, bool DIV, int KC>
__global__ __launch_bounds__(kChainThreads) void k_sweep_chain(uint32_t l0, uint32_t l1, const uint32_t* __restrict__ lvl_ptr,
                                                               const uint32_t* __restrict__ slot_row,
                                                               const uint16_t* __restrict__ len,
                                                               const uint32_t* __restrict__ cds,
                                                               const T* __restrict__ cval, const T* __restrict__ diag,
                                                               T* rslot, const T* y, T* sink, Gate g) {
    if (g.closed()) return;
    __shared__ T ring[kRing + 1];  // + a sink for threads without a row
    __shared__ uint32_t lp[kChainLevels + 1];
    for (uint32_t t = threadIdx.x; t <= l1 - l0; t += blockDim.x) lp[t] = lvl_ptr[l0 + t];
    __syncthreads();
    const uint32_t s0 = lp[0];
    // the kRing slots before the chain (written by earlier launches)
    for (uint32_t t = threadIdx.x; t < kRing; t += blockDim.x) {
        const uint32_t back = kRing - t;  // slot s0 - back
        ring[(s0 - back) % kRing] = back <= s0 ? y[slot_row[s0 - back]] : (T)0;
    }
    __syncthreads();
    struct Row {
        uint32_t on, L;
        T r, d;
        uint32_t ds[KC];
        T v[KC];
    };
    auto fetch = [&](uint32_t l, Row& R) {
        const uint32_t lc = l < l1 ? l : l1 - 1;
        const uint32_t sr = lp[lc - l0] + threadIdx.x;
        R.on = (l < l1 && sr < lp[lc - l0 + 1]) ? 1u : 0u;
        const uint32_t s = R.on ? sr : s0;
        R.L = len[s];
        R.r = rslot[s];
        R.d = DIV ? diag[s] : (T)1;
        const uint64_t o = (uint64_t)(s - s0) * KC;
#pragma unroll
        for (int e = 0; e < KC; ++e) {
            R.ds[e] = cds[o + e];
            R.v[e] = cval[o + e];
        }
    };
    // D rows in flight per thread: buf[k] holds level l0 + q D + k; after a
    // level is done its slot is refilled with the level D further on. The
    // rotation is unrolled so buf stays in registers.
    constexpr int D = KC <= 2 ? 6 : KC <= 4 ? 4 : 2;
    Row buf[D];
#pragma unroll
    for (int k = 0; k < D; ++k) fetch(l0 + k, buf[k]);
    for (uint32_t base = l0; base < l1; base += D) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const uint32_t l = base + k;
            if (l >= l1) break;  // uniform over the workgroup
            Row& cur = buf[k];
            // a wavefront with no row in a level skips its arithmetic (the
            // fetch stays unconditional, which keeps the compiler's waits on
            // the prefetched rows exact): levels are narrower than the
            // workgroup most of the time
            const uint32_t wave0 = threadIdx.x & ~63u;
            if (lp[l - l0] + wave0 < lp[l - l0 + 1]) {
                T acc = cur.r;
#pragma unroll
                for (int e = 0; e < KC; ++e) {
                    const T t = acc - cur.v[e] * ring[cur.ds[e] % kRing];
                    acc = (uint32_t)e < cur.L ? t : acc;
                }
                if constexpr (DIV) acc = acc / cur.d;
                // stores without a branch: a thread without a row writes to
                // the sinks (ring[kRing], and sink[0] in global memory). The
                // result goes to the row's slot of rslot, whose right-hand
                // side was read D levels ago: a level's slots are
                // consecutive, so the store is coalesced; k_scatter_slots
                // moves the chain's results to y afterwards.
                const uint32_t s = lp[l - l0] + threadIdx.x;
                ring[cur.on ? s % kRing : kRing] = acc;
                *(cur.on ? rslot + s : sink) = acc;
            }
            fetch(l + D, cur);
            // Between levels only the LDS ring must be ordered: wait for LDS
            // operations and barrier, leaving the loads fetched ahead in
            // flight (__syncthreads would also wait for them).
            __asm__ volatile("" ::: "memory");
            __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): LDS done
            __builtin_amdgcn_s_barrier();
            __asm__ volatile("" ::: "memory");
        }
    }
}

// A run of consecutive levels of kChainRows..kWideR * 1024 rows in one
// workgroup: kWideR rows per thread per level, an LDS ring of the last
// kRingW results (preloaded like the narrow chain's), each level's rows
// fetched one level ahead (the registers of kWideR rows leave room for one
// level in flight). Stores, sinks and the per-row arithmetic order are the
// narrow chain's.
template <typename T, bool DIV, int KC, int kWideR>
__global__ __launch_bounds__(kChainThreads) void k_sweep_chain_wide(uint32_t l0, uint32_t l1,
                                                                    const uint32_t* __restrict__ lvl_ptr,
                                                                    const uint32_t* __restrict__ slot_row,
                                                                    const uint16_t* __restrict__ len,
                                                                    const uint32_t* __restrict__ cds,
                                                                    const T* __restrict__ cval,
                                                                    const T* __restrict__ diag, T* rslot, const T* y,
                                                                    T* sink, Gate g) {
    if (g.closed()) return;
    __shared__ T ring[kRingW + 1];  // + a sink for threads without a row
    __shared__ uint32_t lp[kChainLevels + 1];
    for (uint32_t t = threadIdx.x; t <= l1 - l0; t += blockDim.x) lp[t] = lvl_ptr[l0 + t];
    __syncthreads();
    const uint32_t s0 = lp[0];
    for (uint32_t t = threadIdx.x; t < kRingW; t += blockDim.x) {
        const uint32_t back = kRingW - t;  // slot s0 - back
        ring[(s0 - back) % kRingW] = back <= s0 ? y[slot_row[s0 - back]] : (T)0;
    }
    __syncthreads();
    struct Row {
        uint32_t on, slot, L;
        T r, d;
        uint32_t ds[KC];
        T v[KC];
    };
    auto fetch = [&](uint32_t l, Row (&R)[kWideR]) {
        const uint32_t lc = l < l1 ? l : l1 - 1;
        const uint32_t a = lp[lc - l0], b = lp[lc - l0 + 1];
#pragma unroll
        for (int q = 0; q < kWideR; ++q) {
            const uint32_t sr = a + threadIdx.x + (uint32_t)q * kChainThreads;
            R[q].on = (l < l1 && sr < b) ? 1u : 0u;
            R[q].slot = sr;
            const uint32_t sl = R[q].on ? sr : s0;
            R[q].L = len[sl];
            R[q].r = rslot[sl];
            R[q].d = DIV ? diag[sl] : (T)1;
            const uint64_t o = (uint64_t)(sl - s0) * KC;
#pragma unroll
            for (int e = 0; e < KC; ++e) {
                R[q].ds[e] = cds[o + e];
                R[q].v[e] = cval[o + e];
            }
        }
    };
    auto level = [&](Row (&R)[kWideR]) {
#pragma unroll
        for (int q = 0; q < kWideR; ++q) {
            T acc = R[q].r;
#pragma unroll
            for (int e = 0; e < KC; ++e) {
                const T t = acc - R[q].v[e] * ring[R[q].ds[e] % kRingW];
                acc = (uint32_t)e < R[q].L ? t : acc;
            }
            if constexpr (DIV) acc = acc / R[q].d;
            ring[R[q].on ? R[q].slot % kRingW : kRingW] = acc;
            *(R[q].on ? rslot + R[q].slot : sink) = acc;
        }
        // LDS-only barrier between levels (the next level's rows stay in flight)
        __asm__ volatile("" ::: "memory");
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): LDS done
        __builtin_amdgcn_s_barrier();
        __asm__ volatile("" ::: "memory");
    };
    Row A[kWideR], B[kWideR];
    fetch(l0, A);
    for (uint32_t l = l0; l < l1; l += 2) {
        fetch(l + 1, B);
        level(A);
        if (l + 1 >= l1) break;  // uniform over the workgroup
        fetch(l + 2, A);
        level(B);
    }
}

template <typename T, bool DIV, int KC>
static void chain_launch(const Sweep& S, const Segment& sg, const T* cval, const T* diag, T* rslot, T* y, hipStream_t s,
                         Gate g) {
    T* sink;
    if constexpr (std::is_same<T, double>::value) sink = S.sink.get();
    else sink = S.sink32.get();
    constexpr int R = wide_r(KC);
    if constexpr (R > 0) {
        if (sg.wide) {
            hipLaunchKernelGGL((k_sweep_chain_wide<T, DIV, KC, R>), dim3(1), dim3(kChainThreads), 0, s, sg.lvl0,
                               sg.lvl1, S.lvl_ptr.get(), S.slot_row.get(), S.len.get(), S.cds.get() + sg.cbase,
                               cval + sg.cbase, diag, rslot, y, sink, g);
            return;
        }
    }
    if (sg.wide) fail(FEM_ERR_RUNTIME, "triangular sweep: wide chain with rows longer than 4 entries");
        hipLaunchKernelGGL((k_sweep_chain<T, DIV, KC>), dim3(1), dim3(kChainThreads), 0, s, sg.lvl0, sg.lvl1,
                           S.lvl_ptr.get(), S.slot_row.get(), S.len.get(), S.cds.get() + sg.cbase, cval + sg.cbase,
                           diag, rslot, y, sink, g);
}

template <typename T, bool DIV>
static void run_sweep(const Sweep& S, const T* val, const T* diag, const T* rhs, T* y, hipStream_t s, Gate g) {
    const T* cval;
    T* rslot;
    if constexpr (std::is_same<T, double>::value) {
        cval = S.cval.get();
        rslot = S.rslot.get();
    } else {
        cval = S.cval32.get();
        rslot = S.rslot32.get();
    }
    for (const Segment& sg : S.segs) {
        // a chain with a dense copy: its stride sg.kc equals the kernel width
        if (sg.lvl1 - sg.lvl0 > 1 && (sg.kc == 2 || sg.kc == 4 || sg.kc == 8)) {
            const uint32_t cnt = sg.s1 - sg.s0;
            hipLaunchKernelGGL((k_gather_rhs<T>), dim3((cnt + 255) / 256), dim3(256), 0, s, S.slot_row.get(), sg.s0,
                               sg.s1, rhs, rslot, g);
            if (sg.kc == 2) chain_launch<T, DIV, 2>(S, sg, cval, diag, rslot, y, s, g);
            else if (sg.kc == 4) chain_launch<T, DIV, 4>(S, sg, cval, diag, rslot, y, s, g);
            else chain_launch<T, DIV, 8>(S, sg, cval, diag, rslot, y, s, g);
            hipLaunchKernelGGL((k_scatter_slots<T>), dim3((cnt + 255) / 256), dim3(256), 0, s, S.slot_row.get(), sg.s0,
                               sg.s1, rslot, y, g);
        } else {
            // one launch per level (also for chains with rows longer than 8)
            for (uint32_t l = sg.lvl0; l < sg.lvl1; ++l) {
                const uint32_t a = S.h_lvl_ptr[l], b = S.h_lvl_ptr[l + 1];
                hipLaunchKernelGGL((k_sweep_level<T, DIV>), dim3((b - a + 255) / 256), dim3(256), 0, s, a, b,
                                   S.slot_row.get(), S.len.get(), S.gptr.get(), S.col.get(), val, diag, rhs, y, g);
            }
        }
    }
    FEM_HIP(hipGetLastError());
}

// levels of a triangular dependency graph given per-row dependency lists;
// forward = dependencies have smaller index (process ascending)
static std::vector<uint32_t> levels_of(uint32_t n, const std::vector<uint32_t>& rp, const std::vector<uint32_t>& cols,
                                       bool forward) {
    std::vector<uint32_t> lev(n, 0);
    if (forward) {
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t l = 0;
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) l = std::max(l, lev[cols[p]] + 1);
            lev[i] = l;
        }
    } else {
        for (int64_t ii = (int64_t)n - 1; ii >= 0; --ii) {
            const uint32_t i = (uint32_t)ii;
            uint32_t l = 0;
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) l = std::max(l, lev[cols[p]] + 1);
            lev[i] = l;
        }
    }
    return lev;
}

template <typename T>
__global__ void k_to_f32(const double* a, float* b, uint64_t n) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
static void to_f32(const DevBuf<double>& a, DevBuf<float>& b, uint64_t n, hipStream_t s) {
    b.alloc(std::max<uint64_t>(n, 1));
    if (!n) return;
    hipLaunchKernelGGL(k_to_f32<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a.get(), b.get(), n);
    FEM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// IC(0)
// ---------------------------------------------------------------------------
// Factor the rows of one level (array.cpp:2582-2654). Lp/Lc/Lv: lower part
// with diagonal, columns ascending; dpos[i]: position of L_ii.
__global__ void k_ic0_level(uint32_t a, uint32_t b, const uint32_t* __restrict__ rows, const uint32_t* __restrict__ Lp,
                            const uint32_t* __restrict__ Lc, double* Lv, double* diag, const uint32_t* __restrict__ dpos,
                            int* bad) {
    const uint32_t t = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b) return;
    const uint32_t i = rows[t];
    const uint32_t rs = Lp[i], dp = dpos[i];
    for (uint32_t p = rs; p < dp; ++p) {
        const uint32_t j = Lc[p];
        double s = 0.0;
        uint32_t pi = rs, pj = Lp[j];
        const uint32_t ej = Lp[j + 1];
        while (pi < p && pj < ej) {
            const uint32_t ki = Lc[pi], kj = Lc[pj];
            if (ki == kj) { s += Lv[pi] * Lv[pj]; ++pi; ++pj; }
            else if (ki < kj) ++pi;
            else ++pj;
        }
        Lv[p] = (Lv[p] - s) / diag[j];
    }
    double sd = 0.0;
    for (uint32_t p = rs; p < dp; ++p) sd += Lv[p] * Lv[p];
    const double dv = Lv[dp] - sd;
    if (dv <= 1e-14) { atomicMin(bad, (int)i); return; }
    const double r = sqrt(dv);
    Lv[dp] = r;
    diag[i] = r;
}

// Algorithmic bytes of one sweep: its off-diagonal entries (uint32 indices),
// the row pointers, the right-hand side read and the result written, the
// level list, and the diagonal when the sweep divides.
static double sweep_bytes(const Sweep& w, int s, bool div) {
    return (double)w.ndep * (s + 4) + 4.0 * (w.n + 1) + 2.0 * w.n * s + 4.0 * w.n + (div ? (double)w.n * s : 0.0);
}

struct PcIC0 : Precond {
    uint32_t n;
    Sweep fw, bw;       // forward L, backward L^T
    DevBuf<double> y;   // intermediate vector
    DevBuf<float> y32;

    PcIC0(CsrMat& A, hipStream_t s) : n(A.n) {
        if (A.n != A.m) fail(FEM_ERR_INVALID_ARGUMENT, "IC(0) preconditioner needs a square matrix");
        A.sync_host();
        const auto& rp = A.h_rp;
        const auto& ci = A.h_ci;
        // lower pattern with diagonal (array.cpp:2539-2574)
        std::vector<uint32_t> Lp(n + 1, 0), Lc, amap, dpos(n, UINT32_MAX);
        Lc.reserve(A.nnz / 2 + n);
        amap.reserve(A.nnz / 2 + n);
        for (uint32_t i = 0; i < n; ++i) {
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
                if (ci[p] <= i) {
                    if (ci[p] == i && dpos[i] == UINT32_MAX) dpos[i] = (uint32_t)Lc.size();
                    Lc.push_back(ci[p]);
                    amap.push_back(p);
                }
            Lp[i + 1] = (uint32_t)Lc.size();
        }
        // The reference stops a row's off-diagonal loop at the first column
        // >= i (array.cpp:2589); with sorted columns that is the diagonal.
        for (uint32_t i = 0; i < n; ++i)
            if (dpos[i] == UINT32_MAX) fail(FEM_ERR_RUNTIME, "IC(0) failed: the matrix lacks a diagonal entry");
        const uint64_t nL = Lc.size();
        // strictly-lower dependency lists (forward) and L^T lists (backward)
        std::vector<uint32_t> Fp(n + 1, 0), Fc, Fpos;
        Fc.reserve(nL);
        Fpos.reserve(nL);
        for (uint32_t i = 0; i < n; ++i) {
            for (uint32_t p = Lp[i]; p < dpos[i]; ++p) { Fc.push_back(Lc[p]); Fpos.push_back(p); }
            Fp[i + 1] = (uint32_t)Fc.size();
        }
        // L^T holds every strictly-lower entry of L (array.cpp:2659-2695)
        std::vector<uint32_t> Bp(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t p = Lp[i]; p < Lp[i + 1]; ++p)
                if (Lc[p] < i) Bp[Lc[p] + 1]++;
        for (uint32_t i = 0; i < n; ++i) Bp[i + 1] += Bp[i];
        std::vector<uint32_t> Bc(Bp[n]), Bpos(Bp[n]);
        {
            std::vector<uint32_t> off(Bp.begin(), Bp.end() - 1);
            for (uint32_t i = 0; i < n; ++i)
                for (uint32_t p = Lp[i]; p < Lp[i + 1]; ++p)
                    if (Lc[p] < i) {
                        const uint32_t q = off[Lc[p]]++;
                        Bc[q] = i;
                        Bpos[q] = p;
                    }
        }
        const auto lf = levels_of(n, Fp, Fc, true);
        const auto lb = levels_of(n, Bp, Bc, false);
        fw.build(n, Fp, Fc, Fpos, lf, &dpos, s);
        bw.build(n, Bp, Bc, Bpos, lb, &dpos, s);
        levels = (int)fw.nlev;
        // numeric factorization on the device, level by level
        DevBuf<uint32_t> dLp, dLc, dmap, ddpos;
        DevBuf<double> Lv(std::max<uint64_t>(nL, 1)), dg(n);
        DevBuf<int> bad(1);
        dLp.upload(Lp, s);
        dLc.upload(Lc, s);
        dmap.upload(amap, s);
        ddpos.upload(dpos, s);
        gather(A.v.get(), dmap.get(), Lv.get(), nL, s);
        const int big = 0x7fffffff;
        FEM_HIP(hipMemcpyAsync(bad.get(), &big, sizeof(int), hipMemcpyHostToDevice, s));
        for (uint32_t l = 0; l < fw.nlev; ++l) {
            const uint32_t a = fw.h_lvl_ptr[l], b = fw.h_lvl_ptr[l + 1];
            hipLaunchKernelGGL(k_ic0_level, dim3((b - a + 255) / 256), dim3(256), 0, s, a, b, fw.slot_row.get(),
                               dLp.get(), dLc.get(), Lv.get(), dg.get(), ddpos.get(), bad.get());
        }
        FEM_HIP(hipGetLastError());
        int hb = 0;
        FEM_HIP(hipMemcpyAsync(&hb, bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        if (hb != big) fail(FEM_ERR_RUNTIME, "IC(0) failed: the matrix may not be SPD, or needs a shifted IC");
        gather(Lv.get(), fw.vmap.get(), fw.val.get(), fw.nent, s);
        gather(Lv.get(), bw.vmap.get(), bw.val.get(), bw.nent, s);
        gather(Lv.get(), fw.dmap.get(), fw.diag.get(), n, s);
        gather(Lv.get(), bw.dmap.get(), bw.diag.get(), n, s);
        densify(fw, s);
        densify(bw, s);
        y.alloc(n);
        FEM_HIP(hipStreamSynchronize(s));
    }
    int kind() const override { return FEM_PC_IC0; }
    uint32_t size() const override { return n; }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override {
        run_sweep<double, true>(fw, fw.val.get(), fw.diag.get(), r, y.get(), s, g);
        run_sweep<double, true>(bw, bw.val.get(), bw.diag.get(), y.get(), z, s, g);
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override {
        run_sweep<float, true>(fw, fw.val32.get(), fw.diag32.get(), r, y32.get(), s, g);
        run_sweep<float, true>(bw, bw.val32.get(), bw.diag32.get(), y32.get(), z, s, g);
    }
    double apply_bytes(int s) override { return sweep_bytes(fw, s, true) + sweep_bytes(bw, s, true); }
    bool has_fp32() const override { return y32.get() != nullptr; }
    void ensure_fp32(hipStream_t s) override {
        if (y32.get()) return;
        to_f32(fw.val, fw.val32, fw.nent, s);
        to_f32(bw.val, bw.val32, bw.nent, s);
        to_f32(fw.diag, fw.diag32, n, s);
        to_f32(bw.diag, bw.diag32, n, s);
        densify(fw, s);
        densify(bw, s);
        y32.alloc(n);
    }
};

// ---------------------------------------------------------------------------
// ILU(0), the reference's "upper-only" IKJ variant
// ---------------------------------------------------------------------------
// Factor the rows of one level (array.cpp:2846-2897). L: strict lower,
// U: upper with diagonal (dU[i] = position of U_ii).
__global__ void k_ilu0_level(uint32_t a, uint32_t b, const uint32_t* __restrict__ rows, const uint32_t* __restrict__ Lp,
                             const uint32_t* __restrict__ Lc, double* Lv, const uint32_t* __restrict__ Up,
                             const uint32_t* __restrict__ Uc, double* Uv, const uint32_t* __restrict__ dU, int* bad) {
    const uint32_t t = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b) return;
    const uint32_t i = rows[t];
    const uint32_t us = Up[i], ue = Up[i + 1];
    for (uint32_t pl = Lp[i]; pl < Lp[i + 1]; ++pl) {
        const uint32_t k = Lc[pl];
        const double ukk = Uv[dU[k]];
        if (fabs(ukk) < 1e-30) { atomicMin(bad, (int)i); return; }
        const double lik = Lv[pl] / ukk;
        Lv[pl] = lik;
        uint32_t pi = us, pk = Up[k];
        const uint32_t ek = Up[k + 1];
        while (pi < ue && pk < ek) {
            const uint32_t ji = Uc[pi], jk = Uc[pk];
            if (ji == jk) { Uv[pi] -= lik * Uv[pk]; ++pi; ++pk; }
            else if (ji < jk) ++pi;
            else ++pk;
        }
    }
}

struct PcILU0 : Precond {
    uint32_t n;
    Sweep fw, bw;  // forward unit-lower L, backward U
    DevBuf<double> y;
    DevBuf<float> y32;

    PcILU0(CsrMat& A, hipStream_t s) : n(A.n) {
        if (A.n != A.m) fail(FEM_ERR_INVALID_ARGUMENT, "ILU(0) preconditioner needs a square matrix");
        A.sync_host();
        const auto& rp = A.h_rp;
        const auto& ci = A.h_ci;
        std::vector<uint32_t> Lp(n + 1, 0), Up(n + 1, 0), Lc, Uc, Lmap, Umap, dU(n, UINT32_MAX);
        for (uint32_t i = 0; i < n; ++i) {
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
                const uint32_t j = ci[p];
                if (j < i) { Lc.push_back(j); Lmap.push_back(p); }
                else {
                    if (j == i) dU[i] = (uint32_t)Uc.size();
                    Uc.push_back(j);
                    Umap.push_back(p);
                }
            }
            Lp[i + 1] = (uint32_t)Lc.size();
            Up[i + 1] = (uint32_t)Uc.size();
        }
        for (uint32_t i = 0; i < n; ++i)
            if (dU[i] == UINT32_MAX) fail(FEM_ERR_RUNTIME, "ILU(0) failed: the matrix lacks a diagonal entry");
        // backward dependency lists: U row without the diagonal
        std::vector<uint32_t> Bp(n + 1, 0), Bc, Bpos;
        for (uint32_t i = 0; i < n; ++i) {
            for (uint32_t p = Up[i]; p < Up[i + 1]; ++p)
                if (Uc[p] != i) { Bc.push_back(Uc[p]); Bpos.push_back(p); }
            Bp[i + 1] = (uint32_t)Bc.size();
        }
        std::vector<uint32_t> Lpos(Lc.size());
        std::iota(Lpos.begin(), Lpos.end(), 0u);
        const auto lf = levels_of(n, Lp, Lc, true);
        const auto lb = levels_of(n, Bp, Bc, false);
        fw.build(n, Lp, Lc, Lpos, lf, nullptr, s);
        bw.build(n, Bp, Bc, Bpos, lb, &dU, s);
        levels = (int)fw.nlev;
        DevBuf<uint32_t> dLp, dLc, dUp, dUc, ddU, dLm, dUm;
        DevBuf<double> Lv(std::max<size_t>(Lc.size(), 1)), Uv(std::max<size_t>(Uc.size(), 1));
        DevBuf<int> bad(1);
        dLp.upload(Lp, s); dLc.upload(Lc, s); dUp.upload(Up, s); dUc.upload(Uc, s); ddU.upload(dU, s);
        dLm.upload(Lmap, s); dUm.upload(Umap, s);
        gather(A.v.get(), dLm.get(), Lv.get(), Lc.size(), s);
        gather(A.v.get(), dUm.get(), Uv.get(), Uc.size(), s);
        const int big = 0x7fffffff;
        FEM_HIP(hipMemcpyAsync(bad.get(), &big, sizeof(int), hipMemcpyHostToDevice, s));
        for (uint32_t l = 0; l < fw.nlev; ++l) {
            const uint32_t a = fw.h_lvl_ptr[l], b = fw.h_lvl_ptr[l + 1];
            hipLaunchKernelGGL(k_ilu0_level, dim3((b - a + 255) / 256), dim3(256), 0, s, a, b, fw.slot_row.get(),
                               dLp.get(), dLc.get(), Lv.get(), dUp.get(), dUc.get(), Uv.get(), ddU.get(), bad.get());
        }
        FEM_HIP(hipGetLastError());
        int hb = 0;
        FEM_HIP(hipMemcpyAsync(&hb, bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        if (hb != big) fail(FEM_ERR_RUNTIME, "ILU(0) breakdown: U_kk close to 0, pivoting may be needed");
        gather(Lv.get(), fw.vmap.get(), fw.val.get(), fw.nent, s);
        gather(Uv.get(), bw.vmap.get(), bw.val.get(), bw.nent, s);
        gather(Uv.get(), bw.dmap.get(), bw.diag.get(), n, s);
        densify(fw, s);
        densify(bw, s);
        y.alloc(n);
        FEM_HIP(hipStreamSynchronize(s));
    }
    int kind() const override { return FEM_PC_ILU0; }
    uint32_t size() const override { return n; }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override {
        run_sweep<double, false>(fw, fw.val.get(), nullptr, r, y.get(), s, g);
        run_sweep<double, true>(bw, bw.val.get(), bw.diag.get(), y.get(), z, s, g);
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override {
        run_sweep<float, false>(fw, fw.val32.get(), nullptr, r, y32.get(), s, g);
        run_sweep<float, true>(bw, bw.val32.get(), bw.diag32.get(), y32.get(), z, s, g);
    }
    double apply_bytes(int s) override { return sweep_bytes(fw, s, false) + sweep_bytes(bw, s, true); }
    bool has_fp32() const override { return y32.get() != nullptr; }
    void ensure_fp32(hipStream_t s) override {
        if (y32.get()) return;
        to_f32(fw.val, fw.val32, fw.nent, s);
        to_f32(bw.val, bw.val32, bw.nent, s);
        to_f32(bw.diag, bw.diag32, n, s);
        densify(fw, s);
        densify(bw, s);
        y32.alloc(n);
    }
};

std::unique_ptr<Precond> make_ic0(CsrMat& A, hipStream_t s) { return std::unique_ptr<Precond>(new PcIC0(A, s)); }
std::unique_ptr<Precond> make_ilu0(CsrMat& A, hipStream_t s) { return std::unique_ptr<Precond>(new PcILU0(A, s)); }

}  // namespace fem

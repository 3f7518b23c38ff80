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
// Apply: a sweep whose rows have at most 16 entries (every stencil matrix
// here) is one launch of the block sweep (k_sweep_blk): the rows, in
// processing order, are cut into blocks of kBlkRows; a workgroup takes a
// block, orders its rows by level, keeps the block's results in LDS and
// reads rows of earlier blocks through a sentinel-initialized array, so no
// launch, grid barrier or L2 round trip separates the levels inside a
// block. Sweeps with longer rows use the level schedule: rows in level order
// as SELL-64 (groups of 64 level-ordered rows, entries column-major within a
// group), runs of small levels in one workgroup with an LDS ring (chains),
// runs of wide levels in one sync-free launch. Every path keeps the
// reference's per-row summation order (ascending column for the forward
// sweep and for U, ascending row for the L^T sweep of IC0).
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
// block sweeps: rows per block (results of a block stay in LDS) and the
// widest row (entries) the block kernel holds in registers
constexpr uint32_t kBlkRows = 8176;
constexpr uint32_t kBlkMaxEntries = 16;
constexpr uint32_t kExtDep = 0x80000000u;  // dependency on a row of an earlier block
constexpr uint16_t kExtUse = 0x8000;       // entry count flag: a later block reads this row
constexpr uint32_t kXLead = 1;             // block sweep: steps between the first external reading and use
constexpr int kBlkWaves = 2;               // block sweep: wavefronts per block, taking steps in turn
constexpr uint32_t wide_rows(uint32_t kc) { return (uint32_t)wide_r(kc) * 1024u; }

using Idx = HostVec<uint32_t>;  // host index arrays of the setup (filled in full before use)

// ---------------------------------------------------------------------------
// Level schedule + SELL-64 sweep layout
// ---------------------------------------------------------------------------
// A row whose input has not arrived after this long (seconds, on the
// device's constant wall clock) gives up and stores NaN: the solve then
// reports a breakdown instead of the kernel never ending. Rows only wait on
// rows of lower tickets, so this is a safety net.
constexpr double kSfWaitSeconds = 2.0;

// (queried at set-up, outside any stream capture)
static int sf_cus() {
    static int c = [] {
        int dev = 0, cu = 0;
        FEM_HIP(hipGetDevice(&dev));
        FEM_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
        return cu;
    }();
    return c;
}

static long long sf_wait_ticks() {
    static long long ticks = [] {
        int dev = 0, khz = 0;
        FEM_HIP(hipGetDevice(&dev));
        FEM_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        return (long long)(kSfWaitSeconds * 1e3 * (khz > 0 ? khz : 100000));
    }();
    return ticks;
}

struct Segment {
    uint32_t s0, s1;       // slot range
    uint32_t lvl0, lvl1;   // level range (chain when lvl1 - lvl0 > 1)
    uint32_t kc = 0;       // chain: entries per slot in the dense layout
    uint64_t cbase = 0;    // chain: offset of slot s0 in the dense arrays
    bool wide = false;     // chain of levels wider than kChainRows (wide-chain kernel)
    bool sf = false;       // run of levels swept by one sync-free launch
    uint32_t tick = 0;     // sync-free: index of its ticket pair in Sweep::tick
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
    // sync-free segments: results in slot order (a sentinel until written)
    // and per segment a (ticket, finished) counter pair
    DevBuf<double> ys;
    DevBuf<float> ys32;
    DevBuf<uint32_t> tick;
    uint32_t nsf = 0;
    // block sweep (used when every row has at most kBlkMaxEntries entries):
    // the rows, in processing order, cut into blocks of kBlkRows; block b
    // holds global slots [b kBlkRows, ...), ordered by level within the
    // block. Per global slot: its row, entry count, the diagonal's position
    // in the factor and bkc entries (dependency: local slot inside the block,
    // or kExtDep | global slot of a row of an earlier block; values and their
    // positions in the factor). bstp: per block (from bsptr[b]) the first
    // local slot of each step (a level, or 64 rows of a wider level).
    bool blk = false;
    uint32_t bkc = 0, nblk = 0;
    DevBuf<uint32_t> brow, bdep, bmap, bdmap, bsptr;
    DevBuf<uint16_t> blen, bstp;
    DevBuf<double> bval, bdiag;
    DevBuf<float> bval32, bdiag32;

    // rows: dependency lists per row (entries in summation order), as
    // (col, position-in-factor-values). lev: level per row.
    void build(uint32_t n_, const Idx& rp, const Idx& cols, const Idx& pos, const Idx& lev, const Idx* diagpos,
               bool forward, hipStream_t s) {
        SetupTrace tr(forward ? "sweep layout (forward)" : "sweep layout (backward)", s);
        n = n_;
        nlev = n ? 1 + *std::max_element(lev.begin(), lev.end()) : 0;
        h_lvl_ptr.assign(nlev + 1, 0);
        for (uint32_t i = 0; i < n; ++i) h_lvl_ptr[lev[i] + 1]++;
        for (uint32_t l = 0; l < nlev; ++l) h_lvl_ptr[l + 1] += h_lvl_ptr[l];
        Idx order(n);
        std::vector<uint32_t> off(h_lvl_ptr.begin(), h_lvl_ptr.end() - 1);
        for (uint32_t i = 0; i < n; ++i) order[off[lev[i]]++] = i;  // stable: ascending row within a level
        tr.lap("level order");
        ndep = n ? (uint64_t)rp[n] - rp[0] : 0;
        uint32_t kmax = 0;
        for (uint32_t i = 0; i < n; ++i) kmax = std::max(kmax, rp[i + 1] - rp[i]);
        if (kmax > 65535) fail(FEM_ERR_RUNTIME, "triangular factor row too long for the sweep layout");
        segs.clear();
        blk = false;
        if (n && kmax <= kBlkMaxEntries) {
            // the block sweep covers this sweep: its own layout, values
            // gathered straight from the factor (no level schedule)
            build_blocks(rp, cols, pos, lev, diagpos, forward, kmax, tr, s);
            slot_row.upload(order, s);  // level order (the numeric factorization's launches)
            tr.lap("blocks, uploads");
            return;
        }
        const uint32_t ngrp = (n + 63) / 64;
        std::vector<uint32_t> g(ngrp + 1, 0);
        std::vector<uint16_t> ln(n);
        for (uint32_t s = 0; s < n; ++s) ln[s] = (uint16_t)(rp[order[s] + 1] - rp[order[s]]);
        for (uint32_t gi = 0; gi < ngrp; ++gi) {
            uint32_t w = 0;
            for (uint32_t s = gi * 64; s < std::min(n, gi * 64 + 64); ++s) w = std::max<uint32_t>(w, ln[s]);
            g[gi + 1] = g[gi] + w * 64;
        }
        nent = g[ngrp];
        Idx slot_of(n);
        for (uint32_t s = 0; s < n; ++s) slot_of[order[s]] = s;
        Idx hc(nent, 0), hm(nent, UINT32_MAX), hd(nent, 0);
        for (uint32_t s = 0; s < n; ++s) {
            const uint32_t i = order[s];
            const uint32_t base = g[s >> 6] + (s & 63);
            for (uint32_t e = 0; e < ln[s]; ++e) {
                hc[base + e * 64] = cols[rp[i] + e];
                hd[base + e * 64] = slot_of[cols[rp[i] + e]];
                hm[base + e * 64] = pos[rp[i] + e];
            }
        }
        tr.lap("level order, SELL layout");
        if (n) build_segments(rp, cols, order, slot_of, g, ln, hd, std::max(kmax, 1u), s);
        tr.lap("segments");
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
        tr.lap("uploads");
    }

    // Level-schedule segments (sweeps with rows longer than the block sweep
    // holds): chains of consecutive small levels, big levels alone.
    void build_segments(const Idx& rp, const Idx& cols, const Idx& order, const Idx& slot_of,
                        const std::vector<uint32_t>& g, const std::vector<uint16_t>& ln, const Idx& hd, uint32_t kcmax,
                        hipStream_t s) {
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
        // every run of segments without a dense chain copy (levels too wide
        // for one workgroup, and chains of rows longer than 8 entries)
        // becomes one sync-free segment
        {
            std::vector<Segment> merged;
            nsf = 0;
            for (const Segment& sg : segs) {
                if (sg.lvl1 - sg.lvl0 > 1 && sg.kc != 0) {
                    merged.push_back(sg);
                } else if (!merged.empty() && merged.back().sf) {
                    merged.back().s1 = sg.s1;
                    merged.back().lvl1 = sg.lvl1;
                } else {
                    Segment m{sg.s0, sg.s1, sg.lvl0, sg.lvl1};
                    m.sf = true;
                    m.tick = nsf++;
                    merged.push_back(m);
                }
            }
            segs.swap(merged);
            if (nsf) {
                tick.alloc(2 * nsf);
                FEM_HIP(hipMemsetAsync(tick.get(), 0, 2 * nsf * sizeof(uint32_t), s));
                ys.alloc(n);
                sf_wait_ticks();
                sf_cus();
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
    }

    // Block layout (see the fields). Within a block, rows are ordered by
    // their level in the whole sweep, so a block runs about one step behind
    // the blocks it reads from wherever the block boundaries fall; a row's
    // entries keep their summation order. Values come straight from the
    // factor: bmap holds each entry's position in it (pos), bdmap the
    // diagonal's (diagpos). Blocks are laid out in parallel on the host.
    void build_blocks(const Idx& rp, const Idx& cols, const Idx& pos, const Idx& lev, const Idx* diagpos,
                      bool forward, uint32_t kmax, SetupTrace& tr, hipStream_t s) {
        bkc = kmax <= 4 ? 4u : kmax <= 8 ? 8u : 16u;
        nblk = (n + kBlkRows - 1) / kBlkRows;
        auto row_at = [&](uint32_t q) { return forward ? q : n - 1 - q; };
        auto pos_of = [&](uint32_t i) { return forward ? i : n - 1 - i; };
        // every slot and entry of these is written below (gslot in step 1,
        // the others in step 2)
        Idx hrow(n), hdep((size_t)n * bkc), hmap((size_t)n * bkc), hdm(diagpos ? n : 0), gslot(n);
        std::vector<uint32_t> sp(nblk + 1, 0);
        HostVec<uint16_t> hlen(n);
        std::vector<uint16_t> stp;
        std::vector<uint8_t> ext(n, 0);
        std::vector<std::vector<uint16_t>> steps(nblk);
        std::atomic<int> later{0};
        // 1. per block: rows by level (counting sort, stable in processing
        //    order) and the first local slot of each step (a level's rows in
        //    this block, 64 at a time)
        host_parallel(nblk, [&](uint32_t b) {
            const uint32_t q0 = b * kBlkRows, q1 = std::min(n, q0 + kBlkRows);
            uint32_t lo = UINT32_MAX, hi = 0;
            for (uint32_t q = q0; q < q1; ++q) {
                const uint32_t i = row_at(q);
                for (uint32_t t = rp[i]; t < rp[i + 1]; ++t)
                    if (pos_of(cols[t]) >= q) later.store(1, std::memory_order_relaxed);
                lo = std::min(lo, lev[i]);
                hi = std::max(hi, lev[i]);
            }
            std::vector<uint32_t> cnt(hi - lo + 2, 0);
            for (uint32_t q = q0; q < q1; ++q) cnt[lev[row_at(q)] - lo + 1]++;
            for (uint32_t l = 0; l + 1 < cnt.size(); ++l) cnt[l + 1] += cnt[l];
            for (uint32_t l = 0; l + 1 < cnt.size(); ++l)
                for (uint32_t c = cnt[l]; c < cnt[l + 1]; c += 64) steps[b].push_back((uint16_t)c);
            for (uint32_t q = q0; q < q1; ++q) gslot[row_at(q)] = q0 + cnt[lev[row_at(q)] - lo]++;
        });
        if (later.load()) fail(FEM_ERR_RUNTIME, "triangular sweep: a row depends on a later row");
        tr.lap("block order");
        for (uint32_t b = 0; b < nblk; ++b) {
            sp[b + 1] = sp[b] + (uint32_t)steps[b].size();
            stp.insert(stp.end(), steps[b].begin(), steps[b].end());
        }
        // 2. per block: rows, entry counts, dependencies (local slot, or
        //    kExtDep | global slot of an earlier block's row), factor positions
        host_parallel(nblk, [&](uint32_t b) {
            const uint32_t q0 = b * kBlkRows, q1 = std::min(n, q0 + kBlkRows);
            for (uint32_t q = q0; q < q1; ++q) {
                const uint32_t i = row_at(q), gs = gslot[i];
                const uint32_t L = rp[i + 1] - rp[i];
                hrow[gs] = i;
                hlen[gs] = (uint16_t)L;
                if (diagpos) hdm[gs] = (*diagpos)[i];
                for (uint32_t e = 0; e < L; ++e) {
                    const uint32_t gj = gslot[cols[rp[i] + e]];
                    hdep[(size_t)gs * bkc + e] = gj >= q0 ? gj - q0 : (kExtDep | gj);
                    // a row of an earlier block: flag it (every writer stores 1)
                    if (gj < q0) __atomic_store_n(&ext[gj], (uint8_t)1, __ATOMIC_RELAXED);
                    hmap[(size_t)gs * bkc + e] = pos[rp[i] + e];
                }
                for (uint32_t e = L; e < bkc; ++e) {  // padding: no dependency, value 0
                    hdep[(size_t)gs * bkc + e] = 0;
                    hmap[(size_t)gs * bkc + e] = UINT32_MAX;
                }
            }
        });
        for (uint32_t q = 0; q < n; ++q)
            if (ext[q]) hlen[q] |= kExtUse;  // a later block reads this row
        tr.lap("block entries");
        brow.upload(hrow, s);
        blen.upload(hlen, s);
        bdep.upload(hdep, s);
        bmap.upload(hmap, s);
        bsptr.upload(sp, s);
        bstp.upload(stp, s);
        bval.alloc((size_t)n * bkc);
        if (diagpos) {
            bdmap.upload(hdm, s);
            bdiag.alloc(n);
        }
        tick.alloc(2);  // the block sweep's counters
        FEM_HIP(hipMemsetAsync(tick.get(), 0, 2 * sizeof(uint32_t), s));
        ys.alloc(n + 2);  // slot n: target of the stores of lanes without a row; n + 1: read dummy
        sink.alloc(1);
        sink32.alloc(1);
        sf_wait_ticks();
        sf_cus();
        blk = true;
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

// copy the sweep's values (and float32 copy when present) into the chain
// layout (level-schedule sweeps)
static void densify(Sweep& S, hipStream_t s) {
    if (S.nsf && S.val32.get() && !S.ys32.get()) S.ys32.alloc(S.n + 2);
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

// ---------------------------------------------------------------------------
// Sync-free sweep of a run of wide levels in one launch (level schedule)
// ---------------------------------------------------------------------------
// Every slot's result is also stored, in slot order, into ys, which holds a
// signalling-NaN sentinel until then: a row waits for exactly the rows it
// reads (their ys entries leave the sentinel) instead of for whole levels,
// so rows of many levels are in flight at once and no launch or grid
// barrier separates levels. Arithmetic never produces a signalling NaN, and
// a result that equals the sentinel bit for bit (a right-hand side copied
// through) is stored as the quiet NaN instead, so a finished row is never
// mistaken for a pending one.
//
// Workgroups take slot blocks in ticket order (an atomic counter), so every
// block a workgroup waits on is running or done: the lowest unfinished slot
// always has its inputs and the sweep makes progress. Within a wavefront a
// row may read a row of a lower lane; each lane therefore polls, computes
// and stores inside one loop, and a lane that is ready finishes while the
// others keep polling. ys is written and polled with relaxed device-scope
// atomics (the value is its own flag, so no fence is needed); rows of
// earlier launches are read from y. Per-row arithmetic is the level
// kernel's: rhs first, then the entries in ascending order.
template <typename T>
struct SfBits;
template <>
struct SfBits<double> {
    using U = unsigned long long;
    static constexpr U kSent = 0x7FF4DEADBEEF0001ull;  // signalling NaN
    static constexpr U kQuiet = 0x7FF8000000000000ull;
};
template <>
struct SfBits<float> {
    using U = unsigned int;
    static constexpr U kSent = 0x7F80DEADu;  // signalling NaN
    static constexpr U kQuiet = 0x7FC00000u;
};

constexpr int kSfThreads = 256;

template <typename T>
__global__ void k_sf_fill(T* ys, uint32_t s0, uint32_t s1, Gate g) {
    if (g.closed()) return;
    using U = typename SfBits<T>::U;
    const uint32_t s = s0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (s < s1) reinterpret_cast<U*>(ys)[s] = SfBits<T>::kSent;
}

template <typename T, bool DIV>
__global__ __launch_bounds__(kSfThreads) void k_sweep_sf(uint32_t s0, uint32_t s1, const uint32_t* __restrict__ slot_row,
                                                         const uint16_t* __restrict__ len,
                                                         const uint32_t* __restrict__ gptr,
                                                         const uint32_t* __restrict__ col,
                                                         const uint32_t* __restrict__ dslot,
                                                         const T* __restrict__ val, const T* __restrict__ diag,
                                                         const T* __restrict__ rhs, T* y, T* ys, uint32_t* tick,
                                                         long long wait_ticks, Gate g) {
    using B = SfBits<T>;
    using U = typename B::U;
    if (g.closed()) return;
    __shared__ uint32_t blk_sh;
    const uint32_t nblk = (s1 - s0 + kSfThreads - 1) / kSfThreads;
    U* yb = reinterpret_cast<U*>(ys);
    bool timed_out = false;
    for (;;) {
        __syncthreads();  // blk_sh is rewritten below
        if (threadIdx.x == 0) blk_sh = atomicAdd(tick, 1u);
        __syncthreads();
        const uint32_t blk = blk_sh;
        if (blk >= nblk) break;  // uniform over the workgroup
        const long long t0 = wall_clock64();  // the waits of this slot block
        const uint32_t first = s0 + blk * kSfThreads;
        const uint32_t s = first + threadIdx.x;
        const bool on = s < s1;
        uint32_t i = 0, base = 0, L = 0;
        T acc = (T)0, d = (T)1;
        if (on) {
            i = slot_row[s];
            base = gptr[s >> 6] + (s & 63);
            L = len[s];
            acc = rhs[i];
            if constexpr (DIV) d = diag[s];
        }
        // rows of at most 8 entries keep the whole row in registers
        const bool reg = L <= 8;
        uint32_t ds[8];
        T vl[8], yl[8];
        uint32_t pend = 0;
        if (on && reg) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t k = base + 64u * ((uint32_t)e < L ? (uint32_t)e : 0u);
                ds[e] = L ? dslot[k] : 0u;
                vl[e] = L ? val[k] : (T)0;
                yl[e] = (T)0;
                if ((uint32_t)e < L) {
                    if (ds[e] < s0) yl[e] = y[col[k]];
                    else pend |= 1u << e;
                }
            }
        }
        bool done = !on;
        uint32_t e = 0;
        while (!done) {
            bool ready;
            if (reg) {
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if ((pend >> q) & 1u) {
                        const U b = __hip_atomic_load(yb + ds[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (b != B::kSent) {
                            yl[q] = __builtin_bit_cast(T, b);
                            pend &= ~(1u << q);
                        }
                    }
                ready = pend == 0;
                if (ready) {
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if ((uint32_t)q < L) acc -= vl[q] * yl[q];
                }
            } else {
                // longer rows: entries in order, each as soon as its input is in
                while (e < L) {
                    const uint32_t k = base + 64u * e;
                    const uint32_t dk = dslot[k];
                    const U b = dk < s0 ? __builtin_bit_cast(U, y[col[k]])
                                        : __hip_atomic_load(yb + dk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (b == B::kSent) break;
                    acc -= val[k] * __builtin_bit_cast(T, b);
                    ++e;
                }
                ready = e == L;
            }
            if (!ready && (timed_out || wall_clock64() - t0 >= wait_ticks)) {
                timed_out = true;
                acc = __builtin_bit_cast(T, B::kQuiet);
                ready = true;
            }
            if (ready) {
                if constexpr (DIV) acc = acc / d;
                U ab = __builtin_bit_cast(U, acc);
                if (ab == B::kSent) ab = B::kQuiet;
                __hip_atomic_store(yb + s, ab, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                y[i] = __builtin_bit_cast(T, ab);
                done = true;
            } else {
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
    // the workgroup that finishes last resets the counters for the next sweep
    if (threadIdx.x == 0 && atomicAdd(tick + 1, 1u) == gridDim.x - 1) {
        __hip_atomic_store(tick, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(tick + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// Block sweep: one workgroup per block of kBlkRows consecutive rows
// ---------------------------------------------------------------------------
// A block's rows, ordered by level, are processed in steps (the block's rows
// of one level, 64 at a time). W wavefronts take the steps in turn (step q
// by wavefront q mod W) and hand over through a step counter in LDS, so a
// step costs an LDS round trip while each wavefront has W steps of time for
// the loads of its next one. Results of the block stay in LDS, where later
// steps read them. A row of an earlier block is read from ys (slot order, a
// signalling-NaN sentinel until the row is done): rows a later block reads
// are published there with relaxed device-scope stores (the value is its own
// flag), first read with a plain load kXLead steps before use (it returns
// the sentinel or the final value, the only two values the slot holds) and,
// if still pending, polled with device-scope loads. Blocks are taken in
// ticket order by persistent workgroups, so every block a workgroup waits
// on is running or done and the sweep always progresses; a wait that runs
// out of time (wait_ticks) stores NaN, so the solve reports a breakdown
// instead of hanging. Per-row arithmetic is the level kernel's: rhs first,
// then the entries in ascending order, then the division.
template <typename T, bool DIV, int KC, int D, int W>
__global__ __launch_bounds__(64 * W) void k_sweep_blk(uint32_t n, uint32_t nblk, const uint32_t* __restrict__ sptr,
                                                       const uint16_t* __restrict__ stp,
                                                       const uint32_t* __restrict__ brow,
                                                       const uint16_t* __restrict__ blen,
                                                       const uint32_t* __restrict__ bdep, const T* __restrict__ bval,
                                                       const T* __restrict__ bdiag, const T* __restrict__ rhs,
                                                       T* __restrict__ y, T* ys, uint32_t* tick, long long wait_ticks,
                                                       Gate g) {
    static_assert(D >= 2, "stage B reads the row fetched one own step earlier");
    using B = SfBits<T>;
    using U = typename B::U;
    if (g.closed()) return;
    // results and step starts of one block, the steps done and the block
    // index: two workgroups fit one CU's 160 KB
    __shared__ T res[kBlkRows];
    __shared__ uint16_t st[kBlkRows];
    __shared__ uint32_t done_sh, blk_sh;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    U* yb = reinterpret_cast<U*>(ys);
    const U* yplain = yb;
    bool timed_out = false;
    struct Row {
        uint32_t on, L, row, gs;
        T d, r;
        uint32_t dep[KC];
        T v[KC];
        U x[KC];
    };
    for (;;) {
        if (threadIdx.x == 0) {
            blk_sh = atomicAdd(tick, 1u);
            done_sh = 0;
        }
        __syncthreads();
        const uint32_t b = blk_sh;
        if (b >= nblk) break;  // uniform over the workgroup
        const uint32_t g0 = b * kBlkRows;
        const uint32_t cnt = min(kBlkRows, n - g0);
        const uint32_t s0 = sptr[b], ns = sptr[b + 1] - s0;
        for (uint32_t t = threadIdx.x; t < ns; t += 64 * W) st[t] = stp[s0 + t];
        __syncthreads();
        // Wavefront w takes steps w, w + W, w + 2W, ... ("own steps"). Stage
        // A (D own steps ahead) loads what the slot addresses; stage B (one
        // own step ahead) the right-hand side, which stage A's row index
        // addresses; stage X the first readings of the external inputs,
        // kXLead steps before use. Loads are issued by every lane (lanes
        // without a row use slot g0) and in the order they are consumed, so
        // the compiler's waits (which count loads in issue order) never drain
        // the loads of later steps.
        auto stage_a = [&](uint32_t q, Row& R) {
            const uint32_t qc = q < ns ? q : ns - 1;
            const uint32_t ls = (uint32_t)st[qc] + lane;
            const uint32_t end = qc + 1 < ns ? (uint32_t)st[qc + 1] : cnt;
            R.on = (q < ns && ls < end) ? 1u : 0u;
            const uint32_t gs = g0 + (R.on ? ls : 0u);
            R.gs = gs;
            R.row = brow[gs];
            R.L = blen[gs];
            R.d = DIV ? bdiag[gs] : (T)1;
            const uint64_t o = (uint64_t)gs * KC;
#pragma unroll
            for (int e = 0; e < KC; ++e) {
                R.dep[e] = bdep[o + e];
                R.v[e] = bval[o + e];
            }
        };
        auto stage_b = [&](Row& R) { R.r = rhs[R.row]; };
        auto stage_x = [&](Row& R) {
#pragma unroll
            for (int e = 0; e < KC; ++e) {
                const uint32_t dp = R.dep[e];
                // a plain load: it returns the sentinel or the final value
                // (the only two values the slot holds during the sweep);
                // internal inputs read the never-written slot n + 1
                R.x[e] = yplain[(dp & kExtDep) ? (dp & ~kExtDep) : n + 1];
            }
        };
        const uint32_t nown = ns > wave ? (ns - wave + W - 1) / W : 0u;
        Row buf[D];
#pragma unroll
        for (int k = 0; k < D; ++k) stage_a(wave + W * (uint32_t)k, buf[k]);
        stage_b(buf[0]);
        for (uint32_t jb = 0; jb < nown; jb += D) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const uint32_t j = jb + (uint32_t)k;
                if (j >= nown) break;  // uniform over the wavefront
                const uint32_t q = wave + W * j;
                Row& cur = buf[k];
                const uint32_t L = cur.L & ~(uint32_t)kExtUse;
                // first readings of the external inputs kXLead steps before
                // this one (late enough to find most of them published, early
                // enough for the load to return), then wait until the steps
                // before q are done (their results are in res)
                while (__hip_atomic_load(&done_sh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) + kXLead < q)
                    __builtin_amdgcn_s_sleep(1);
                stage_x(cur);
                while (__hip_atomic_load(&done_sh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != q) {
                }
                // external inputs still pending at stage B: poll them
                bool late = false;
#pragma unroll
                for (int e = 0; e < KC; ++e)
                    late |= cur.on && (uint32_t)e < L && (cur.dep[e] & kExtDep) && cur.x[e] == B::kSent;
                if (__any(late)) {
                    const long long w0 = wall_clock64();
#pragma unroll
                    for (int e = 0; e < KC; ++e) {
                        if (cur.on && (uint32_t)e < L && (cur.dep[e] & kExtDep)) {
                            while (cur.x[e] == B::kSent) {
                                if (timed_out || wall_clock64() - w0 >= wait_ticks) {
                                    timed_out = true;
                                    cur.x[e] = B::kQuiet;
                                    break;
                                }
                                cur.x[e] = __hip_atomic_load(yb + (cur.dep[e] & ~kExtDep), __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT);
                            }
                        }
                    }
                }
                T acc = cur.r;
#pragma unroll
                for (int e = 0; e < KC; ++e) {
                    const uint32_t dp = cur.dep[e];
                    const T in = res[(dp & kExtDep) ? 0u : dp];
                    const T yv = (dp & kExtDep) ? __builtin_bit_cast(T, cur.x[e]) : in;
                    const T t = acc - cur.v[e] * yv;
                    acc = (uint32_t)e < L ? t : acc;
                }
                if constexpr (DIV) acc = acc / cur.d;
                U ab = __builtin_bit_cast(U, acc);
                if (ab == B::kSent) ab = B::kQuiet;
                const T out = __builtin_bit_cast(T, ab);
                if (cur.on) res[cur.gs - g0] = out;
                // the step's LDS results, then the hand-off to the next step:
                // a wavefront's LDS operations complete in issue order, so a
                // reader that sees the new count also sees the results (no
                // wait for the stores in between)
                __asm__ volatile("" ::: "memory");
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) __hip_atomic_store(&done_sh, q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (cur.on) {
                    y[cur.row] = out;
                    // only rows a later block reads are published (write-through)
                    if (cur.L & kExtUse)
                        __hip_atomic_store(yb + cur.gs, ab, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                // the next own step's stage B after this step's polls (a poll
                // waits for every load issued before it), then a stage A
                stage_b(buf[(k + 1) % D]);
                stage_a(wave + W * (j + D), cur);
            }
        }
        __syncthreads();  // st and res are rewritten by the next block
    }
    // the workgroup that finishes last resets the counters for the next sweep
    if (threadIdx.x == 0 && atomicAdd(tick + 1, 1u) == gridDim.x - 1) {
        __hip_atomic_store(tick, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(tick + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <typename T, bool DIV, int KC, int D>
static void blk_launch(const Sweep& S, const T* val, const T* diag, const T* rhs, T* y, T* ys, hipStream_t s, Gate g) {
    const uint32_t grid = std::min<uint32_t>(S.nblk, (uint32_t)(2 * sf_cus()));
    hipLaunchKernelGGL((k_sweep_blk<T, DIV, KC, D, kBlkWaves>), dim3(grid), dim3(64 * kBlkWaves), 0, s, S.n, S.nblk,
                       S.bsptr.get(), S.bstp.get(), S.brow.get(), S.blen.get(), S.bdep.get(), val, diag, rhs, y, ys,
                       S.tick.get(), sf_wait_ticks(), g);
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
static void run_sweep(const Sweep& S, const T* rhs, T* y, hipStream_t s, Gate g) {
    const T *cval, *val, *diag;
    T* rslot;
    if constexpr (std::is_same<T, double>::value) {
        cval = S.cval.get();
        rslot = S.rslot.get();
        val = S.val.get();
        diag = DIV ? S.diag.get() : nullptr;
    } else {
        cval = S.cval32.get();
        rslot = S.rslot32.get();
        val = S.val32.get();
        diag = DIV ? S.diag32.get() : nullptr;
    }
    T* ys;
    if constexpr (std::is_same<T, double>::value) ys = S.ys.get();
    else ys = S.ys32.get();
    if (S.blk) {
        const T *bv, *bd;
        if constexpr (std::is_same<T, double>::value) {
            bv = S.bval.get();
            bd = DIV ? S.bdiag.get() : nullptr;
        } else {
            bv = S.bval32.get();
            bd = DIV ? S.bdiag32.get() : nullptr;
        }
        hipLaunchKernelGGL((k_sf_fill<T>), dim3((S.n + 255) / 256), dim3(256), 0, s, ys, 0u, S.n, g);
        if (S.bkc == 4) blk_launch<T, DIV, 4, 3>(S, bv, bd, rhs, y, ys, s, g);
        else if (S.bkc == 8) blk_launch<T, DIV, 8, 3>(S, bv, bd, rhs, y, ys, s, g);
        else blk_launch<T, DIV, 16, 2>(S, bv, bd, rhs, y, ys, s, g);
        FEM_HIP(hipGetLastError());
        return;
    }
    for (const Segment& sg : S.segs) {
        if (sg.sf) {
            const uint32_t cnt = sg.s1 - sg.s0;
            hipLaunchKernelGGL((k_sf_fill<T>), dim3((cnt + 255) / 256), dim3(256), 0, s, ys, sg.s0, sg.s1, g);
            // persistent: one workgroup per CU takes slot blocks in ticket order
            const uint32_t grid = std::min<uint32_t>((cnt + kSfThreads - 1) / kSfThreads, (uint32_t)sf_cus());
            hipLaunchKernelGGL((k_sweep_sf<T, DIV>), dim3(grid), dim3(kSfThreads), 0, s, sg.s0, sg.s1,
                               S.slot_row.get(), S.len.get(), S.gptr.get(), S.col.get(), S.dslot.get(), val, diag,
                               rhs, y, ys, S.tick.get() + 2 * sg.tick, sf_wait_ticks(), g);
            continue;
        }
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

// A row-wise selection from a CSR pattern, built in parallel over chunks of
// rows with the serial loop's result: keep(i, p) selects entry p of row i;
// out_p receives the row pointers of the selection, size(total) is called
// once with the number of kept entries, then emit(i, p, q) for every kept
// entry with q its position in the selection (rows in order, entries in
// stored order).
template <typename R, typename Keep, typename Size, typename Emit>
static void select_rows(uint32_t n, const R& rp, Idx& out_p, Keep keep,
                        Size size, Emit emit) {
    constexpr uint32_t kRowsPerChunk = 8192;
    const uint32_t nch = std::max(1u, (n + kRowsPerChunk - 1) / kRowsPerChunk);
    out_p.assign((size_t)n + 1, 0);
    std::vector<uint64_t> base(nch + 1, 0);
    host_parallel(nch, [&](uint32_t c) {
        const uint32_t i0 = c * kRowsPerChunk, i1 = std::min(n, i0 + kRowsPerChunk);
        uint64_t k = 0;
        for (uint32_t i = i0; i < i1; ++i)
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) k += keep(i, p) ? 1 : 0;
        base[c + 1] = k;
    });
    for (uint32_t c = 0; c < nch; ++c) base[c + 1] += base[c];
    if (base[nch] > UINT32_MAX) fail(FEM_ERR_RUNTIME, "triangular factor too large for 32-bit positions");
    size(base[nch]);
    host_parallel(nch, [&](uint32_t c) {
        const uint32_t i0 = c * kRowsPerChunk, i1 = std::min(n, i0 + kRowsPerChunk);
        uint32_t q = (uint32_t)base[c];
        for (uint32_t i = i0; i < i1; ++i) {
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
                if (keep(i, p)) emit(i, p, q++);
            out_p[i + 1] = q;
        }
    });
}

// The transposed lists of a lower-triangular selection: for every entry p of
// row i with column j = cols[p] < i, row i is listed under j with position
// p; each list ascending in i (then p), as a serial pass over the rows in
// order produces. Counted and filled in parallel, each list then sorted.
static void transpose_lists(uint32_t n, const Idx& rp, const Idx& cols, Idx& Tp, Idx& Tc, Idx& Tpos) {
    constexpr uint32_t kRowsPerChunk = 8192;
    const uint32_t nch = std::max(1u, (n + kRowsPerChunk - 1) / kRowsPerChunk);
    Tp.assign((size_t)n + 1, 0);
    host_parallel(nch, [&](uint32_t c) {
        const uint32_t i0 = c * kRowsPerChunk, i1 = std::min(n, i0 + kRowsPerChunk);
        for (uint32_t i = i0; i < i1; ++i)
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
                if (cols[p] < i) __atomic_fetch_add(&Tp[cols[p] + 1], 1u, __ATOMIC_RELAXED);
    });
    for (uint32_t i = 0; i < n; ++i) Tp[i + 1] += Tp[i];
    Tc.resize(Tp[n]);
    Tpos.resize(Tp[n]);
    std::vector<uint32_t> off(Tp.begin(), Tp.end() - 1);
    host_parallel(nch, [&](uint32_t c) {
        const uint32_t i0 = c * kRowsPerChunk, i1 = std::min(n, i0 + kRowsPerChunk);
        for (uint32_t i = i0; i < i1; ++i)
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
                if (cols[p] < i) {
                    const uint32_t q = __atomic_fetch_add(&off[cols[p]], 1u, __ATOMIC_RELAXED);
                    Tc[q] = i;
                    Tpos[q] = p;
                }
    });
    host_parallel(nch, [&](uint32_t c) {
        const uint32_t j0 = c * kRowsPerChunk, j1 = std::min(n, j0 + kRowsPerChunk);
        for (uint32_t j = j0; j < j1; ++j)  // insertion sort by (row, position): lists are short
            for (uint32_t a = Tp[j] + 1; a < Tp[j + 1]; ++a)
                for (uint32_t b = a; b > Tp[j] && (Tc[b - 1] > Tc[b] || (Tc[b - 1] == Tc[b] && Tpos[b - 1] > Tpos[b]));
                     --b) {
                    std::swap(Tc[b - 1], Tc[b]);
                    std::swap(Tpos[b - 1], Tpos[b]);
                }
    });
}

// levels of a triangular dependency graph given per-row dependency lists;
// forward = dependencies have smaller index (process ascending)
static Idx levels_of(uint32_t n, const Idx& rp, const Idx& cols,
                                       bool forward) {
    Idx lev(n);  // every row's level is written before a later row reads it
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

// a sweep's values (and diagonal, if it divides) from the factor values fac
static void sweep_values(Sweep& S, const double* fac, hipStream_t s) {
    if (S.blk) {
        gather(fac, S.bmap.get(), S.bval.get(), (uint64_t)S.n * S.bkc, s);
        if (S.bdiag.get()) gather(fac, S.bdmap.get(), S.bdiag.get(), S.n, s);
        return;
    }
    gather(fac, S.vmap.get(), S.val.get(), S.nent, s);
    if (S.diag.get()) gather(fac, S.dmap.get(), S.diag.get(), S.n, s);
    densify(S, s);
}

// the float32 copies of a sweep's values (and its float32 work arrays)
static void sweep_values_f32(Sweep& S, hipStream_t s) {
    if (S.blk) {
        to_f32(S.bval, S.bval32, (uint64_t)S.n * S.bkc, s);
        if (S.bdiag.get()) to_f32(S.bdiag, S.bdiag32, S.n, s);
        S.ys32.alloc(S.n + 2);
        return;
    }
    to_f32(S.val, S.val32, S.nent, s);
    if (S.diag.get()) to_f32(S.diag, S.diag32, S.n, s);
    densify(S, s);
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
        SetupTrace tr("IC(0) setup", s);
        A.sync_host();
        tr.lap("matrix on the host");
        const auto& rp = A.h_rp;
        const auto& ci = A.h_ci;
        // lower pattern with diagonal (array.cpp:2539-2574)
        Idx Lp, Lc, amap, dpos(n, UINT32_MAX);
        select_rows(
            n, rp, Lp, [&](uint32_t i, uint32_t p) { return ci[p] <= i; },
            [&](uint64_t k) {
                Lc.resize(k);
                amap.resize(k);
            },
            [&](uint32_t i, uint32_t p, uint32_t q) {
                if (ci[p] == i && dpos[i] == UINT32_MAX) dpos[i] = q;
                Lc[q] = ci[p];
                amap[q] = p;
            });
        // The reference stops a row's off-diagonal loop at the first column
        // >= i (array.cpp:2589); with sorted columns that is the diagonal.
        for (uint32_t i = 0; i < n; ++i)
            if (dpos[i] == UINT32_MAX) fail(FEM_ERR_RUNTIME, "IC(0) failed: the matrix lacks a diagonal entry");
        const uint64_t nL = Lc.size();
        tr.lap("lower pattern");
        // strictly-lower dependency lists (forward) and L^T lists (backward)
        Idx Fp, Fc, Fpos;
        select_rows(
            n, Lp, Fp, [&](uint32_t i, uint32_t p) { return p < dpos[i]; },
            [&](uint64_t k) {
                Fc.resize(k);
                Fpos.resize(k);
            },
            [&](uint32_t, uint32_t p, uint32_t q) {
                Fc[q] = Lc[p];
                Fpos[q] = p;
            });
        tr.lap("forward lists");
        // L^T holds every strictly-lower entry of L (array.cpp:2659-2695)
        Idx Bp, Bc, Bpos;
        transpose_lists(n, Lp, Lc, Bp, Bc, Bpos);
        tr.lap("dependency lists");
        Idx lb;  // the two sweeps' levels side by side
        std::thread tb([&] { lb = levels_of(n, Bp, Bc, false); });
        const auto lf = levels_of(n, Fp, Fc, true);
        tb.join();
        tr.lap("levels");
        fw.build(n, Fp, Fc, Fpos, lf, &dpos, true, s);
        bw.build(n, Bp, Bc, Bpos, lb, &dpos, false, s);
        levels = (int)fw.nlev;
        tr.lap("sweep layouts");
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
        tr.lap("numeric factorization");
        sweep_values(fw, Lv.get(), s);
        sweep_values(bw, Lv.get(), s);
        y.alloc(n);
        FEM_HIP(hipStreamSynchronize(s));
        tr.lap("sweep values");
    }
    int kind() const override { return FEM_PC_IC0; }
    uint32_t size() const override { return n; }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override {
        run_sweep<double, true>(fw, r, y.get(), s, g);
        run_sweep<double, true>(bw, y.get(), z, s, g);
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override {
        run_sweep<float, true>(fw, r, y32.get(), s, g);
        run_sweep<float, true>(bw, y32.get(), z, s, g);
    }
    double apply_bytes(int s) override { return sweep_bytes(fw, s, true) + sweep_bytes(bw, s, true); }
    bool has_fp32() const override { return y32.get() != nullptr; }
    void ensure_fp32(hipStream_t s) override {
        if (y32.get()) return;
        sweep_values_f32(fw, s);
        sweep_values_f32(bw, s);
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
        SetupTrace tr("ILU(0) setup", s);
        A.sync_host();
        tr.lap("matrix on the host");
        const auto& rp = A.h_rp;
        const auto& ci = A.h_ci;
        Idx Lp, Up, Lc, Uc, Lmap, Umap, dU(n, UINT32_MAX);
        select_rows(
            n, rp, Lp, [&](uint32_t i, uint32_t p) { return ci[p] < i; },
            [&](uint64_t k) {
                Lc.resize(k);
                Lmap.resize(k);
            },
            [&](uint32_t, uint32_t p, uint32_t q) {
                Lc[q] = ci[p];
                Lmap[q] = p;
            });
        select_rows(
            n, rp, Up, [&](uint32_t i, uint32_t p) { return ci[p] >= i; },
            [&](uint64_t k) {
                Uc.resize(k);
                Umap.resize(k);
            },
            [&](uint32_t i, uint32_t p, uint32_t q) {
                if (ci[p] == i) dU[i] = q;
                Uc[q] = ci[p];
                Umap[q] = p;
            });
        for (uint32_t i = 0; i < n; ++i)
            if (dU[i] == UINT32_MAX) fail(FEM_ERR_RUNTIME, "ILU(0) failed: the matrix lacks a diagonal entry");
        // backward dependency lists: U row without the diagonal
        Idx Bp, Bc, Bpos;
        select_rows(
            n, Up, Bp, [&](uint32_t i, uint32_t p) { return Uc[p] != i; },
            [&](uint64_t k) {
                Bc.resize(k);
                Bpos.resize(k);
            },
            [&](uint32_t, uint32_t p, uint32_t q) {
                Bc[q] = Uc[p];
                Bpos[q] = p;
            });
        tr.lap("patterns");
        Idx Lpos(Lc.size());
        std::iota(Lpos.begin(), Lpos.end(), 0u);
        Idx lb;  // the two sweeps' levels side by side
        std::thread tb([&] { lb = levels_of(n, Bp, Bc, false); });
        const auto lf = levels_of(n, Lp, Lc, true);
        tb.join();
        tr.lap("levels");
        fw.build(n, Lp, Lc, Lpos, lf, nullptr, true, s);
        bw.build(n, Bp, Bc, Bpos, lb, &dU, false, s);
        levels = (int)fw.nlev;
        tr.lap("sweep layouts");
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
        tr.lap("numeric factorization");
        sweep_values(fw, Lv.get(), s);
        sweep_values(bw, Uv.get(), s);
        y.alloc(n);
        FEM_HIP(hipStreamSynchronize(s));
        tr.lap("sweep values");
    }
    int kind() const override { return FEM_PC_ILU0; }
    uint32_t size() const override { return n; }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override {
        run_sweep<double, false>(fw, r, y.get(), s, g);
        run_sweep<double, true>(bw, y.get(), z, s, g);
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override {
        run_sweep<float, false>(fw, r, y32.get(), s, g);
        run_sweep<float, true>(bw, y32.get(), z, s, g);
    }
    double apply_bytes(int s) override { return sweep_bytes(fw, s, false) + sweep_bytes(bw, s, true); }
    bool has_fp32() const override { return y32.get() != nullptr; }
    void ensure_fp32(hipStream_t s) override {
        if (y32.get()) return;
        sweep_values_f32(fw, s);
        sweep_values_f32(bw, s);
        y32.alloc(n);
    }
};

std::unique_ptr<Precond> make_ic0(CsrMat& A, hipStream_t s) { return std::unique_ptr<Precond>(new PcIC0(A, s)); }
std::unique_ptr<Precond> make_ilu0(CsrMat& A, hipStream_t s) { return std::unique_ptr<Precond>(new PcILU0(A, s)); }

}  // namespace fem

// amg_setup.cpp — host construction of the smoothed-aggregation hierarchy
// (block AMG, ode-solver/include/math/blockMatrix.hpp:1051-1179): the
// reference's greedy aggregation (explicit weights) and the host build of the
// robust variant. The C/F hierarchy and the robust SA hierarchy are built on
// the device (amg_device.hip).
// The coarse products are exact sparse products accumulated in the
// reference's traversal order (row p of A, then row q of B). The coarsest
// solve is stored as an explicit dense operator C (C f = coarse solve of f,
// column k = solve of e_k), computed here in float64.
#include <algorithm>
#include <cmath>
#include <map>
#include <queue>

#include "amg.hpp"

namespace fem {

HCsr transpose(const HCsr& A) {
    HCsr T;
    T.n = A.m;
    T.m = A.n;
    T.rp.assign(A.m + 1, 0);
    for (uint32_t c : A.ci) T.rp[c + 1]++;
    for (uint32_t j = 0; j < A.m; ++j) T.rp[j + 1] += T.rp[j];
    T.ci.resize(A.ci.size());
    T.v.resize(A.v.size());
    std::vector<uint32_t> off(T.rp.begin(), T.rp.end() - 1);
    for (uint32_t i = 0; i < A.n; ++i)
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            const uint32_t q = off[A.ci[p]]++;
            T.ci[q] = i;
            T.v[q] = A.v[p];
        }
    return T;
}

HCsr spgemm(const HCsr& A, const HCsr& B) {
    HCsr C;
    C.n = A.n;
    C.m = B.m;
    C.rp.assign(A.n + 1, 0);
    std::vector<int64_t> mark(B.m, -1);
    std::vector<double> acc(B.m, 0.0);
    std::vector<uint32_t> cols;
    C.ci.reserve(A.ci.size() * 2);
    C.v.reserve(A.ci.size() * 2);
    for (uint32_t i = 0; i < A.n; ++i) {
        cols.clear();
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            const uint32_t j = A.ci[p];
            const double a = A.v[p];
            for (uint32_t q = B.rp[j]; q < B.rp[j + 1]; ++q) {
                const uint32_t k = B.ci[q];
                if (mark[k] != (int64_t)i) { mark[k] = i; cols.push_back(k); acc[k] = 0.0; }
                acc[k] += a * B.v[q];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (uint32_t k : cols) { C.ci.push_back(k); C.v.push_back(acc[k]); }
        C.rp[i + 1] = (uint32_t)C.ci.size();
    }
    return C;
}

static HCsr from_mat(const CsrMat& A) {
    A.sync_host();
    HCsr H;
    H.n = A.n;
    H.m = A.m;
    H.rp = A.h_rp;
    H.ci = A.h_ci;
    H.v = A.h_v;
    return H;
}

static void check_coarse(uint32_t nc) {
    if (nc > 4096) fail(FEM_ERR_RUNTIME, "AMG: coarsest level has " + std::to_string(nc) +
                                             " rows; raise max_levels or coarse_size");
}

// 1x1 condensation (blockMatrix.hpp:269-315 with BS = 1): columns sorted and
// unique per row, duplicates summed, explicit zeros kept.
static HCsr condense1(const HCsr& A) {
    HCsr B;
    B.n = A.n;
    B.m = A.m;
    B.rp.assign(A.n + 1, 0);
    std::vector<std::pair<uint32_t, double>> row;
    for (uint32_t i = 0; i < A.n; ++i) {
        row.clear();
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) row.emplace_back(A.ci[p], A.v[p]);
        std::stable_sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < row.size();) {
            const uint32_t j = row[k].first;
            double s = 0.0 + row[k].second;
            size_t e = k + 1;
            while (e < row.size() && row[e].first == j) s += row[e++].second;
            B.ci.push_back(j);
            B.v.push_back(s);
            k = e;
        }
        B.rp[i + 1] = (uint32_t)B.ci.size();
    }
    return B;
}

// coarsest solve of the block AMG (blockMatrix.hpp:958-999) as a dense
// operator. exact = 0 reproduces the reference's sequence (the right-hand
// side runs through the already-eliminated sub-diagonal entries without the
// row swaps); exact = 1 is the pivoted LU solve. Pivots below 1e-14 in
// magnitude are replaced by +-1e-14 in both.
std::vector<double> sa_coarse_operator_host(const HCsr& A, bool exact) {
    const int sz = (int)A.n;
    check_coarse(A.n);
    std::vector<double> d((size_t)sz * sz, 0.0);
    for (int i = 0; i < sz; ++i)
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) d[(size_t)i * sz + A.ci[p]] = A.v[p];
    auto fl = [](double x) { return std::fabs(x) < 1e-14 ? (x >= 0 ? 1e-14 : -1e-14) : x; };
    std::vector<int> pv(sz);
    for (int i = 0; i < sz; ++i) pv[i] = i;
    for (int k = 0; k < sz; ++k) {
        int pk = k;
        for (int i = k + 1; i < sz; ++i)
            if (std::fabs(d[(size_t)i * sz + k]) > std::fabs(d[(size_t)pk * sz + k])) pk = i;
        if (pk != k) {
            std::swap(pv[k], pv[pk]);
            for (int j = 0; j < sz; ++j) std::swap(d[(size_t)k * sz + j], d[(size_t)pk * sz + j]);
        }
        if (exact) {
            d[(size_t)k * sz + k] = fl(d[(size_t)k * sz + k]);
            for (int i = k + 1; i < sz; ++i) {
                d[(size_t)i * sz + k] /= d[(size_t)k * sz + k];
                for (int j = k + 1; j < sz; ++j) d[(size_t)i * sz + j] -= d[(size_t)i * sz + k] * d[(size_t)k * sz + j];
            }
        } else {
            const double dk = fl(d[(size_t)k * sz + k]);
            for (int i = k + 1; i < sz; ++i) {
                const double f = d[(size_t)i * sz + k] / dk;
                for (int j = k; j < sz; ++j) d[(size_t)i * sz + j] -= f * d[(size_t)k * sz + j];
            }
        }
    }
    // C = A^-1: every column of the identity solved at once, row by row
    // (row i of C holds element i of every column's solution), so the inner
    // loops run over contiguous columns; each element sees the same
    // operations in the same order as a column-by-column solve.
    std::vector<double> C((size_t)sz * sz, 0.0);
    auto row = [&](int i) { return C.data() + (size_t)i * sz; };
    if (exact) {
        for (int i = 0; i < sz; ++i) row(i)[pv[i]] = 1.0;
        for (int i = 0; i < sz; ++i)
            for (int j = 0; j < i; ++j) {
                const double f = d[(size_t)i * sz + j];
                double* ci = row(i);
                const double* cj = row(j);
                for (int c = 0; c < sz; ++c) ci[c] -= f * cj[c];
            }
        for (int i = sz - 1; i >= 0; --i) {
            double* ci = row(i);
            for (int j = i + 1; j < sz; ++j) {
                const double f = d[(size_t)i * sz + j];
                const double* cj = row(j);
                for (int c = 0; c < sz; ++c) ci[c] -= f * cj[c];
            }
            const double di = d[(size_t)i * sz + i];
            for (int c = 0; c < sz; ++c) ci[c] /= di;
        }
    } else {
        for (int i = 0; i < sz; ++i) row(i)[i] = 1.0;
        for (int k = 0; k < sz; ++k) {
            const double dk = fl(d[(size_t)k * sz + k]);
            const double* ck = row(k);
            for (int i = k + 1; i < sz; ++i) {
                const double f = d[(size_t)i * sz + k] / dk;
                double* ci = row(i);
                for (int c = 0; c < sz; ++c) ci[c] -= f * ck[c];
            }
        }
        for (int i = sz - 1; i >= 0; --i) {
            double* ci = row(i);
            for (int j = i + 1; j < sz; ++j) {
                const double f = d[(size_t)i * sz + j];
                const double* cj = row(j);
                for (int c = 0; c < sz; ++c) ci[c] -= f * cj[c];
            }
            const double di = fl(d[(size_t)i * sz + i]);
            for (int c = 0; c < sz; ++c) ci[c] /= di;
        }
    }
    return C;
}

// Node priority for the aggregation rounds: the "lowbias32" integer hash
// (C. Wellons, "Prospecting for hash functions", 2018) of the row index.
static inline uint32_t node_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Aggregation by a distance-2 maximal independent set of the strength graph
// (Bell, Dalton & Olson, "Exposing fine-grained parallelism in algebraic
// multigrid methods", SIAM J. Sci. Comput. 34(4), 2012). No counterpart in
// the reference, whose serial greedy aggregates of at most max_agg rows
// coarsen a 3-D 7-point operator by about 2x per level while the Galerkin
// operators fill in.
//   strength: j != i is strong for i when |a_ij| >= theta * sqrt(m_i m_j),
//             m_i = max_{k != i} |a_ik| (symmetric for symmetric A);
//   MIS-2:    synchronous rounds over keys (state, hash, index), state
//             2 = root > 1 = undecided > 0 = out; a node whose two-hop maximum
//             is itself becomes a root, one that sees a root within two hops
//             leaves;
//   phase 1:  a non-root joins its strong root neighbour of largest key;
//   phase 2:  a node still free joins the phase-1 neighbour of largest key;
//   phase 3:  what is left (only possible for unsymmetric strength) becomes
//             singleton aggregates in index order.
// Every step is a function of the previous round's arrays only, so the
// result does not depend on the order rows are visited in.
int mis2_aggregate(const HCsr& A, double theta, std::vector<int>& agg) {
    const uint32_t n = A.n;
    std::vector<double> mrow(n, 0.0);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
            if (A.ci[p] != i) mrow[i] = std::max(mrow[i], std::sqrt(A.v[p] * A.v[p]));
    std::vector<uint32_t> Sp(n + 1, 0), Sc;
    Sc.reserve(A.ci.size());
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            const uint32_t j = A.ci[p];
            const double a = std::sqrt(A.v[p] * A.v[p]);
            if (j != i && a != 0.0 && a >= theta * std::sqrt(mrow[i] * mrow[j])) Sc.push_back(j);
        }
        Sp[i + 1] = (uint32_t)Sc.size();
    }
    auto key = [](uint64_t st, uint32_t i) { return (st << 62) | ((uint64_t)(node_hash(i) & 0x3fffffffU) << 32) | i; };
    std::vector<uint8_t> st(n, 1);
    std::vector<uint64_t> t0(n), t1(n);
    for (bool busy = true; busy;) {
        for (uint32_t i = 0; i < n; ++i) t0[i] = key(st[i], i);
        for (uint32_t i = 0; i < n; ++i) {
            uint64_t m = t0[i];
            for (uint32_t p = Sp[i]; p < Sp[i + 1]; ++p) m = std::max(m, t0[Sc[p]]);
            t1[i] = m;
        }
        busy = false;
        for (uint32_t i = 0; i < n; ++i) {
            if (st[i] != 1) continue;
            uint64_t m = t1[i];
            for (uint32_t p = Sp[i]; p < Sp[i + 1]; ++p) m = std::max(m, t1[Sc[p]]);
            t0[i] = m;  // reuse as the two-hop maximum
        }
        for (uint32_t i = 0; i < n; ++i) {
            if (st[i] != 1) continue;
            if ((uint32_t)t0[i] == i) st[i] = 2;
            else if ((t0[i] >> 62) == 2) st[i] = 0;
            else busy = true;
        }
    }
    agg.assign(n, -1);
    int na = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (st[i] == 2) agg[i] = na++;
    std::vector<int> a1(agg);
    for (uint32_t i = 0; i < n; ++i) {
        if (st[i] == 2) continue;
        uint64_t best = 0;
        int g = -1;
        for (uint32_t p = Sp[i]; p < Sp[i + 1]; ++p) {
            const uint32_t j = Sc[p];
            if (st[j] == 2 && key(0, j) >= best) { best = key(0, j); g = agg[j]; }
        }
        a1[i] = g;
    }
    agg = a1;
    for (uint32_t i = 0; i < n; ++i) {
        if (a1[i] >= 0) continue;
        uint64_t best = 0;
        int g = -1;
        for (uint32_t p = Sp[i]; p < Sp[i + 1]; ++p) {
            const uint32_t j = Sc[p];
            if (a1[j] >= 0 && key(0, j) >= best) { best = key(0, j); g = a1[j]; }
        }
        agg[i] = g;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (agg[i] < 0) agg[i] = na++;
    return na;
}

AmgHierarchy build_sa_hierarchy(const CsrMat& A0, const fem_pc_params& prm) {
    AmgHierarchy H;
    HCsr cur = condense1(from_mat(A0));
    const double theta = prm.theta;
    const int maxl = std::max(1, prm.max_levels);
    const int max_agg = std::max(1, prm.max_agg);
    const uint32_t cmax = (uint32_t)std::max(1, prm.coarse_size);
    // p_omega <= 0 / omega <= 0 select the robust variant (no counterpart in
    // the reference): per level w = (4/3) / rho with rho = max_i sum_j
    // |a_ij / a_ii| (Gershgorin bound of D^{-1} A); coarsening also stops when
    // a level keeps more than 85% of its rows.
    const bool auto_p = prm.p_omega <= 0.0, auto_s = prm.omega <= 0.0;
    for (int l = 0; l < maxl; ++l) {
        const uint32_t n = cur.n;
        std::vector<double> dinv(n, 0.0);
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t dp = UINT32_MAX;
            for (uint32_t p = cur.rp[i]; p < cur.rp[i + 1]; ++p)
                if (cur.ci[p] == i) { dp = p; break; }
            if (dp == UINT32_MAX) fail(FEM_ERR_RUNTIME, "Block AMG: block row " + std::to_string(i) + " has no diagonal block");
            // 1x1 Gauss-Jordan inverse (blockMatrix.hpp:70-100); a failed
            // inversion leaves 0, as the reference ignores the failure
            const double a = cur.v[dp];
            if (std::fabs(a) >= 1e-300) dinv[i] = 1.0 / a;
        }
        double rho = 0.0;
        if (auto_p || auto_s) {
            for (uint32_t i = 0; i < n; ++i) {
                double rs = 0.0;
                for (uint32_t p = cur.rp[i]; p < cur.rp[i + 1]; ++p) rs += std::fabs(0.0 + dinv[i] * cur.v[p]);
                rho = std::max(rho, rs);
            }
        }
        const double wauto = rho > 0.0 ? (4.0 / 3.0) / rho : 2.0 / 3.0;
        H.omega.push_back(auto_s ? wauto : prm.omega);
        const double pw = auto_p ? wauto : prm.p_omega;
        H.A.push_back(cur);
        if (n <= cmax) break;
        std::vector<int> agg;
        int na = 0;
        if (auto_p) {
            na = mis2_aggregate(cur, theta, agg);
        } else {
            // strength by (1x1) Frobenius norm (blockMatrix.hpp:1084-1099)
            std::vector<uint32_t> Sp(n + 1, 0), Sc;
            Sc.reserve(cur.ci.size());
            for (uint32_t i = 0; i < n; ++i) {
                double mx = 0.0;
                for (uint32_t p = cur.rp[i]; p < cur.rp[i + 1]; ++p)
                    if (cur.ci[p] != i) mx = std::max(mx, std::sqrt(cur.v[p] * cur.v[p]));
                const double thr = theta * mx;
                for (uint32_t p = cur.rp[i]; p < cur.rp[i + 1]; ++p)
                    if (cur.ci[p] != i && std::sqrt(cur.v[p] * cur.v[p]) >= thr) Sc.push_back(cur.ci[p]);
                Sp[i + 1] = (uint32_t)Sc.size();
            }
            // greedy aggregation (blockMatrix.hpp:1101-1111)
            agg.assign(n, -1);
            for (uint32_t i = 0; i < n; ++i) {
                if (agg[i] != -1) continue;
                agg[i] = na;
                int sz = 1;
                for (uint32_t p = Sp[i]; p < Sp[i + 1]; ++p) {
                    const uint32_t j = Sc[p];
                    if (agg[j] == -1 && sz < max_agg) { agg[j] = na; ++sz; }
                }
                ++na;
            }
        }
        // smoothed prolongator P = (I - w D^{-1} A) P_t (blockMatrix.hpp:1125-1164)
        HCsr P;
        P.n = n;
        P.m = (uint32_t)na;
        P.rp.assign(n + 1, 0);
        std::vector<std::pair<uint32_t, double>> row;
        for (uint32_t i = 0; i < n; ++i) {
            row.clear();
            row.emplace_back((uint32_t)agg[i], 1.0);
            for (uint32_t p = cur.rp[i]; p < cur.rp[i + 1]; ++p) {
                const uint32_t g = (uint32_t)agg[cur.ci[p]];
                const double sv = 0.0 + dinv[i] * cur.v[p];
                auto it = std::find_if(row.begin(), row.end(), [g](auto& e) { return e.first == g; });
                if (it == row.end()) { row.emplace_back(g, 0.0); it = row.end() - 1; }
                it->second -= pw * sv;
            }
            std::sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
            for (auto& e : row) { P.ci.push_back(e.first); P.v.push_back(e.second); }
            P.rp[i + 1] = (uint32_t)P.ci.size();
        }
        if (auto_p && (double)na > 0.85 * (double)n) break;  // coarsening stalled
        HCsr R = transpose(P);
        HCsr AP = spgemm(cur, P);
        HCsr Ac = spgemm(R, AP);
        H.d.push_back(std::move(dinv));
        H.P.push_back(std::move(P));
        cur = std::move(Ac);
    }
    H.coarse = sa_coarse_operator_host(H.A.back(), prm.coarse_exact != 0);
    return H;
}

}  // namespace fem

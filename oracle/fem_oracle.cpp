// fem_oracle.cpp — CPU restatement of the reference's sparse solve path.
//
// TEST INFRASTRUCTURE ONLY. The GPU library's results are checked against
// it: only tests/, the smoke check of __graft_entry__ and the cpu_baseline
// leg of bench.py may load it.
// The product (libfemgpu) never links or calls it.
//
// Each function cites the reference file:line it restates
// (array.cpp = ode-solver/src/math/array.cpp,
//  block.hpp = ode-solver/include/math/blockMatrix.hpp,
//  test_BlockMatrix_mpi.cpp = ode-solver/test/test_BlockMatrix_mpi.cpp).
// Deliberate differences from the reference (documented, numerically inert):
//  * SpGEMM computes the mathematically correct product; the reference's
//    marker reuse between its two passes (array.cpp:2356,2380-2383) can drop an
//    entry whose column is touched by a single row.
//  * C/F-AMG prolongators carry no padding zeros (array.cpp:3282 reserves one
//    slot per strong neighbour but writes only strong C neighbours).
//  * The block SpGEMM uses a sparse accumulator instead of clearing a dense
//    hit array per row (block.hpp:896); per-entry summation order is the same.
// Built with -ffp-contract=off so a*b+c is never fused, like the x86-64
// reference build.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <queue>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using vec = std::vector<double>;

// ---------------------------------------------------------------------------
// CSR container (array.hpp:266-306)
// ---------------------------------------------------------------------------
struct Csr {
    uint32_t n = 0, m = 0;
    std::vector<uint32_t> rp{0}, ci;
    vec v;
    uint32_t nnz() const { return (uint32_t)v.size(); }
};

// COO constructor (array.cpp:2072-2182): validate, drop v==0, sort (i,j),
// merge duplicates by sum, drop merged zeros, build CSR.
static Csr csr_from_coo(uint32_t n, uint32_t m, const vec& vals,
                        const std::vector<uint32_t>& rl,
                        const std::vector<uint32_t>& cl) {
    if (vals.size() != rl.size() || vals.size() != cl.size())
        throw std::invalid_argument("COO input: values, rowlist, collist lengths differ");
    using E = std::tuple<uint32_t, uint32_t, double>;
    std::vector<E> e;
    e.reserve(vals.size());
    for (size_t k = 0; k < vals.size(); ++k) {
        if (rl[k] >= n || cl[k] >= m) throw std::out_of_range("COO input: row or column index out of range");
        if (vals[k] != 0.0) e.emplace_back(rl[k], cl[k], vals[k]);
    }
    std::sort(e.begin(), e.end(), [](const E& a, const E& b) {
        if (std::get<0>(a) != std::get<0>(b)) return std::get<0>(a) < std::get<0>(b);
        return std::get<1>(a) < std::get<1>(b);
    });
    std::vector<E> mg;
    mg.reserve(e.size());
    for (auto& x : e) {
        if (!mg.empty() && std::get<0>(mg.back()) == std::get<0>(x) && std::get<1>(mg.back()) == std::get<1>(x))
            std::get<2>(mg.back()) += std::get<2>(x);
        else
            mg.push_back(x);
    }
    Csr A;
    A.n = n;
    A.m = m;
    A.rp.assign(n + 1, 0);
    for (auto& x : mg)
        if (std::get<2>(x) != 0.0) A.rp[std::get<0>(x) + 1]++;
    for (uint32_t i = 0; i < n; ++i) A.rp[i + 1] += A.rp[i];
    A.ci.resize(A.rp[n]);
    A.v.resize(A.rp[n]);
    std::vector<uint32_t> off(A.rp.begin(), A.rp.end() - 1);
    for (auto& x : mg) {
        if (std::get<2>(x) == 0.0) continue;
        uint32_t p = off[std::get<0>(x)]++;
        A.ci[p] = std::get<1>(x);
        A.v[p] = std::get<2>(x);
    }
    return A;
}

// transpose (array.cpp:2202-2235): counting sort by column; rows ascend within
// each output row.
static Csr transpose(const Csr& A) {
    Csr T;
    T.n = A.m;
    T.m = A.n;
    T.rp.assign(A.m + 1, 0);
    for (uint32_t p = 0; p < A.nnz(); ++p) T.rp[A.ci[p] + 1]++;
    for (uint32_t j = 0; j < A.m; ++j) T.rp[j + 1] += T.rp[j];
    T.ci.resize(A.nnz());
    T.v.resize(A.nnz());
    std::vector<uint32_t> off(T.rp.begin(), T.rp.end() - 1);
    for (uint32_t i = 0; i < A.n; ++i)
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            uint32_t q = off[A.ci[p]]++;
            T.ci[q] = i;
            T.v[q] = A.v[p];
        }
    return T;
}

// SpGEMM C = A*B (array.cpp:2336-2447), correct form: per row, accumulate in
// traversal order (p over A's row, q over B's row), columns sorted ascending.
static Csr spgemm(const Csr& A, const Csr& B) {
    if (A.m != B.n) throw std::invalid_argument("spgemm: inner dimensions differ");
    Csr C;
    C.n = A.n;
    C.m = B.m;
    C.rp.assign(A.n + 1, 0);
    std::vector<int64_t> mark(B.m, -1);
    vec acc(B.m, 0.0);
    std::vector<uint32_t> cols;
    for (uint32_t i = 0; i < A.n; ++i) {
        cols.clear();
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            const uint32_t j = A.ci[p];
            const double a = A.v[p];
            for (uint32_t q = B.rp[j]; q < B.rp[j + 1]; ++q) {
                const uint32_t k = B.ci[q];
                if (mark[k] != (int64_t)i) {
                    mark[k] = i;
                    cols.push_back(k);
                    acc[k] = 0.0;
                }
                acc[k] += a * B.v[q];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (uint32_t k : cols) {
            C.ci.push_back(k);
            C.v.push_back(acc[k]);
        }
        C.rp[i + 1] = (uint32_t)C.ci.size();
    }
    return C;
}

// dot (array.cpp:3401-3416)
static double dot(const vec& a, const vec& b) {
    if (a.size() != b.size()) throw std::invalid_argument("dot: vector lengths differ");
    double s = 0.0;
    const int64_t n = (int64_t)a.size();
#pragma omp parallel for reduction(+ : s) schedule(static) if (n > 4000)
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// spmv (array.cpp:3418-3442): per row, sequential sum in stored order from 0.0.
static void spmv(const Csr& A, const vec& x, vec& y) {
    if (A.m != x.size()) throw std::invalid_argument("spmv: matrix columns differ from vector length");
    y.assign(A.n, 0.0);
    const int64_t n = A.n;
#pragma omp parallel for schedule(static) if (n > 1000)
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) s += A.v[k] * x[A.ci[k]];
        y[i] = s;
    }
}

// ---------------------------------------------------------------------------
// Preconditioners
// ---------------------------------------------------------------------------
struct Pc {
    virtual ~Pc() {}
    virtual void apply(const vec& r, vec& z) = 0;
};

struct PcNone : Pc {
    void apply(const vec& r, vec& z) override { z = r; }
};

// make_diag_preconditioner (array.cpp:2450-2508)
struct PcDiag : Pc {
    vec d;
    explicit PcDiag(const Csr& A) {
        if (A.n != A.m) throw std::invalid_argument("diag preconditioner needs a square matrix");
        d.assign(A.n, 0.0);
        for (uint32_t i = 0; i < A.n; ++i)
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
                if (A.ci[p] == i) { d[i] = A.v[p]; break; }
        for (uint32_t i = 0; i < A.n; ++i) {
            if (std::fabs(d[i]) < 1e-14) throw std::runtime_error("diag preconditioner: tiny or zero diagonal entry");
            if (d[i] <= 0.0) throw std::runtime_error("diag preconditioner: diagonal must be positive");
        }
    }
    void apply(const vec& r, vec& z) override {
        z.resize(r.size());
        const int64_t n = r.size();
#pragma omp parallel for schedule(static) if (n > 1000)
        for (int64_t i = 0; i < n; ++i) z[i] = r[i] / d[i];
    }
};

// make_ic0_preconditioner (array.cpp:2509-2732): lower CSR with diagonal,
// left-looking IC(0) by two-pointer merges, explicit L^T for the backward sweep.
struct PcIC0 : Pc {
    std::vector<uint32_t> Lp, Lc, Tp, Tc;
    vec Lv, Tv, diag;
    explicit PcIC0(const Csr& A) {
        if (A.n != A.m) throw std::invalid_argument("IC(0) needs a square matrix");
        const uint32_t n = A.n;
        Lp.assign(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
                if (A.ci[p] <= i) Lp[i + 1]++;
        for (uint32_t i = 0; i < n; ++i) Lp[i + 1] += Lp[i];
        Lc.resize(Lp[n]);
        Lv.resize(Lp[n]);
        {
            std::vector<uint32_t> off(Lp.begin(), Lp.end() - 1);
            for (uint32_t i = 0; i < n; ++i)
                for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
                    if (A.ci[p] <= i) {
                        uint32_t d = off[i]++;
                        Lc[d] = A.ci[p];
                        Lv[d] = A.v[p];
                    }
        }
        diag.assign(n, 0.0);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t rs = Lp[i], re = Lp[i + 1];
            for (uint32_t p = rs; p < re; ++p) {
                const uint32_t j = Lc[p];
                if (j >= i) break;
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
            uint32_t dp = re;
            for (uint32_t p = rs; p < re; ++p)
                if (Lc[p] == i) { dp = p; break; }
            if (dp == re) throw std::runtime_error("IC(0): missing diagonal entry");
            double sd = 0.0;
            for (uint32_t p = rs; p < dp; ++p) sd += Lv[p] * Lv[p];
            const double dv = Lv[dp] - sd;
            if (dv <= 1e-14) throw std::runtime_error("IC(0): matrix may not be SPD (non-positive pivot)");
            Lv[dp] = std::sqrt(dv);
            diag[i] = Lv[dp];
        }
        Tp.assign(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t p = Lp[i]; p < Lp[i + 1]; ++p)
                if (Lc[p] < i) Tp[Lc[p] + 1]++;
        for (uint32_t i = 0; i < n; ++i) Tp[i + 1] += Tp[i];
        Tc.resize(Tp[n]);
        Tv.resize(Tp[n]);
        std::vector<uint32_t> off(Tp.begin(), Tp.end() - 1);
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t p = Lp[i]; p < Lp[i + 1]; ++p)
                if (Lc[p] < i) {
                    uint32_t d = off[Lc[p]]++;
                    Tc[d] = i;
                    Tv[d] = Lv[p];
                }
    }
    void apply(const vec& r, vec& z) override {
        const uint32_t n = (uint32_t)diag.size();
        vec y(n, 0.0);
        z.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            double s = r[i];
            for (uint32_t p = Lp[i]; p < Lp[i + 1]; ++p) {
                const uint32_t j = Lc[p];
                if (j >= i) break;
                s -= Lv[p] * y[j];
            }
            y[i] = s / diag[i];
        }
        for (int64_t ii = (int64_t)n - 1; ii >= 0; --ii) {
            const uint32_t i = (uint32_t)ii;
            double s = y[i];
            for (uint32_t p = Tp[i]; p < Tp[i + 1]; ++p) s -= Tv[p] * z[Tc[p]];
            z[i] = s / diag[i];
        }
    }
};

// make_ilu0_preconditioner (array.cpp:2733-2934): "upper-only" IKJ variant;
// L strictly lower (unit diagonal implied), U upper with diagonal.
struct PcILU0 : Pc {
    std::vector<uint32_t> Lp, Lc, Up, Uc, dU;
    vec Lv, Uv;
    explicit PcILU0(const Csr& A) {
        if (A.n != A.m) throw std::invalid_argument("ILU(0) needs a square matrix");
        const uint32_t n = A.n;
        Lp.assign(n + 1, 0);
        Up.assign(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
                (A.ci[p] < i ? Lp : Up)[i + 1]++;
        for (uint32_t i = 0; i < n; ++i) { Lp[i + 1] += Lp[i]; Up[i + 1] += Up[i]; }
        Lc.resize(Lp[n]); Lv.resize(Lp[n]);
        Uc.resize(Up[n]); Uv.resize(Up[n]);
        dU.assign(n, 0);
        std::vector<char> has(n, 0);
        std::vector<uint32_t> oL(Lp.begin(), Lp.end() - 1), oU(Up.begin(), Up.end() - 1);
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
                const uint32_t j = A.ci[p];
                if (j < i) { uint32_t d = oL[i]++; Lc[d] = j; Lv[d] = A.v[p]; }
                else {
                    uint32_t d = oU[i]++;
                    Uc[d] = j; Uv[d] = A.v[p];
                    if (j == i) { dU[i] = d; has[i] = 1; }
                }
            }
        for (uint32_t i = 0; i < n; ++i)
            if (!has[i]) throw std::runtime_error("ILU(0): missing diagonal entry");
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t us = Up[i], ue = Up[i + 1];
            for (uint32_t pl = Lp[i]; pl < Lp[i + 1]; ++pl) {
                const uint32_t k = Lc[pl];
                const double ukk = Uv[dU[k]];
                if (std::fabs(ukk) < 1e-30) throw std::runtime_error("ILU(0) breakdown: tiny pivot U_kk");
                Lv[pl] /= ukk;
                uint32_t pi = us, pk = Up[k];
                const uint32_t ek = Up[k + 1];
                while (pi < ue && pk < ek) {
                    const uint32_t ji = Uc[pi], jk = Uc[pk];
                    if (ji == jk) { Uv[pi] -= Lv[pl] * Uv[pk]; ++pi; ++pk; }
                    else if (ji < jk) ++pi;
                    else ++pk;
                }
            }
        }
    }
    void apply(const vec& r, vec& z) override {
        const uint32_t n = (uint32_t)dU.size();
        vec y(n, 0.0);
        z.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            double s = r[i];
            for (uint32_t p = Lp[i]; p < Lp[i + 1]; ++p) s -= Lv[p] * y[Lc[p]];
            y[i] = s;
        }
        for (int64_t ii = (int64_t)n - 1; ii >= 0; --ii) {
            const uint32_t i = (uint32_t)ii;
            double s = y[i];
            for (uint32_t p = Up[i]; p < Up[i + 1]; ++p) {
                if (Uc[p] == i) continue;
                s -= Uv[p] * z[Uc[p]];
            }
            z[i] = s / Uv[dU[i]];
        }
    }
};

// make_redblack_data (array.cpp:2947-3067): colouring of pattern(A+A^T) by BFS
// 2-colouring with greedy multicolour fallback, stable renumbering by colour,
// A2 = P A P^T through the COO constructor.
struct RBData {
    Csr A2;
    std::vector<uint32_t> color_ptr, new_id, perm;
    vec diag2;
    uint32_t ncolor = 0;
};

static std::shared_ptr<RBData> make_redblack(const Csr& A) {
    const uint32_t n = A.n, nnz = A.nnz();
    auto D = std::make_shared<RBData>();
    std::vector<uint32_t> tp(n + 1, 0), tr(nnz);
    for (uint32_t p = 0; p < nnz; ++p) tp[A.ci[p] + 1]++;
    for (uint32_t i = 0; i < n; ++i) tp[i + 1] += tp[i];
    {
        std::vector<uint32_t> off(tp.begin(), tp.end() - 1);
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) tr[off[A.ci[p]]++] = i;
    }
    std::vector<int> color(n, -1);
    bool bip = true;
    std::vector<uint32_t> q;
    for (uint32_t seed = 0; seed < n && bip; ++seed) {
        if (color[seed] != -1) continue;
        color[seed] = 0;
        q.clear();
        q.push_back(seed);
        for (size_t h = 0; h < q.size() && bip; ++h) {
            const uint32_t i = q[h];
            auto visit = [&](uint32_t j) {
                if (j == i) return;
                if (color[j] == -1) { color[j] = 1 - color[i]; q.push_back(j); }
                else if (color[j] == color[i]) bip = false;
            };
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) visit(A.ci[p]);
            for (uint32_t p = tp[i]; p < tp[i + 1]; ++p) visit(tr[p]);
        }
    }
    if (!bip) {
        std::fill(color.begin(), color.end(), -1);
        std::vector<char> used(n + 1, 0);
        std::vector<int> touched;
        for (uint32_t i = 0; i < n; ++i) {
            touched.clear();
            auto mark = [&](uint32_t j) {
                if (j == i) return;
                const int c = color[j];
                if (c >= 0 && !used[c]) { used[c] = 1; touched.push_back(c); }
            };
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) mark(A.ci[p]);
            for (uint32_t p = tp[i]; p < tp[i + 1]; ++p) mark(tr[p]);
            int c = 0;
            while (used[c]) ++c;
            color[i] = c;
            for (int c2 : touched) used[c2] = 0;
        }
    }
    for (uint32_t i = 0; i < n; ++i)
        if ((uint32_t)color[i] + 1 > D->ncolor) D->ncolor = color[i] + 1;
    D->color_ptr.assign(D->ncolor + 1, 0);
    for (uint32_t i = 0; i < n; ++i) D->color_ptr[color[i] + 1]++;
    for (uint32_t c = 0; c < D->ncolor; ++c) D->color_ptr[c + 1] += D->color_ptr[c];
    D->new_id.resize(n);
    {
        std::vector<uint32_t> off(D->color_ptr.begin(), D->color_ptr.end() - 1);
        for (uint32_t i = 0; i < n; ++i) D->new_id[i] = off[color[i]]++;
    }
    D->perm.resize(n);
    for (uint32_t i = 0; i < n; ++i) D->perm[D->new_id[i]] = i;
    vec v2;
    std::vector<uint32_t> r2, c2;
    v2.reserve(nnz); r2.reserve(nnz); c2.reserve(nnz);
    for (uint32_t ni = 0; ni < n; ++ni) {
        const uint32_t i = D->perm[ni];
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            r2.push_back(ni);
            c2.push_back(D->new_id[A.ci[p]]);
            v2.push_back(A.v[p]);
        }
    }
    D->A2 = csr_from_coo(n, n, v2, r2, c2);
    D->diag2.assign(n, 0.0);
    for (uint32_t ni = 0; ni < n; ++ni) {
        bool found = false;
        for (uint32_t p = D->A2.rp[ni]; p < D->A2.rp[ni + 1]; ++p)
            if (D->A2.ci[p] == ni) { D->diag2[ni] = D->A2.v[p]; found = true; break; }
        if (!found || std::fabs(D->diag2[ni]) < 1e-300)
            throw std::runtime_error("SSOR/red-black: missing nonzero diagonal entry");
    }
    return D;
}

// make_ssor_preconditioner (array.cpp:3068-3164): forward (D+wL)w = w(2-w)f
// colour by colour, backward (D+wU)z2 = Dw in reverse colour order, permute back.
struct PcSSOR : Pc {
    std::shared_ptr<RBData> D;
    double omega, fw;
    vec f, w, z2;
    PcSSOR(const Csr& A, double om) : omega(om), fw(om * (2.0 - om)) {
        if (A.n != A.m) throw std::invalid_argument("SSOR needs a square matrix");
        if (!(om > 0.0 && om < 2.0)) throw std::invalid_argument("SSOR: relaxation factor must satisfy 0 < omega < 2");
        D = make_redblack(A);
        f.resize(A.n); w.resize(A.n); z2.resize(A.n);
    }
    void apply(const vec& r, vec& z) override {
        const uint32_t n = (uint32_t)D->diag2.size();
        const Csr& B = D->A2;
        for (uint32_t i = 0; i < n; ++i) f[i] = r[D->perm[i]];
        for (uint32_t c = 0; c < D->ncolor; ++c) {
            const int64_t lo = D->color_ptr[c], hi = D->color_ptr[c + 1];
#pragma omp parallel for schedule(static) if (hi - lo > 1000)
            for (int64_t ii = lo; ii < hi; ++ii) {
                double s = fw * f[ii];
                for (uint32_t p = B.rp[ii]; p < B.rp[ii + 1]; ++p)
                    if (B.ci[p] < (uint32_t)ii) s -= omega * B.v[p] * w[B.ci[p]];
                w[ii] = s / D->diag2[ii];
            }
        }
        for (int64_t c = (int64_t)D->ncolor - 1; c >= 0; --c) {
            const int64_t lo = D->color_ptr[c], hi = D->color_ptr[c + 1];
#pragma omp parallel for schedule(static) if (hi - lo > 1000)
            for (int64_t ii = lo; ii < hi; ++ii) {
                double s = 0.0;
                for (uint32_t p = B.rp[ii]; p < B.rp[ii + 1]; ++p)
                    if (B.ci[p] > (uint32_t)ii) s += B.v[p] * z2[B.ci[p]];
                z2[ii] = w[ii] - (omega / D->diag2[ii]) * s;
            }
        }
        z.resize(n);
        for (uint32_t i = 0; i < n; ++i) z[i] = z2[D->new_id[i]];
    }
};

// dense_lu_factor / dense_lu_solve (array.cpp:2252-2333): partial pivoting,
// pivots below 1e-14 in magnitude are replaced by +1e-14.
static void dense_lu_factor(const Csr& A, vec& mat, std::vector<uint32_t>& piv) {
    const uint32_t n = A.n;
    mat.assign((size_t)n * n, 0.0);
    piv.resize(n);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) mat[(size_t)i * n + A.ci[p]] = A.v[p];
    for (uint32_t i = 0; i < n; ++i) piv[i] = i;
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t pm = k;
        double vm = std::fabs(mat[(size_t)k * n + k]);
        for (uint32_t i = k + 1; i < n; ++i)
            if (std::fabs(mat[(size_t)i * n + k]) > vm) { vm = std::fabs(mat[(size_t)i * n + k]); pm = i; }
        if (pm != k) {
            std::swap(piv[k], piv[pm]);
            for (uint32_t j = 0; j < n; ++j) std::swap(mat[(size_t)k * n + j], mat[(size_t)pm * n + j]);
        }
        if (std::fabs(mat[(size_t)k * n + k]) < 1e-14) mat[(size_t)k * n + k] = 1e-14;
        for (uint32_t i = k + 1; i < n; ++i) {
            mat[(size_t)i * n + k] /= mat[(size_t)k * n + k];
            for (uint32_t j = k + 1; j < n; ++j) mat[(size_t)i * n + j] -= mat[(size_t)i * n + k] * mat[(size_t)k * n + j];
        }
    }
}

static void dense_lu_solve(const vec& mat, const std::vector<uint32_t>& piv, const vec& b, vec& x) {
    const uint32_t n = (uint32_t)piv.size();
    x.resize(n);
    for (uint32_t i = 0; i < n; ++i) x[i] = b[piv[i]];
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < i; ++j) x[i] -= mat[(size_t)i * n + j] * x[j];
    for (int64_t i = (int64_t)n - 1; i >= 0; --i) {
        for (uint32_t j = i + 1; j < n; ++j) x[i] -= mat[(size_t)i * n + j] * x[j];
        x[i] /= mat[(size_t)i * n + i];
    }
}

// C/F AMG (array.cpp:3165-3398): strength theta=0.25, priority-queue C/F
// splitting, direct interpolation, Galerkin A_c = P^T (A P), dense LU at the
// coarsest level, damped-Jacobi V-cycle (omega=2/3, nu1=nu2=1).
struct AmgLevel {
    Csr A, P, R;
    vec diag;
};

struct PcAmgCF : Pc {
    std::vector<AmgLevel> lv;
    std::vector<std::vector<char>> split;  // per non-coarsest level: 1 C, 2 F
    vec lu;
    std::vector<uint32_t> piv;
    double omega = 2.0 / 3.0;
    int nu1 = 1, nu2 = 1;

    PcAmgCF(const Csr& A0, double theta, uint32_t coarse_thresh, int max_levels) {
        if (A0.n != A0.m) throw std::invalid_argument("AMG needs a square matrix");
        lv.emplace_back();
        lv[0].A = A0;
        for (int k = 0;; ++k) {
            const Csr& A = lv[k].A;
            const uint32_t n = A.n;
            lv[k].diag.assign(n, 0.0);
            for (uint32_t i = 0; i < n; ++i)
                for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
                    if (A.ci[p] == i) { lv[k].diag[i] = A.v[p]; break; }
            if (n <= coarse_thresh || k + 1 >= max_levels) break;
            std::vector<uint32_t> Sp(n + 1, 0), Sc;
            for (uint32_t i = 0; i < n; ++i) {
                double mx = 0.0;
                for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
                    if (A.ci[p] != i) mx = std::max(mx, std::fabs(A.v[p]));
                const double thr = theta * mx;
                for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
                    if (A.ci[p] != i && std::fabs(A.v[p]) >= thr) Sc.push_back(A.ci[p]);
                Sp[i + 1] = (uint32_t)Sc.size();
            }
            std::vector<char> st(n, 0);  // 0 undecided, 1 C, 2 F
            std::vector<uint32_t> w(n);
            for (uint32_t i = 0; i < n; ++i) w[i] = Sp[i + 1] - Sp[i];
            std::priority_queue<std::pair<uint32_t, uint32_t>> pq;
            for (uint32_t i = 0; i < n; ++i) pq.push({w[i], i});
            while (!pq.empty()) {
                const uint32_t i = pq.top().second;
                pq.pop();
                if (st[i] != 0) continue;
                st[i] = 1;
                for (uint32_t p = Sp[i]; p < Sp[i + 1]; ++p) {
                    const uint32_t j = Sc[p];
                    if (st[j] != 0) continue;
                    st[j] = 2;
                    for (uint32_t q = Sp[j]; q < Sp[j + 1]; ++q) {
                        const uint32_t kk = Sc[q];
                        if (st[kk] == 0) { w[kk]++; pq.push({w[kk], kk}); }
                    }
                }
            }
            for (uint32_t i = 0; i < n; ++i)
                if (st[i] == 0) st[i] = 1;
            for (uint32_t i = 0; i < n; ++i) {
                if (st[i] != 2) continue;
                bool hasC = false;
                for (uint32_t p = Sp[i]; p < Sp[i + 1]; ++p)
                    if (st[Sc[p]] == 1) { hasC = true; break; }
                if (!hasC) st[i] = 1;
            }
            std::vector<uint32_t> mapC(n, 0);
            uint32_t nc = 0;
            for (uint32_t i = 0; i < n; ++i) { mapC[i] = nc; if (st[i] == 1) nc++; }
            if (nc == 0 || nc == n) break;
            split.push_back(st);
            Csr P;
            P.n = n;
            P.m = nc;
            P.rp.assign(n + 1, 0);
            for (uint32_t i = 0; i < n; ++i) {
                if (st[i] == 1) { P.ci.push_back(mapC[i]); P.v.push_back(1.0); }
                else {
                    double den = lv[k].diag[i];
                    for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
                        const uint32_t j = A.ci[p];
                        if (j == i) continue;
                        bool strong = false;
                        for (uint32_t s = Sp[i]; s < Sp[i + 1]; ++s)
                            if (Sc[s] == j) { strong = true; break; }
                        if (!strong) den += A.v[p];
                    }
                    if (std::fabs(den) < 1e-300) den = lv[k].diag[i];
                    for (uint32_t s = Sp[i]; s < Sp[i + 1]; ++s) {
                        const uint32_t j = Sc[s];
                        if (st[j] != 1) continue;
                        double aij = 0.0;
                        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
                            if (A.ci[p] == j) { aij = A.v[p]; break; }
                        P.ci.push_back(mapC[j]);
                        P.v.push_back(-aij / den);
                    }
                }
                P.rp[i + 1] = (uint32_t)P.ci.size();
            }
            Csr AP = spgemm(A, P);
            Csr R = transpose(P);
            Csr Ac = spgemm(R, AP);
            lv[k].P = std::move(P);
            lv[k].R = std::move(R);
            lv.emplace_back();
            lv[k + 1].A = std::move(Ac);
        }
        dense_lu_factor(lv.back().A, lu, piv);
    }

    void jacobi(const AmgLevel& L, const vec& f, vec& u, int nu) {
        const uint32_t n = L.A.n;
        vec Ax, t(n);
        for (int s = 0; s < nu; ++s) {
            spmv(L.A, u, Ax);
            for (uint32_t i = 0; i < n; ++i) t[i] = u[i] + omega * (f[i] - Ax[i]) / L.diag[i];
            u = t;
        }
    }
    void vcycle(size_t k, const vec& f, vec& u) {
        const AmgLevel& L = lv[k];
        if (k + 1 == lv.size()) { dense_lu_solve(lu, piv, f, u); return; }
        const uint32_t n = L.A.n;
        jacobi(L, f, u, nu1);
        vec Ax, rr(n), rc, uc, pc;
        spmv(L.A, u, Ax);
        for (uint32_t i = 0; i < n; ++i) rr[i] = f[i] - Ax[i];
        spmv(L.R, rr, rc);
        uc.assign(L.P.m, 0.0);
        vcycle(k + 1, rc, uc);
        spmv(L.P, uc, pc);
        for (uint32_t i = 0; i < n; ++i) u[i] += pc[i];
        jacobi(L, f, u, nu2);
    }
    void apply(const vec& r, vec& z) override {
        z.assign(r.size(), 0.0);
        vcycle(0, r, z);
    }
};

// ---------------------------------------------------------------------------
// Block (BSR) path (block.hpp)
// ---------------------------------------------------------------------------
template <int BS>
struct Bsr {
    static constexpr int BV = BS * BS;
    uint32_t nbr = 0, nbc = 0;
    std::vector<uint32_t> rp{0}, ci, dptr;
    vec v;
    uint32_t nnzb() const { return (uint32_t)ci.size(); }
    const double* blk(uint32_t p) const { return &v[(size_t)p * BV]; }
    // cache_diag (block.hpp:199-205)
    void cache_diag() {
        dptr.assign(nbr, UINT32_MAX);
        for (uint32_t b = 0; b < nbr; ++b)
            for (uint32_t p = rp[b]; p < rp[b + 1]; ++p)
                if (ci[p] == b) { dptr[b] = p; break; }
    }
    // matvec (block.hpp:208-234): per block row, per block, per row r the
    // inner product over c is formed first, then added to y.
    void matvec(const vec& x, vec& y) const {
        y.assign((size_t)nbr * BS, 0.0);
        const int64_t nb = nbr;
#pragma omp parallel for schedule(static)
        for (int64_t b = 0; b < nb; ++b) {
            double* yi = &y[(size_t)b * BS];
            for (uint32_t p = rp[b]; p < rp[b + 1]; ++p) {
                const double* B = blk(p);
                const double* xj = &x[(size_t)ci[p] * BS];
                for (int r = 0; r < BS; ++r) {
                    double acc = 0.0;
                    for (int c = 0; c < BS; ++c) acc += B[r * BS + c] * xj[c];
                    yi[r] += acc;
                }
            }
        }
    }
    // transpose (block.hpp:237-267)
    Bsr transpose() const {
        Bsr T;
        T.nbr = nbc;
        T.nbc = nbr;
        T.rp.assign(nbc + 1, 0);
        for (uint32_t p = 0; p < nnzb(); ++p) T.rp[ci[p] + 1]++;
        for (uint32_t j = 0; j < nbc; ++j) T.rp[j + 1] += T.rp[j];
        T.ci.resize(nnzb());
        T.v.resize(v.size());
        std::vector<uint32_t> off(T.rp.begin(), T.rp.end() - 1);
        for (uint32_t i = 0; i < nbr; ++i)
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
                const uint32_t q = off[ci[p]]++;
                T.ci[q] = i;
                const double* B = blk(p);
                double* C = &T.v[(size_t)q * BV];
                for (int r = 0; r < BS; ++r)
                    for (int c = 0; c < BS; ++c) C[r * BS + c] = B[c * BS + r];
            }
        T.cache_diag();
        return T;
    }
};

// condense (block.hpp:269-315): block column = j/BS, duplicates summed in
// scalar CSR order, block columns per block row sorted and unique.
template <int BS>
static Bsr<BS> condense(const Csr& A, uint32_t nbr, uint32_t nbc) {
    Bsr<BS> B;
    B.nbr = nbr;
    B.nbc = nbc;
    B.rp.assign(nbr + 1, 0);
    std::vector<std::vector<uint32_t>> bc(nbr);
    for (uint32_t b = 0; b < nbr; ++b) {
        auto& o = bc[b];
        for (uint32_t i = b * BS; i < (b + 1) * BS; ++i)
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) o.push_back(A.ci[p] / BS);
        std::sort(o.begin(), o.end());
        o.erase(std::unique(o.begin(), o.end()), o.end());
        B.rp[b + 1] = B.rp[b] + (uint32_t)o.size();
    }
    B.ci.resize(B.rp[nbr]);
    B.v.assign((size_t)B.rp[nbr] * Bsr<BS>::BV, 0.0);
    for (uint32_t b = 0; b < nbr; ++b) {
        for (size_t k = 0; k < bc[b].size(); ++k) B.ci[B.rp[b] + k] = bc[b][k];
        for (uint32_t i = b * BS; i < (b + 1) * BS; ++i) {
            const int r = i % BS;
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
                const uint32_t bj = A.ci[p] / BS, c = A.ci[p] % BS;
                auto it = std::lower_bound(bc[b].begin(), bc[b].end(), bj);
                const uint32_t off = B.rp[b] + (uint32_t)(it - bc[b].begin());
                B.v[(size_t)off * Bsr<BS>::BV + (size_t)r * BS + c] += A.v[p];
            }
        }
    }
    B.cache_diag();
    return B;
}

// blk_invert (block.hpp:70-100): Gauss elimination with partial pivoting on
// the augmented [M | I], then back substitution; fails on a pivot < 1e-300.
template <int BS>
static bool blk_invert(const double* M, double* Mi) {
    double a[BS][2 * BS];
    for (int i = 0; i < BS; ++i)
        for (int j = 0; j < BS; ++j) { a[i][j] = M[i * BS + j]; a[i][BS + j] = (i == j) ? 1.0 : 0.0; }
    for (int k = 0; k < BS; ++k) {
        int pv = k;
        for (int i = k + 1; i < BS; ++i)
            if (std::fabs(a[i][k]) > std::fabs(a[pv][k])) pv = i;
        if (std::fabs(a[pv][k]) < 1e-300) return false;
        if (pv != k)
            for (int j = 0; j < 2 * BS; ++j) std::swap(a[k][j], a[pv][j]);
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

// C -= A*B (block.hpp:37-46)
template <int BS>
static void blk_mm_sub(const double* A, const double* B, double* C) {
    for (int r = 0; r < BS; ++r)
        for (int c = 0; c < BS; ++c) {
            double s = 0.0;
            for (int k = 0; k < BS; ++k) s += A[r * BS + k] * B[k * BS + c];
            C[r * BS + c] -= s;
        }
}
// C := C*B (block.hpp:48-59)
template <int BS>
static void blk_right_mul(double* C, const double* B) {
    double t[BS * BS];
    for (int r = 0; r < BS; ++r)
        for (int c = 0; c < BS; ++c) {
            double s = 0.0;
            for (int k = 0; k < BS; ++k) s += C[r * BS + k] * B[k * BS + c];
            t[r * BS + c] = s;
        }
    for (int i = 0; i < BS * BS; ++i) C[i] = t[i];
}
// y = M x (block.hpp:61-69)
template <int BS>
static void blk_mv(const double* M, const double* x, double* y) {
    for (int r = 0; r < BS; ++r) {
        double s = 0.0;
        for (int c = 0; c < BS; ++c) s += M[r * BS + c] * x[c];
        y[r] = s;
    }
}
template <int BS>
static uint32_t blk_find(const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci, uint32_t row, uint32_t col) {
    uint32_t lo = rp[row], hi = rp[row + 1];
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (ci[mid] == col) return mid;
        if (ci[mid] < col) lo = mid + 1; else hi = mid;
    }
    return UINT32_MAX;
}

// make_block_ilu0_preconditioner (block.hpp:454-623): full block IKJ ILU(0)
// with left sum L_IK -= sum_J L_IJ U_JK, right multiply by inv(U_KK),
// two-pointer update of U-row I, cached inverse diagonal blocks.
template <int BS>
struct PcBlockILU0 : Pc {
    static constexpr int BV = BS * BS;
    uint32_t nb;
    std::vector<uint32_t> Lp, Lc, Up, Uc, Ud;
    vec Lv, Uv, inv;
    explicit PcBlockILU0(const Bsr<BS>& A) {
        if (A.nbr != A.nbc) throw std::invalid_argument("Block-ILU(0) needs a square block matrix");
        nb = A.nbr;
        Lp.assign(nb + 1, 0);
        Up.assign(nb + 1, 0);
        for (uint32_t I = 0; I < nb; ++I)
            for (uint32_t p = A.rp[I]; p < A.rp[I + 1]; ++p) (A.ci[p] < I ? Lp : Up)[I + 1]++;
        for (uint32_t I = 0; I < nb; ++I) { Lp[I + 1] += Lp[I]; Up[I + 1] += Up[I]; }
        Lc.resize(Lp[nb]); Lv.resize((size_t)Lp[nb] * BV);
        Uc.resize(Up[nb]); Uv.resize((size_t)Up[nb] * BV);
        Ud.assign(nb, UINT32_MAX);
        std::vector<uint32_t> oL(Lp.begin(), Lp.end() - 1), oU(Up.begin(), Up.end() - 1);
        for (uint32_t I = 0; I < nb; ++I)
            for (uint32_t p = A.rp[I]; p < A.rp[I + 1]; ++p) {
                const uint32_t J = A.ci[p];
                const double* s = A.blk(p);
                if (J < I) { uint32_t d = oL[I]++; Lc[d] = J; std::copy(s, s + BV, &Lv[(size_t)d * BV]); }
                else {
                    uint32_t d = oU[I]++;
                    Uc[d] = J;
                    std::copy(s, s + BV, &Uv[(size_t)d * BV]);
                    if (J == I) Ud[I] = d;
                }
            }
        for (uint32_t I = 0; I < nb; ++I)
            if (Ud[I] == UINT32_MAX) throw std::runtime_error("Block-ILU(0): block row " + std::to_string(I) + " has no diagonal block");
        inv.resize((size_t)nb * BV);
        for (uint32_t I = 0; I < nb; ++I) {
            const uint32_t Ls = Lp[I], Le = Lp[I + 1], Us = Up[I], Ue = Up[I + 1];
            for (uint32_t pL = Ls; pL < Le; ++pL) {
                const uint32_t K = Lc[pL];
                double* LIK = &Lv[(size_t)pL * BV];
                for (uint32_t q = Ls; q < pL; ++q) {
                    const uint32_t J = Lc[q];
                    const uint32_t pJK = blk_find<BS>(Up, Uc, J, K);
                    if (pJK != UINT32_MAX) blk_mm_sub<BS>(&Lv[(size_t)q * BV], &Uv[(size_t)pJK * BV], LIK);
                }
                blk_right_mul<BS>(LIK, &inv[(size_t)K * BV]);
                uint32_t pi = Us, pk = Up[K];
                const uint32_t ek = Up[K + 1];
                while (pi < Ue && pk < ek) {
                    if (Uc[pi] == Uc[pk]) { blk_mm_sub<BS>(LIK, &Uv[(size_t)pk * BV], &Uv[(size_t)pi * BV]); ++pi; ++pk; }
                    else if (Uc[pi] < Uc[pk]) ++pi;
                    else ++pk;
                }
            }
            if (!blk_invert<BS>(&Uv[(size_t)Ud[I] * BV], &inv[(size_t)I * BV]))
                throw std::runtime_error("Block-ILU(0) breakdown: singular diagonal block U_" + std::to_string(I));
        }
    }
    void apply(const vec& rhs, vec& z) override {
        vec y((size_t)nb * BS);
        z.resize((size_t)nb * BS);
        for (uint32_t I = 0; I < nb; ++I) {
            double* yI = &y[(size_t)I * BS];
            for (int c = 0; c < BS; ++c) yI[c] = rhs[(size_t)I * BS + c];
            for (uint32_t p = Lp[I]; p < Lp[I + 1]; ++p) {
                const double* L = &Lv[(size_t)p * BV];
                const double* yJ = &y[(size_t)Lc[p] * BS];
                for (int r = 0; r < BS; ++r) {
                    double s = 0.0;
                    for (int c = 0; c < BS; ++c) s += L[r * BS + c] * yJ[c];
                    yI[r] -= s;
                }
            }
        }
        for (uint32_t II = nb; II-- > 0;) {
            double t[BS];
            for (int c = 0; c < BS; ++c) t[c] = y[(size_t)II * BS + c];
            for (uint32_t p = Up[II]; p < Up[II + 1]; ++p) {
                const uint32_t J = Uc[p];
                if (J == II) continue;
                const double* U = &Uv[(size_t)p * BV];
                const double* zJ = &z[(size_t)J * BS];
                for (int r = 0; r < BS; ++r) {
                    double s = 0.0;
                    for (int c = 0; c < BS; ++c) s += U[r * BS + c] * zJ[c];
                    t[r] -= s;
                }
            }
            blk_mv<BS>(&inv[(size_t)II * BV], t, &z[(size_t)II * BS]);
        }
    }
};

// Block-Jacobi with cached inverses (test_BlockMatrix_mpi.cpp:132-155); the
// "blockjacobi" lambda of BlockMatrix::gmres (block.hpp:714-735) computes the
// same inverse on every apply.
template <int BS>
struct PcBlockJacobi : Pc {
    static constexpr int BV = BS * BS;
    uint32_t nb;
    vec inv;
    explicit PcBlockJacobi(const Bsr<BS>& A) : nb(A.nbr) {
        inv.resize((size_t)nb * BV);
        for (uint32_t b = 0; b < nb; ++b) {
            if (A.dptr[b] == UINT32_MAX) throw std::runtime_error("block-Jacobi: missing diagonal block");
            if (!blk_invert<BS>(A.blk(A.dptr[b]), &inv[(size_t)b * BV]))
                throw std::runtime_error("block-Jacobi: singular diagonal block");
        }
    }
    void apply(const vec& r, vec& z) override {
        z.resize(r.size());
        const int64_t n = nb;
#pragma omp parallel for schedule(static)
        for (int64_t b = 0; b < n; ++b) blk_mv<BS>(&inv[(size_t)b * BV], &r[(size_t)b * BS], &z[(size_t)b * BS]);
    }
};

// blk_matmul (block.hpp:883-932): per row, accumulate block products in
// traversal order, columns sorted ascending.
template <int BS>
static Bsr<BS> blk_matmul(const Bsr<BS>& A, const Bsr<BS>& B) {
    constexpr int BV = BS * BS;
    Bsr<BS> C;
    C.nbr = A.nbr;
    C.nbc = B.nbc;
    C.rp.assign(A.nbr + 1, 0);
    vec acc((size_t)B.nbc * BV);
    std::vector<int64_t> mark(B.nbc, -1);
    std::vector<uint32_t> hl;
    for (uint32_t i = 0; i < A.nbr; ++i) {
        hl.clear();
        for (uint32_t pa = A.rp[i]; pa < A.rp[i + 1]; ++pa) {
            const uint32_t k = A.ci[pa];
            const double* ab = A.blk(pa);
            for (uint32_t pb = B.rp[k]; pb < B.rp[k + 1]; ++pb) {
                const uint32_t j = B.ci[pb];
                const double* bb = B.blk(pb);
                if (mark[j] != (int64_t)i) {
                    mark[j] = i;
                    hl.push_back(j);
                    std::fill(acc.begin() + (size_t)j * BV, acc.begin() + (size_t)(j + 1) * BV, 0.0);
                }
                double* cb = &acc[(size_t)j * BV];
                for (int r = 0; r < BS; ++r)
                    for (int c = 0; c < BS; ++c) {
                        double s = 0.0;
                        for (int mm = 0; mm < BS; ++mm) s += ab[r * BS + mm] * bb[mm * BS + c];
                        cb[r * BS + c] += s;
                    }
            }
        }
        std::sort(hl.begin(), hl.end());
        for (uint32_t j : hl) {
            C.ci.push_back(j);
            for (int t = 0; t < BV; ++t) C.v.push_back(acc[(size_t)j * BV + t]);
        }
        C.rp[i + 1] = (uint32_t)C.ci.size();
    }
    C.cache_diag();
    return C;
}

// Robust-mode aggregation (no counterpart in the reference): aggregates
// grown around a distance-2 maximal independent set of the strength graph,
// following Bell, Dalton & Olson, SIAM J. Sci. Comput. 34(4), 2012.
//   strong(i,j), j != i:  |a_ij| >= theta * sqrt(m_i * m_j), a_ij != 0,
//                         m_i = max_{k != i} |a_ik|   (block: Frobenius norm)
//   priority of node i:   h(i) = lowbias32(i) (C. Wellons' integer hash),
//                         key = state<<62 | (h & 0x3fffffff)<<32 | i with
//                         state 2 root, 1 undecided, 0 out
//   each round: two max-propagations of the keys over the strong neighbours;
//   an undecided node whose two-hop max is its own key becomes a root, one
//   whose two-hop max has state 2 goes out; repeat until nothing undecided.
//   Roots are numbered in index order; a non-root joins the strong root
//   neighbour with the largest key; a node still free joins the neighbour
//   (assigned in that first pass) with the largest key; leftovers become
//   singletons in index order. Written out with per-node lists, one node at
//   a time, as a restatement independent of the product's array sweeps.
static uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

template <class Norm>
static int mis2_agg(uint32_t n, const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci, Norm nrm,
                    double theta, std::vector<int>& agg, std::vector<int>* roots = nullptr) {
    std::vector<double> m(n, 0.0);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
            if (ci[p] != i) m[i] = std::max(m[i], nrm(p));
    std::vector<std::vector<uint32_t>> S(n);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
            const uint32_t j = ci[p];
            const double a = nrm(p);
            if (j != i && a != 0.0 && a >= theta * std::sqrt(m[i] * m[j])) S[i].push_back(j);
        }
    auto key = [](int state, uint32_t i) {
        return ((uint64_t)state << 62) | ((uint64_t)(lowbias32(i) & 0x3fffffffU) << 32) | (uint64_t)i;
    };
    std::vector<int> state(n, 1);
    while (true) {
        std::vector<uint64_t> one(n), two(n);
        for (uint32_t i = 0; i < n; ++i) {
            one[i] = key(state[i], i);
            for (uint32_t j : S[i]) one[i] = std::max(one[i], key(state[j], j));
        }
        for (uint32_t i = 0; i < n; ++i) {
            two[i] = one[i];
            for (uint32_t j : S[i]) two[i] = std::max(two[i], one[j]);
        }
        int undecided = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (state[i] != 1) continue;
            if ((two[i] & 0xffffffffULL) == i) state[i] = 2;
            else if ((two[i] >> 62) == 2) state[i] = 0;
            else ++undecided;
        }
        if (!undecided) break;
    }
    agg.assign(n, -1);
    int na = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (state[i] == 2) agg[i] = na++;
    if (roots) roots->assign(state.begin(), state.end());
    std::vector<int> first = agg;
    for (uint32_t i = 0; i < n; ++i) {
        if (state[i] == 2) continue;
        bool found = false;
        uint64_t best = 0;
        for (uint32_t j : S[i])
            if (state[j] == 2 && (!found || key(0, j) > best)) { found = true; best = key(0, j); first[i] = agg[j]; }
    }
    agg = first;
    for (uint32_t i = 0; i < n; ++i) {
        if (first[i] >= 0) continue;
        bool found = false;
        uint64_t best = 0;
        for (uint32_t j : S[i])
            if (first[j] >= 0 && (!found || key(0, j) > best)) { found = true; best = key(0, j); agg[i] = first[j]; }
    }
    for (uint32_t i = 0; i < n; ++i)
        if (agg[i] < 0) agg[i] = na++;
    return na;
}

// Block SA-AMG (block.hpp:1051-1179) with the smoother made configurable:
// reference defaults theta=0.25, omega=2/3, one block-Jacobi sweep pre and
// post, aggregates of at most 4 block rows, coarsest level <= 50 block rows
// solved by Gauss elimination with pivot floor +-1e-14 (block.hpp:946-1000).
template <int BS>
struct PcBlockAMG : Pc {
    static constexpr int BV = BS * BS;
    struct Level {
        Bsr<BS> A, P, R;
        vec dinv;
        uint32_t n;
        double om;  // smoother weight of this level
    };
    // SGS smoother (BS = 1 only; no counterpart in the reference): per level
    // the greedy colouring of pattern(A + A^T) in row order (row i takes the
    // least colour none of its lower-numbered neighbours holds). With at most
    // kSgsColours colours the level is smoothed by Gauss-Seidel colour by
    // colour (ascending before the coarse correction, descending after it),
    // u_i <- u_i + (w (f_i - sum_j a_ij u_j)) / a_ii with w the level's
    // smoother weight; otherwise by damped Jacobi with a_ii and the weight
    // (4/3) / max_i sum_j |a_ij / a_ii|.
    static constexpr int kSgsColours = 8;
    struct Sgs {
        bool gs = false;
        std::vector<std::vector<uint32_t>> rows;  // rows of each colour, ascending
        vec diag;
        double jw = 0.0;
    };
    std::vector<Level> lv;
    std::vector<Sgs> sg;
    double omega;
    int npre, npost;
    bool coarse_exact;
    int smoother;  // 0 Jacobi (the reference's), 1 SGS
    PcBlockAMG(const Bsr<BS>& A0, double theta, double om, int pre, int post, int max_levels, int max_agg,
               uint32_t coarse_max, double p_omega, bool exact = false, int smoother_kind = 0)
        : omega(om), npre(pre), npost(post), coarse_exact(exact), smoother(smoother_kind) {
        if (smoother != 0 && BS != 1) throw std::invalid_argument("the SGS smoother is for 1x1 blocks only");
        build(A0, theta, om, max_levels, max_agg, coarse_max, p_omega);
        if (smoother != 0)
            for (size_t l = 0; l + 1 < lv.size(); ++l) sg.push_back(sgs_setup(lv[l].A));
    }
    static Sgs sgs_setup(const Bsr<BS>& A) {
        Sgs S;
        const uint32_t n = A.nbr;
        S.diag.assign(n, 0.0);
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
                if (A.ci[p] == i) { S.diag[i] = A.blk(p)[0]; break; }
        double rho = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            double rs = 0.0;
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) rs += std::fabs(A.blk(p)[0] / S.diag[i]);
            rho = std::max(rho, rs);
        }
        S.jw = rho > 0.0 ? (4.0 / 3.0) / rho : 2.0 / 3.0;
        std::vector<std::vector<uint32_t>> low(n);
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
                const uint32_t j = A.ci[p];
                if (j != i) low[std::max(i, j)].push_back(std::min(i, j));
            }
        std::vector<int> col(n, -1);
        int ncol = 0;
        for (uint32_t i = 0; i < n; ++i) {
            std::vector<char> used(low[i].size() + 2, 0);
            for (uint32_t j : low[i])
                if ((size_t)col[j] < used.size()) used[col[j]] = 1;
            int c = 0;
            while (used[c]) ++c;
            col[i] = c;
            ncol = std::max(ncol, c + 1);
        }
        S.gs = ncol <= kSgsColours;
        if (S.gs) {
            S.rows.assign(ncol, {});
            for (uint32_t i = 0; i < n; ++i) S.rows[col[i]].push_back(i);
        }
        return S;
    }
    void build(const Bsr<BS>& A0, double theta, double om, int max_levels, int max_agg, uint32_t coarse_max,
               double p_omega) {
        // Weights <= 0 select the robust variant (no counterpart in the
        // reference): per level w = (4/3) / rho with rho the Gershgorin bound
        // max_row sum |D^{-1} A|, for the smoother (om <= 0) and for the
        // prolongator (p_omega <= 0); the latter also stops coarsening when a
        // level keeps more than 85% of its rows.
        if (A0.nbr != A0.nbc) throw std::invalid_argument("Block AMG needs a square matrix");
        const bool auto_p = p_omega <= 0.0;
        Bsr<BS> cur = A0;
        for (int l = 0; l < max_levels; ++l) {
            const uint32_t n = cur.nbr;
            Level L;
            L.A = cur;
            L.n = n;
            L.dinv.resize((size_t)n * BV);
            for (uint32_t i = 0; i < n; ++i) blk_invert<BS>(cur.blk(cur.dptr[i]), &L.dinv[(size_t)i * BV]);
            double rho = 0.0;
            if (om <= 0.0 || auto_p) {
                for (uint32_t i = 0; i < n; ++i) {
                    double rs[BS] = {0};
                    const double* di = &L.dinv[(size_t)i * BV];
                    for (uint32_t p = cur.rp[i]; p < cur.rp[i + 1]; ++p) {
                        const double* ab = cur.blk(p);
                        for (int r = 0; r < BS; ++r)
                            for (int c = 0; c < BS; ++c) {
                                double s = 0.0;
                                for (int mm = 0; mm < BS; ++mm) s += di[r * BS + mm] * ab[mm * BS + c];
                                rs[r] += std::fabs(s);
                            }
                    }
                    for (int r = 0; r < BS; ++r) rho = std::max(rho, rs[r]);
                }
            }
            L.om = om > 0.0 ? om : (rho > 0.0 ? (4.0 / 3.0) / rho : 2.0 / 3.0);
            const double pw = auto_p ? (rho > 0.0 ? (4.0 / 3.0) / rho : 2.0 / 3.0) : p_omega;
            if (n <= coarse_max) { lv.push_back(std::move(L)); break; }
            std::vector<int> agg(n, -1);
            int na = 0;
            if (auto_p) {
                na = mis2_agg(n, cur.rp, cur.ci, [&](uint32_t p) { return fro(cur.blk(p)); }, theta, agg);
            } else {
                std::vector<std::vector<uint32_t>> S(n);
                for (uint32_t i = 0; i < n; ++i) {
                    double mx = 0.0;
                    for (uint32_t p = cur.rp[i]; p < cur.rp[i + 1]; ++p)
                        if (cur.ci[p] != i) mx = std::max(mx, fro(cur.blk(p)));
                    const double thr = theta * mx;
                    for (uint32_t p = cur.rp[i]; p < cur.rp[i + 1]; ++p)
                        if (cur.ci[p] != i && fro(cur.blk(p)) >= thr) S[i].push_back(cur.ci[p]);
                }
                for (uint32_t i = 0; i < n; ++i) {
                    if (agg[i] != -1) continue;
                    agg[i] = na;
                    int sz = 1;
                    for (uint32_t j : S[i])
                        if (agg[j] == -1 && sz < max_agg) { agg[j] = na; ++sz; }
                    ++na;
                }
            }
            Bsr<BS> P;
            P.nbr = n;
            P.nbc = (uint32_t)na;
            P.rp.assign(n + 1, 0);
            for (uint32_t i = 0; i < n; ++i) {
                std::map<uint32_t, std::vector<double>> row;
                row[(uint32_t)agg[i]] = std::vector<double>(BV, 0.0);
                for (int r = 0; r < BS; ++r) row[(uint32_t)agg[i]][r * BS + r] = 1.0;
                const double* di = &L.dinv[(size_t)i * BV];
                for (uint32_t p = cur.rp[i]; p < cur.rp[i + 1]; ++p) {
                    const uint32_t g = (uint32_t)agg[cur.ci[p]];
                    const double* ab = cur.blk(p);
                    if (!row.count(g)) row[g] = std::vector<double>(BV, 0.0);
                    for (int r = 0; r < BS; ++r)
                        for (int c = 0; c < BS; ++c) {
                            double s = 0.0;
                            for (int mm = 0; mm < BS; ++mm) s += di[r * BS + mm] * ab[mm * BS + c];
                            row[g][r * BS + c] -= pw * s;
                        }
                }
                for (auto& kv : row) {
                    P.ci.push_back(kv.first);
                    for (int t = 0; t < BV; ++t) P.v.push_back(kv.second[t]);
                }
                P.rp[i + 1] = (uint32_t)P.ci.size();
            }
            P.cache_diag();
            if (auto_p && (double)na > 0.85 * (double)n) { lv.push_back(std::move(L)); break; }
            L.P = P;
            L.R = P.transpose();
            lv.push_back(std::move(L));
            Bsr<BS> AP = blk_matmul<BS>(cur, lv.back().P);
            cur = blk_matmul<BS>(lv.back().R, AP);
        }
    }
    static double fro(const double* b) {
        double s = 0.0;
        for (int i = 0; i < BV; ++i) s += b[i] * b[i];
        return std::sqrt(s);
    }
    void smooth_sgs(size_t l, const vec& rhs, vec& sol, int ns, bool reverse) {
        const Level& L = lv[l];
        const Sgs& S = sg[l];
        const uint32_t n = L.n;
        if (!S.gs) {
            vec tmp;
            for (int s = 0; s < ns; ++s) {
                L.A.matvec(sol, tmp);
                for (uint32_t i = 0; i < n; ++i) sol[i] = sol[i] + (S.jw * (rhs[i] - tmp[i])) / S.diag[i];
            }
            return;
        }
        const int nc = (int)S.rows.size();
        for (int s = 0; s < ns; ++s)
            for (int cc = 0; cc < nc; ++cc)
                for (uint32_t i : S.rows[reverse ? nc - 1 - cc : cc]) {
                    double acc = 0.0;
                    for (uint32_t p = L.A.rp[i]; p < L.A.rp[i + 1]; ++p) acc += L.A.blk(p)[0] * sol[L.A.ci[p]];
                    sol[i] = sol[i] + (L.om * (rhs[i] - acc)) / S.diag[i];
                }
    }
    void smooth(const Level& L, const vec& rhs, vec& sol, int ns) {
        const uint32_t n = L.n, nd = n * BS;
        vec tmp, r(nd);
        for (int s = 0; s < ns; ++s) {
            L.A.matvec(sol, tmp);
            for (uint32_t i = 0; i < nd; ++i) r[i] = rhs[i] - tmp[i];
            for (uint32_t i = 0; i < n; ++i) {
                const double* d = &L.dinv[(size_t)i * BV];
                for (int c = 0; c < BS; ++c) {
                    double corr = 0.0;
                    for (int rr = 0; rr < BS; ++rr) corr += d[c * BS + rr] * r[i * BS + rr];
                    sol[i * BS + c] += L.om * corr;
                }
            }
        }
    }
    void coarse_solve(const Level& L, const vec& rhs, vec& sol) {
        const int sz = (int)(L.n * BS);
        vec d((size_t)sz * sz, 0.0);
        for (uint32_t i = 0; i < L.n; ++i)
            for (uint32_t p = L.A.rp[i]; p < L.A.rp[i + 1]; ++p) {
                const uint32_t j = L.A.ci[p];
                const double* b = L.A.blk(p);
                for (int r = 0; r < BS; ++r)
                    for (int c = 0; c < BS; ++c) d[(size_t)(i * BS + r) * sz + j * BS + c] = b[r * BS + c];
            }
        auto fl = [](double x) { return std::fabs(x) < 1e-14 ? (x >= 0 ? 1e-14 : -1e-14) : x; };
        if (coarse_exact) {
            // Exact alternative: partial-pivot LU applied to the right-hand
            // side with the row swaps, pivots floored to +-1e-14.
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
                d[(size_t)k * sz + k] = fl(d[(size_t)k * sz + k]);
                for (int i = k + 1; i < sz; ++i) {
                    d[(size_t)i * sz + k] /= d[(size_t)k * sz + k];
                    for (int j = k + 1; j < sz; ++j) d[(size_t)i * sz + j] -= d[(size_t)i * sz + k] * d[(size_t)k * sz + j];
                }
            }
            vec x(sz);
            for (int i = 0; i < sz; ++i) x[i] = rhs[pv[i]];
            for (int i = 0; i < sz; ++i)
                for (int j = 0; j < i; ++j) x[i] -= d[(size_t)i * sz + j] * x[j];
            for (int i = sz - 1; i >= 0; --i) {
                for (int j = i + 1; j < sz; ++j) x[i] -= d[(size_t)i * sz + j] * x[j];
                x[i] /= d[(size_t)i * sz + i];
            }
            sol = x;
            return;
        }
        for (int k = 0; k < sz; ++k) {
            int pk = k;
            for (int i = k + 1; i < sz; ++i)
                if (std::fabs(d[(size_t)i * sz + k]) > std::fabs(d[(size_t)pk * sz + k])) pk = i;
            if (pk != k)
                for (int j = 0; j < sz; ++j) std::swap(d[(size_t)k * sz + j], d[(size_t)pk * sz + j]);
            const double dk = fl(d[(size_t)k * sz + k]);
            for (int i = k + 1; i < sz; ++i) {
                const double f = d[(size_t)i * sz + k] / dk;
                for (int j = k; j < sz; ++j) d[(size_t)i * sz + j] -= f * d[(size_t)k * sz + j];
            }
        }
        // The reference eliminates the matrix rows with pivoting but runs the
        // right-hand side through the multipliers without the row swaps
        // (block.hpp:985-991); after elimination the stored sub-diagonal
        // entries are the post-update values, which the forward pass reuses.
        vec x(rhs.begin(), rhs.end());
        for (int k = 0; k < sz; ++k) {
            const double dk = fl(d[(size_t)k * sz + k]);
            for (int i = k + 1; i < sz; ++i) x[i] -= (d[(size_t)i * sz + k] / dk) * x[k];
        }
        for (int i = sz - 1; i >= 0; --i) {
            for (int j = i + 1; j < sz; ++j) x[i] -= d[(size_t)i * sz + j] * x[j];
            x[i] /= fl(d[(size_t)i * sz + i]);
        }
        sol = x;
    }
    void vcycle(size_t l, const vec& rhs, vec& sol) {
        const Level& L = lv[l];
        const uint32_t nd = L.n * BS;
        sol.resize(nd);
        if (l + 1 == lv.size()) { coarse_solve(L, rhs, sol); return; }
        if (smoother) smooth_sgs(l, rhs, sol, npre, false);
        else smooth(L, rhs, sol, npre);
        vec tmp, r(nd), rc, ec, pec;
        L.A.matvec(sol, tmp);
        for (uint32_t i = 0; i < nd; ++i) r[i] = rhs[i] - tmp[i];
        L.R.matvec(r, rc);
        vcycle(l + 1, rc, ec);
        L.P.matvec(ec, pec);
        for (uint32_t i = 0; i < nd; ++i) sol[i] += pec[i];
        if (smoother) smooth_sgs(l, rhs, sol, npost, true);
        else smooth(L, rhs, sol, npost);
    }
    void apply(const vec& r, vec& z) override {
        z.assign(r.size(), 0.0);
        vcycle(0, r, z);
    }
};

// make_schur_preconditioner (block.hpp:625-685): block upper-triangular
// preconditioner of the saddle-point matrix [[A, B^T], [B, 0]] from a
// velocity preconditioner A_pc and a Schur-complement preconditioner S_pc:
//   w_u = A_pc(r_u); z_p = S_pc(B w_u - r_p); z_u = A_pc(r_u - B^T z_p).
// S_pc may be absent (the identity, the reference test's baseline); its
// output is multiplied by s_scale (the reference test's scaled AMG,
// test_BlockMatrix_schur3d.cpp:149-160).
struct PcSchur : Pc {
    const Csr* B;
    const Csr* Bt;
    Pc* A;
    Pc* S;
    double scale;
    PcSchur(const Csr* B_, const Csr* Bt_, Pc* A_, Pc* S_, double sc) : B(B_), Bt(Bt_), A(A_), S(S_), scale(sc) {
        if (Bt->n != B->m || Bt->m != B->n) throw std::invalid_argument("Schur: B^T must be the transpose shape of B");
    }
    void apply(const vec& r, vec& z) override {
        const size_t nv = B->m, np = B->n;
        vec ru(r.begin(), r.begin() + nv), rp(r.begin() + nv, r.begin() + nv + np);
        vec wu, wp, zp, wbt, zu;
        A->apply(ru, wu);
        spmv(*B, wu, wp);
        vec rmod(np);
        for (size_t i = 0; i < np; ++i) rmod[i] = wp[i] - rp[i];
        if (S) {
            vec w;
            S->apply(rmod, w);
            zp.resize(np);
            for (size_t i = 0; i < np; ++i) zp[i] = scale * w[i];
        } else {
            zp.resize(np);
            for (size_t i = 0; i < np; ++i) zp[i] = scale * rmod[i];
        }
        spmv(*Bt, zp, wbt);
        vec umod(nv);
        for (size_t i = 0; i < nv; ++i) umod[i] = ru[i] - wbt[i];
        A->apply(umod, zu);
        z.resize(nv + np);
        std::copy(zu.begin(), zu.end(), z.begin());
        std::copy(zp.begin(), zp.end(), z.begin() + nv);
    }
};

// ---------------------------------------------------------------------------
// Krylov solvers
// ---------------------------------------------------------------------------
using MatVec = std::function<void(const vec&, vec&)>;

// pcg (array.cpp:3445-3601); returns the reference's iteration index.
static int pcg(const MatVec& A, uint32_t n, const vec& b, Pc& M, double tol, int max_iter, vec& x) {
    if (b.size() != n) throw std::invalid_argument("PCG: matrix columns differ from vector length");
    x.assign(n, 0.0);
    const double normb = std::sqrt(dot(b, b));
    if (normb < 1e-300) return 0;
    vec r = b;
    double normr = std::sqrt(dot(r, r));
    if (normr / normb < tol) return 0;
    vec z(n, 0.0), p, q;
    M.apply(r, z);
    p = z;
    double rz = dot(r, z);
    if (rz <= 0.0) throw std::runtime_error("PCG: r^T M^{-1} r <= 0, preconditioner may not be SPD");
    int it_end = 0;
    for (int it = 0; it < max_iter; ++it) {
        A(p, q);
        const double pAp = dot(p, q);
        if (std::fabs(pAp) < 1e-30) throw std::runtime_error("PCG breakdown: p^T A p too small");
        if (pAp <= 0.0) throw std::runtime_error("PCG: p^T A p <= 0, matrix may not be SPD");
        const double alpha = rz / pAp;
        const int64_t nn = n;
#pragma omp parallel for schedule(static) if (nn > 1000)
        for (int64_t i = 0; i < nn; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        normr = std::sqrt(dot(r, r));
        if (normr / normb < tol) break;
        M.apply(r, z);
        const double rz2 = dot(r, z);
        if (rz2 <= 0.0) throw std::runtime_error("PCG: r^T M^{-1} r <= 0, preconditioner may not be SPD");
        const double beta = rz2 / rz;
#pragma omp parallel for schedule(static) if (nn > 1000)
        for (int64_t i = 0; i < nn; ++i) p[i] = z[i] + beta * p[i];
        rz = rz2;
        it_end = it;
    }
    return it_end;
}

// bicgstab (array.cpp:3602-3842) with the true-residual replacement rule.
static int bicgstab(const MatVec& A, uint32_t n, const vec& b, Pc& M, double tol, int max_iter, vec& x) {
    if (b.size() != n) throw std::invalid_argument("BiCGSTAB: matrix columns differ from vector length");
    x.assign(n, 0.0);
    const double normb = std::sqrt(dot(b, b));
    if (normb < 1e-300) return 0;
    vec r = b;
    double normr = std::sqrt(dot(r, r));
    if (normr / normb < tol) return 0;
    const vec rt = r;
    vec p(n, 0.0), v(n, 0.0), y(n, 0.0), z(n, 0.0), s(n, 0.0), t(n, 0.0), Ax;
    double rho_old = 1.0, alpha = 1.0, omega = 1.0;
    const double bt = 1e-30;
    int it_end = 0;
    const int64_t nn = n;
    for (int it = 0; it < max_iter; ++it) {
        const double rho = dot(rt, r);
        if (std::fabs(rho) < bt) throw std::runtime_error("BiCGSTAB breakdown: rho too small");
        if (std::fabs(rho_old) < bt) throw std::runtime_error("BiCGSTAB breakdown: previous rho too small");
        if (std::fabs(omega) < bt) throw std::runtime_error("BiCGSTAB breakdown: omega = 0");
        const double beta = (rho / rho_old) * (alpha / omega);
#pragma omp parallel for schedule(static) if (nn > 1000)
        for (int64_t i = 0; i < nn; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        M.apply(p, y);
        A(y, v);
        const double rv = dot(rt, v);
        if (std::fabs(rv) < bt) throw std::runtime_error("BiCGSTAB breakdown: rtilde^T v too small");
        alpha = rho / rv;
#pragma omp parallel for schedule(static) if (nn > 1000)
        for (int64_t i = 0; i < nn; ++i) s[i] = r[i] - alpha * v[i];
        M.apply(s, z);
        A(z, t);
        const double tt = dot(t, t);
        if (tt < bt) throw std::runtime_error("BiCGSTAB breakdown: t^T t too small");
        omega = dot(t, s) / tt;
#pragma omp parallel for schedule(static) if (nn > 1000)
        for (int64_t i = 0; i < nn; ++i) { x[i] += alpha * y[i] + omega * z[i]; r[i] = s[i] - omega * t[i]; }
        normr = std::sqrt(dot(r, r));
        it_end = it;
        bool conv = false;
        if (normr / normb < tol) {
            A(x, Ax);
            double tr = 0.0;
#pragma omp parallel for reduction(+ : tr) schedule(static) if (nn > 1000)
            for (int64_t i = 0; i < nn; ++i) { const double d = b[i] - Ax[i]; tr += d * d; }
            if (std::sqrt(tr) / normb < tol) conv = true;
            else {
                for (int64_t i = 0; i < nn; ++i) r[i] = b[i] - Ax[i];
                const double rc = dot(rt, r);
                if (std::fabs(rc) < bt) conv = true;
                else { rho_old = rc; continue; }
            }
        } else {
            rho_old = rho;
        }
        if (conv) break;
    }
    return it_end;
}

// gmres (array.cpp:4190-4401): restarted, right-preconditioned, MGS + Givens.
static int gmres(const MatVec& A, uint32_t n, const vec& b, Pc& M, int m, double tol, int max_iter, vec& x) {
    if (b.size() != n) throw std::invalid_argument("GMRES: matrix columns differ from vector length");
    x.assign(n, 0.0);
    const double normb = std::sqrt(dot(b, b));
    if (normb < 1e-300) return 0;
    const double bt = 1e-30;
    std::vector<vec> V(m + 1, vec(n, 0.0));
    std::vector<vec> H(m + 1, vec(m, 0.0));
    vec cs(m, 0.0), sn(m, 0.0), g(m + 1, 0.0), y(m, 0.0), w(n), wt(n), z(n), dz(n), Ax;
    int total = 0, it_end = 0;
    bool conv = false;
    while (total < max_iter && !conv) {
        double beta;
        if (total == 0) {
            beta = normb;
            for (uint32_t i = 0; i < n; ++i) V[0][i] = b[i] / beta;
        } else {
            A(x, Ax);
            double rr = 0.0;
            for (uint32_t i = 0; i < n; ++i) { const double ri = b[i] - Ax[i]; V[0][i] = ri; rr += ri * ri; }
            beta = std::sqrt(rr);
            if (beta / normb < tol) { conv = true; break; }
            for (uint32_t i = 0; i < n; ++i) V[0][i] /= beta;
        }
        g.assign(m + 1, 0.0);
        g[0] = beta;
        for (int i = 0; i < m + 1; ++i) H[i].assign(m, 0.0);
        int ju = 0;
        for (int j = 0; j < m && total < max_iter; ++j) {
            M.apply(V[j], wt);
            A(wt, w);
            for (int i = 0; i <= j; ++i) {
                H[i][j] = dot(w, V[i]);
                const double h = H[i][j];
                for (uint32_t t = 0; t < n; ++t) w[t] -= h * V[i][t];
            }
            H[j + 1][j] = std::sqrt(dot(w, w));
            if (H[j + 1][j] > bt) {
                const double inv = 1.0 / H[j + 1][j];
                for (uint32_t t = 0; t < n; ++t) V[j + 1][t] = w[t] * inv;
            }
            for (int i = 0; i < j; ++i) {
                const double t1 = cs[i] * H[i][j] + sn[i] * H[i + 1][j];
                const double t2 = -sn[i] * H[i][j] + cs[i] * H[i + 1][j];
                H[i][j] = t1;
                H[i + 1][j] = t2;
            }
            const double den = std::hypot(H[j][j], H[j + 1][j]);
            if (den < bt) throw std::runtime_error("GMRES breakdown: Hessenberg pivot too small");
            cs[j] = H[j][j] / den;
            sn[j] = H[j + 1][j] / den;
            H[j][j] = cs[j] * H[j][j] + sn[j] * H[j + 1][j];
            H[j + 1][j] = 0.0;
            const double g1 = cs[j] * g[j] + sn[j] * g[j + 1];
            const double g2 = -sn[j] * g[j] + cs[j] * g[j + 1];
            g[j] = g1;
            g[j + 1] = g2;
            ju = j + 1;
            it_end = total;
            ++total;
            if (std::fabs(g[j + 1]) / normb < tol) { conv = true; break; }
        }
        for (int i = ju - 1; i >= 0; --i) {
            double s = g[i];
            for (int k = i + 1; k < ju; ++k) s -= H[i][k] * y[k];
            if (std::fabs(H[i][i]) < bt) throw std::runtime_error("GMRES breakdown: upper triangular pivot too small");
            y[i] = s / H[i][i];
        }
        for (uint32_t t = 0; t < n; ++t) {
            double a = 0.0;
            for (int i = 0; i < ju; ++i) a += y[i] * V[i][t];
            z[t] = a;
        }
        M.apply(z, dz);
        for (uint32_t t = 0; t < n; ++t) x[t] += dz[t];
    }
    return it_end;
}

// mini_cg (test_BlockMatrix_mpi.cpp:157-182): plain PCG, no breakdown checks,
// returns the raw loop index at which it stopped.
static int mini_cg(const MatVec& A, const vec& b, Pc& M, double tol, int maxit, vec& x) {
    const size_t n = b.size();
    x.assign(n, 0.0);
    vec r = b, z, p, Ap;
    M.apply(r, z);
    p = z;
    double rz = 0.0;
    for (size_t i = 0; i < n; ++i) rz += r[i] * z[i];
    double nb = 0.0;
    for (size_t i = 0; i < n; ++i) nb += b[i] * b[i];
    nb = std::sqrt(nb);
    int it = 0;
    for (; it < maxit; ++it) {
        A(p, Ap);
        double pAp = 0.0;
        for (size_t i = 0; i < n; ++i) pAp += p[i] * Ap[i];
        const double alpha = rz / pAp;
        for (size_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        double nr = 0.0;
        for (size_t i = 0; i < n; ++i) nr += r[i] * r[i];
        if (std::sqrt(nr) / nb < tol) break;
        M.apply(r, z);
        double rz2 = 0.0;
        for (size_t i = 0; i < n; ++i) rz2 += r[i] * z[i];
        const double beta = rz2 / rz;
        for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rz = rz2;
    }
    return it;
}

// to_lower + alias tables (array.cpp:2238-2250, 3484-3511, 3641-3669, 4215-4238)
static std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}
enum Kind { K_NONE = 0, K_DIAG, K_IC0, K_ILU0, K_SSOR, K_AMG, K_SAAMG };
// solver: 0 pcg, 1 bicgstab, 2 gmres
static int alias(const std::string& name0, int solver) {
    const std::string s = lower(name0);
    if (s == "none" || s == "identity" || s == "noprecond") return K_NONE;
    if (solver == 0 && s == "cg") return K_NONE;
    if (s == "diag" || s == "diagonal" || s == "jacobi") return K_DIAG;
    if (s == "ic" || s == "ic0" || s == "ichol") return K_IC0;
    if (solver != 2 && s == "incomplete_cholesky") return K_IC0;
    if (s == "ilu" || s == "ilu0") return K_ILU0;
    if (s == "sor" || s == "ssor") return K_SSOR;
    if (s == "amg") return K_AMG;
    throw std::invalid_argument("unknown preconditioner type: " + name0);
}

}  // namespace orc

// ===========================================================================
// C ABI for the test harness (ctypes)
// ===========================================================================
using namespace orc;

static thread_local std::string g_err;

template <class F>
static int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return 1;
    } catch (const std::out_of_range& e) {
        g_err = e.what();
        return 2;
    } catch (const std::runtime_error& e) {
        g_err = e.what();
        return 3;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 4;
    }
}

struct OrPc {
    std::unique_ptr<Pc> pc;
    int kind = 0;
};

struct OrBsr {
    int bs;
    std::unique_ptr<Bsr<1>> b1;
    std::unique_ptr<Bsr<3>> b3;
};

extern "C" {

const char* or_last_error(void) { return g_err.c_str(); }

void or_set_threads(int t) {
#ifdef _OPENMP
    omp_set_num_threads(t);
#else
    (void)t;
#endif
}

int or_csr_from_coo(uint32_t n, uint32_t m, uint64_t nin, const double* v, const uint32_t* r, const uint32_t* c, Csr** out) {
    return guard([&] {
        vec vv(v, v + nin);
        std::vector<uint32_t> rr(r, r + nin), cc(c, c + nin);
        *out = new Csr(csr_from_coo(n, m, vv, rr, cc));
    });
}

int or_csr_create(uint32_t n, uint32_t m, const uint32_t* rp, const uint32_t* ci, const double* v, Csr** out) {
    return guard([&] {
        Csr* A = new Csr;
        A->n = n;
        A->m = m;
        A->rp.assign(rp, rp + n + 1);
        A->ci.assign(ci, ci + rp[n]);
        A->v.assign(v, v + rp[n]);
        *out = A;
    });
}

void or_csr_destroy(Csr* A) { delete A; }
uint64_t or_csr_nnz(const Csr* A) { return A->nnz(); }
uint32_t or_csr_rows(const Csr* A) { return A->n; }
uint32_t or_csr_cols(const Csr* A) { return A->m; }
void or_csr_export(const Csr* A, uint32_t* rp, uint32_t* ci, double* v) {
    std::copy(A->rp.begin(), A->rp.end(), rp);
    std::copy(A->ci.begin(), A->ci.end(), ci);
    std::copy(A->v.begin(), A->v.end(), v);
}

int or_csr_transpose(const Csr* A, Csr** out) {
    return guard([&] { *out = new Csr(transpose(*A)); });
}
int or_csr_spgemm(const Csr* A, const Csr* B, Csr** out) {
    return guard([&] { *out = new Csr(spgemm(*A, *B)); });
}
int or_spmv(const Csr* A, const double* x, double* y) {
    return guard([&] {
        vec xx(x, x + A->m), yy;
        spmv(*A, xx, yy);
        std::copy(yy.begin(), yy.end(), y);
    });
}
int or_dot(uint64_t n, const double* a, const double* b, double* out) {
    return guard([&] {
        vec aa(a, a + n), bb(b, b + n);
        *out = dot(aa, bb);
    });
}

// params: [0] omega (SSOR) | AMG: [0] theta [1] coarse_thresh [2] max_levels
// SAAMG: [0] theta [1] smoother omega [2] npre [3] npost [4] max_levels
//        [5] max_agg [6] coarse_max [7] prolongator omega [8] exact coarse solve
//        [9] smoother (0 Jacobi, 1 SGS)
int or_pc_create(const Csr* A, int kind, const double* prm, OrPc** out) {
    return guard([&] {
        auto h = std::make_unique<OrPc>();
        h->kind = kind;
        switch (kind) {
            case K_NONE: h->pc.reset(new PcNone); break;
            case K_DIAG: h->pc.reset(new PcDiag(*A)); break;
            case K_IC0: h->pc.reset(new PcIC0(*A)); break;
            case K_ILU0: h->pc.reset(new PcILU0(*A)); break;
            case K_SSOR: h->pc.reset(new PcSSOR(*A, prm ? prm[0] : 1.2)); break;
            case K_AMG:
                h->pc.reset(new PcAmgCF(*A, prm ? prm[0] : 0.25, prm ? (uint32_t)prm[1] : 300u, prm ? (int)prm[2] : 25));
                break;
            case K_SAAMG: {
                if (A->n != A->m) throw std::invalid_argument("Block AMG needs a square matrix");
                Bsr<1> B = condense<1>(*A, A->n, A->m);
                const double d[10] = {0.25, 2.0 / 3.0, 1, 1, 20, 4, 50, 2.0 / 3.0, 0, 0};
                const double* q = prm ? prm : d;
                h->pc.reset(new PcBlockAMG<1>(B, q[0], q[1], (int)q[2], (int)q[3], (int)q[4], (int)q[5], (uint32_t)q[6], q[7],
                                              q[8] != 0.0, (int)q[9]));
                break;
            }
            default: throw std::invalid_argument("unknown preconditioner kind");
        }
        *out = h.release();
    });
}

int or_pc_apply(OrPc* h, uint64_t n, const double* r, double* z) {
    return guard([&] {
        vec rr(r, r + n), zz;
        h->pc->apply(rr, zz);
        std::copy(zz.begin(), zz.end(), z);
    });
}
void or_pc_destroy(OrPc* h) { delete h; }

// IC(0) factor export: L (lower with diagonal) in CSR
uint64_t or_ic0_nnz(OrPc* h) { return static_cast<PcIC0*>(h->pc.get())->Lv.size(); }
void or_ic0_export(OrPc* h, uint32_t* rp, uint32_t* ci, double* v) {
    auto* p = static_cast<PcIC0*>(h->pc.get());
    std::copy(p->Lp.begin(), p->Lp.end(), rp);
    std::copy(p->Lc.begin(), p->Lc.end(), ci);
    std::copy(p->Lv.begin(), p->Lv.end(), v);
}
// ILU(0) export: L strict lower, U upper with diagonal
void or_ilu0_sizes(OrPc* h, uint64_t* lnnz, uint64_t* unnz) {
    auto* p = static_cast<PcILU0*>(h->pc.get());
    *lnnz = p->Lv.size();
    *unnz = p->Uv.size();
}
void or_ilu0_export(OrPc* h, uint32_t* lp, uint32_t* lc, double* lv, uint32_t* up, uint32_t* uc, double* uv) {
    auto* p = static_cast<PcILU0*>(h->pc.get());
    std::copy(p->Lp.begin(), p->Lp.end(), lp);
    std::copy(p->Lc.begin(), p->Lc.end(), lc);
    std::copy(p->Lv.begin(), p->Lv.end(), lv);
    std::copy(p->Up.begin(), p->Up.end(), up);
    std::copy(p->Uc.begin(), p->Uc.end(), uc);
    std::copy(p->Uv.begin(), p->Uv.end(), uv);
}
// SSOR export: colour count, new_id (old -> new)
uint32_t or_ssor_export(OrPc* h, uint32_t* new_id, uint32_t* color_ptr) {
    auto* p = static_cast<PcSSOR*>(h->pc.get());
    if (new_id) std::copy(p->D->new_id.begin(), p->D->new_id.end(), new_id);
    if (color_ptr) std::copy(p->D->color_ptr.begin(), p->D->color_ptr.end(), color_ptr);
    return p->D->ncolor;
}
// AMG hierarchy summary: per level rows and nnz of A
int or_amg_levels(OrPc* h, uint32_t* rows, uint64_t* nnz, int cap) {
    if (h->kind == K_AMG) {
        auto* p = static_cast<PcAmgCF*>(h->pc.get());
        for (int i = 0; i < (int)p->lv.size() && i < cap; ++i) { rows[i] = p->lv[i].A.n; nnz[i] = p->lv[i].A.nnz(); }
        return (int)p->lv.size();
    }
    if (auto* p3 = dynamic_cast<PcBlockAMG<3>*>(h->pc.get())) {
        for (int i = 0; i < (int)p3->lv.size() && i < cap; ++i) { rows[i] = p3->lv[i].n; nnz[i] = p3->lv[i].A.nnzb(); }
        return (int)p3->lv.size();
    }
    auto* p = static_cast<PcBlockAMG<1>*>(h->pc.get());
    for (int i = 0; i < (int)p->lv.size() && i < cap; ++i) { rows[i] = p->lv[i].n; nnz[i] = p->lv[i].A.nnzb(); }
    return (int)p->lv.size();
}
// C/F split of level l of a C/F hierarchy (1 C, 2 F); -1 if l has none.
int or_amg_cf_split(OrPc* h, int l, int8_t* st) {
    if (h->kind != K_AMG) return -1;
    auto* p = static_cast<PcAmgCF*>(h->pc.get());
    if (l < 0 || l >= (int)p->split.size()) return -1;
    for (size_t i = 0; i < p->split[l].size(); ++i) st[i] = (int8_t)p->split[l][i];
    return 0;
}
// Block ILU(0) factors: L (strict lower blocks) and U (upper blocks with the
// diagonal), each as block CSR with bs*bs row-major values. With null arrays
// only the block counts are returned.
int or_bilu0_export(OrPc* h, uint64_t* lnnzb, uint64_t* unnzb, uint32_t* lp, uint32_t* lc, double* lv, uint32_t* up,
                    uint32_t* uc, double* uv) {
    auto exp = [&](auto* f) {
        *lnnzb = f->Lc.size();
        *unnzb = f->Uc.size();
        if (!lp) return;
        std::copy(f->Lp.begin(), f->Lp.end(), lp);
        std::copy(f->Lc.begin(), f->Lc.end(), lc);
        std::copy(f->Lv.begin(), f->Lv.end(), lv);
        std::copy(f->Up.begin(), f->Up.end(), up);
        std::copy(f->Uc.begin(), f->Uc.end(), uc);
        std::copy(f->Uv.begin(), f->Uv.end(), uv);
    };
    if (auto* f3 = dynamic_cast<PcBlockILU0<3>*>(h->pc.get())) exp(f3);
    else if (auto* f1 = dynamic_cast<PcBlockILU0<1>*>(h->pc.get())) exp(f1);
    else return -1;
    return 0;
}
// Robust-mode aggregation of a scalar CSR matrix (mis2_agg): agg[n], the
// final MIS-2 state per node (2 root, 0 other) and the number of aggregates.
int or_mis2_aggregate(const Csr* A, double theta, int32_t* agg, int32_t* state, int* na) {
    return guard([&] {
        std::vector<int> g, st;
        *na = mis2_agg(A->n, A->rp, A->ci, [&](uint32_t p) { return std::sqrt(A->v[p] * A->v[p]); }, theta, g, &st);
        std::copy(g.begin(), g.end(), agg);
        std::copy(st.begin(), st.end(), state);
    });
}
// Export one AMG level: A (and P when not coarsest) as CSR.
int or_amg_level_export(OrPc* h, int l, int which, uint32_t* rp, uint32_t* ci, double* v, uint32_t* dims) {
    const Csr* M = nullptr;
    Csr tmp;
    if (h->kind == K_AMG) {
        auto* p = static_cast<PcAmgCF*>(h->pc.get());
        M = which == 0 ? &p->lv[l].A : which == 1 ? &p->lv[l].P : &p->lv[l].R;
    } else if (auto* p3 = dynamic_cast<PcBlockAMG<3>*>(h->pc.get())) {
        // 3x3 blocks expanded to scalar CSR: scalar row 3I+a holds, block by
        // block in block-column order, the columns 3J+0..2 of row a of block (I, J)
        const Bsr<3>& B = which == 0 ? p3->lv[l].A : which == 1 ? p3->lv[l].P : p3->lv[l].R;
        tmp.n = B.nbr * 3;
        tmp.m = B.nbc * 3;
        tmp.rp.assign(1, 0);
        for (uint32_t I = 0; I < B.nbr; ++I)
            for (int a = 0; a < 3; ++a) {
                for (uint32_t q = B.rp[I]; q < B.rp[I + 1]; ++q)
                    for (int b = 0; b < 3; ++b) {
                        tmp.ci.push_back(B.ci[q] * 3 + b);
                        tmp.v.push_back(B.blk(q)[a * 3 + b]);
                    }
                tmp.rp.push_back((uint32_t)tmp.ci.size());
            }
        M = &tmp;
    } else {
        auto* p = static_cast<PcBlockAMG<1>*>(h->pc.get());
        const Bsr<1>& B = which == 0 ? p->lv[l].A : which == 1 ? p->lv[l].P : p->lv[l].R;
        tmp.n = B.nbr; tmp.m = B.nbc; tmp.rp = B.rp; tmp.ci = B.ci; tmp.v = B.v;
        M = &tmp;
    }
    dims[0] = M->n;
    dims[1] = M->m;
    dims[2] = M->nnz();
    if (rp) {
        std::copy(M->rp.begin(), M->rp.end(), rp);
        std::copy(M->ci.begin(), M->ci.end(), ci);
        std::copy(M->v.begin(), M->v.end(), v);
    }
    return 0;
}

// solver: 0 pcg, 1 bicgstab, 2 gmres
int or_solve(const Csr* A, int solver, OrPc* h, const double* b, double tol, int max_iter, int restart, double* x, int* it) {
    return guard([&] {
        if (A->n != A->m) throw std::invalid_argument("solver needs a square matrix");
        vec bb(b, b + A->n), xx;
        MatVec mv = [A](const vec& u, vec& w) { spmv(*A, u, w); };
        int r = 0;
        if (solver == 0) r = pcg(mv, A->n, bb, *h->pc, tol, max_iter, xx);
        else if (solver == 1) r = bicgstab(mv, A->n, bb, *h->pc, tol, max_iter, xx);
        else r = gmres(mv, A->n, bb, *h->pc, restart, tol, max_iter, xx);
        std::copy(xx.begin(), xx.end(), x);
        *it = r;
    });
}

// Named-preconditioner solve with the reference's per-solver alias table.
int or_solve_named(const Csr* A, int solver, const char* name, const double* b, double tol, int max_iter, int restart,
                   double* x, int* it) {
    return guard([&] {
        if (A->n != A->m) throw std::invalid_argument("solver needs a square matrix");
        // The reference checks ||b|| (and, for pcg/bicgstab, ||r0||/||b||)
        // before it resolves the preconditioner name (array.cpp:3462-3480).
        vec bb(b, b + A->n);
        const double normb = std::sqrt(dot(bb, bb));
        if (normb < 1e-300 || (solver != 2 && normb / normb < tol)) {
            std::fill(x, x + A->n, 0.0);
            *it = 0;
            return;
        }
        int k = alias(name, solver);
        OrPc* h = nullptr;
        const double om = 1.2;
        int rc = or_pc_create(A, k, k == K_SSOR ? &om : nullptr, &h);
        if (rc) {
            std::string e = g_err;
            if (rc == 1) throw std::invalid_argument(e);
            if (rc == 2) throw std::out_of_range(e);
            throw std::runtime_error(e);
        }
        std::unique_ptr<OrPc> hold(h);
        rc = or_solve(A, solver, h, b, tol, max_iter, restart, x, it);
        if (rc) {
            std::string e = g_err;
            if (rc == 1) throw std::invalid_argument(e);
            throw std::runtime_error(e);
        }
    });
}

// alias lookup exposed for tests: returns kind or -1
int or_alias(const char* name, int solver) {
    try {
        return alias(name, solver);
    } catch (...) {
        return -1;
    }
}

// ---- block path ----
int or_bsr_condense(const Csr* A, int bs, uint32_t nbr, uint32_t nbc, OrBsr** out) {
    return guard([&] {
        auto h = std::make_unique<OrBsr>();
        h->bs = bs;
        if (bs == 1) h->b1.reset(new Bsr<1>(condense<1>(*A, nbr, nbc)));
        else if (bs == 3) h->b3.reset(new Bsr<3>(condense<3>(*A, nbr, nbc)));
        else throw std::invalid_argument("block size must be 1 or 3");
        *out = h.release();
    });
}
void or_bsr_destroy(OrBsr* h) { delete h; }
uint64_t or_bsr_nnzb(const OrBsr* h) { return h->bs == 1 ? h->b1->nnzb() : h->b3->nnzb(); }
void or_bsr_export(const OrBsr* h, uint32_t* rp, uint32_t* ci, double* v) {
    if (h->bs == 1) {
        std::copy(h->b1->rp.begin(), h->b1->rp.end(), rp);
        std::copy(h->b1->ci.begin(), h->b1->ci.end(), ci);
        std::copy(h->b1->v.begin(), h->b1->v.end(), v);
    } else {
        std::copy(h->b3->rp.begin(), h->b3->rp.end(), rp);
        std::copy(h->b3->ci.begin(), h->b3->ci.end(), ci);
        std::copy(h->b3->v.begin(), h->b3->v.end(), v);
    }
}
int or_bsr_create(int bs, uint32_t nbr, uint32_t nbc, const uint32_t* rp, const uint32_t* ci, const double* v, OrBsr** out) {
    return guard([&] {
        auto h = std::make_unique<OrBsr>();
        h->bs = bs;
        const uint64_t nz = rp[nbr];
        if (bs == 3) {
            h->b3.reset(new Bsr<3>);
            auto& B = *h->b3;
            B.nbr = nbr; B.nbc = nbc;
            B.rp.assign(rp, rp + nbr + 1);
            B.ci.assign(ci, ci + nz);
            B.v.assign(v, v + nz * 9);
            B.cache_diag();
        } else if (bs == 1) {
            h->b1.reset(new Bsr<1>);
            auto& B = *h->b1;
            B.nbr = nbr; B.nbc = nbc;
            B.rp.assign(rp, rp + nbr + 1);
            B.ci.assign(ci, ci + nz);
            B.v.assign(v, v + nz);
            B.cache_diag();
        } else throw std::invalid_argument("block size must be 1 or 3");
        *out = h.release();
    });
}
int or_bsr_transpose(const OrBsr* h, OrBsr** out) {
    return guard([&] {
        auto t = std::make_unique<OrBsr>();
        t->bs = h->bs;
        if (h->bs == 1) t->b1.reset(new Bsr<1>(h->b1->transpose()));
        else t->b3.reset(new Bsr<3>(h->b3->transpose()));
        *out = t.release();
    });
}
int or_bsr_matvec(const OrBsr* h, const double* x, double* y) {
    return guard([&] {
        if (h->bs == 1) { vec xx(x, x + h->b1->nbc), yy; h->b1->matvec(xx, yy); std::copy(yy.begin(), yy.end(), y); }
        else { vec xx(x, x + (size_t)h->b3->nbc * 3), yy; h->b3->matvec(xx, yy); std::copy(yy.begin(), yy.end(), y); }
    });
}
// kind: 0 block-Jacobi, 1 block-ILU(0), 2 block SA-AMG (reference defaults)
int or_bsr_pc_create(const OrBsr* h, int kind, OrPc** out) {
    return guard([&] {
        auto p = std::make_unique<OrPc>();
        p->kind = 100 + kind;
        if (h->bs == 3) {
            if (kind == 0) p->pc.reset(new PcBlockJacobi<3>(*h->b3));
            else if (kind == 1) p->pc.reset(new PcBlockILU0<3>(*h->b3));
            else p->pc.reset(new PcBlockAMG<3>(*h->b3, 0.25, 2.0 / 3.0, 1, 1, 20, 4, 50, 2.0 / 3.0));
        } else {
            if (kind == 0) p->pc.reset(new PcBlockJacobi<1>(*h->b1));
            else if (kind == 1) p->pc.reset(new PcBlockILU0<1>(*h->b1));
            else p->pc.reset(new PcBlockAMG<1>(*h->b1, 0.25, 2.0 / 3.0, 1, 1, 20, 4, 50, 2.0 / 3.0));
        }
        *out = p.release();
    });
}
// Schur-complement preconditioner over oracle matrices and preconditioners
// (S may be null); it keeps pointers to them
int or_pc_create_schur(const Csr* B, const Csr* Bt, OrPc* A, OrPc* S, double scale, OrPc** out) {
    return guard([&] {
        auto p = std::make_unique<OrPc>();
        p->kind = 200;
        p->pc.reset(new PcSchur(B, Bt, A->pc.get(), S ? S->pc.get() : nullptr, scale));
        *out = p.release();
    });
}
// block SA-AMG with explicit parameters (as or_pc_create's SAAMG list:
// theta, smoother omega, npre, npost, max_levels, max_agg, coarse_max,
// prolongator omega, exact coarse solve)
int or_bsr_pc_create_amg(const OrBsr* h, const double* q, OrPc** out) {
    return guard([&] {
        auto p = std::make_unique<OrPc>();
        p->kind = 102;
        if (h->bs == 3)
            p->pc.reset(new PcBlockAMG<3>(*h->b3, q[0], q[1], (int)q[2], (int)q[3], (int)q[4], (int)q[5], (uint32_t)q[6],
                                          q[7], q[8] != 0.0));
        else
            p->pc.reset(new PcBlockAMG<1>(*h->b1, q[0], q[1], (int)q[2], (int)q[3], (int)q[4], (int)q[5], (uint32_t)q[6],
                                          q[7], q[8] != 0.0));
        *out = p.release();
    });
}
// solver: 0 mini_cg (test semantics), 1 BlockMatrix::gmres loop
int or_bsr_solve(const OrBsr* h, int solver, OrPc* pc, const double* b, double tol, int max_iter, int restart, double* x,
                 int* it) {
    return guard([&] {
        const uint32_t n = h->bs == 1 ? h->b1->nbr : h->b3->nbr * 3;
        vec bb(b, b + n), xx;
        MatVec mv = [h](const vec& u, vec& w) {
            if (h->bs == 1) h->b1->matvec(u, w); else h->b3->matvec(u, w);
        };
        if (solver == 0) *it = mini_cg(mv, bb, *pc->pc, tol, max_iter, xx);
        else *it = gmres(mv, n, bb, *pc->pc, restart, tol, max_iter, xx);
        std::copy(xx.begin(), xx.end(), x);
    });
}

}  // extern "C"

// precond_ssor.hip — multicolour SSOR preconditioner.
//   make_redblack_data        ode-solver/src/math/array.cpp:2935-3067
//   make_ssor_preconditioner  ode-solver/src/math/array.cpp:3068-3164
//
// Setup (host): colouring of pattern(A + A^T) by breadth-first 2-colouring
// from the lowest uncoloured seed, falling back to greedy multicolouring if
// the graph is not bipartite; rows renumbered stably by colour; A2 = P A P^T
// with columns ascending and explicit zeros dropped, split into its strictly
// lower part, strictly upper part and diagonal.
// Apply (device): one launch per colour and sweep. Rows of one colour are
// independent because the colour classes are independent sets:
//   forward   w_i  = (w(2-w) r[perm i] - sum_{j<i} (w a_ij) w_j) / d_i
//   backward  z2_i = w_i - (w / d_i) sum_{j>i} a_ij z2_j ;  z[perm i] = z2_i
// with the reference's operation order in every row.
#include <algorithm>
#include <cmath>

#include "internal.hpp"
#include "precond.hpp"

namespace fem {

template <typename T>
__global__ __launch_bounds__(256) void k_ssor_fw(uint32_t lo, uint32_t hi, const uint32_t* __restrict__ rp,
                                                 const uint32_t* __restrict__ ci, const T* __restrict__ v,
                                                 const T* __restrict__ d, const uint32_t* __restrict__ perm,
                                                 const T* __restrict__ r, T* w, T om, T fw, Gate g) {
    if (g.closed()) return;
    const uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    T s = fw * r[perm[i]];
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) s -= om * v[p] * w[ci[p]];
    w[i] = s / d[i];
}

template <typename T>
__global__ __launch_bounds__(256) void k_ssor_bw(uint32_t lo, uint32_t hi, const uint32_t* __restrict__ rp,
                                                 const uint32_t* __restrict__ ci, const T* __restrict__ v,
                                                 const T* __restrict__ d, const uint32_t* __restrict__ perm,
                                                 const T* __restrict__ w, T* z2, T* __restrict__ z, T om, Gate g) {
    if (g.closed()) return;
    const uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    T s = (T)0;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) s += v[p] * z2[ci[p]];
    const T zi = w[i] - (om / d[i]) * s;
    z2[i] = zi;
    z[perm[i]] = zi;
}

struct PcSSOR : Precond {
    uint32_t n, ncolor = 0;
    double omega, fw;
    std::vector<uint32_t> cptr;
    DevBuf<uint32_t> Lp, Lc, Up, Uc, perm;
    DevBuf<double> Lv, Uv, d, w, z2;
    DevBuf<float> Lv32, Uv32, d32, w32, z232;

    PcSSOR(CsrMat& A, double om, hipStream_t s) : n(A.n), omega(om), fw(om * (2.0 - om)) {
        if (A.n != A.m) fail(FEM_ERR_INVALID_ARGUMENT, "SSOR preconditioner needs a square matrix");
        if (!(om > 0.0 && om < 2.0)) fail(FEM_ERR_INVALID_ARGUMENT, "SSOR: the relaxation factor must satisfy 0 < omega < 2");
        A.sync_host();
        const auto& rp = A.h_rp;
        const auto& ci = A.h_ci;
        const auto& av = A.h_v;
        const uint64_t nnz = A.nnz;
        // transpose pattern
        std::vector<uint32_t> tp(n + 1, 0), tr(nnz);
        for (uint64_t p = 0; p < nnz; ++p) tp[ci[p] + 1]++;
        for (uint32_t i = 0; i < n; ++i) tp[i + 1] += tp[i];
        {
            std::vector<uint32_t> off(tp.begin(), tp.end() - 1);
            for (uint32_t i = 0; i < n; ++i)
                for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) tr[off[ci[p]]++] = i;
        }
        std::vector<int> color(n, -1);
        bool bip = true;
        std::vector<uint32_t> q;
        q.reserve(n);
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
                for (uint32_t p = rp[i]; p < rp[i + 1] && bip; ++p) visit(ci[p]);
                for (uint32_t p = tp[i]; p < tp[i + 1] && bip; ++p) visit(tr[p]);
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
                for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) mark(ci[p]);
                for (uint32_t p = tp[i]; p < tp[i + 1]; ++p) mark(tr[p]);
                int c = 0;
                while (used[c]) ++c;
                color[i] = c;
                for (int c2 : touched) used[c2] = 0;
            }
        }
        for (uint32_t i = 0; i < n; ++i) ncolor = std::max<uint32_t>(ncolor, (uint32_t)color[i] + 1);
        cptr.assign(ncolor + 1, 0);
        for (uint32_t i = 0; i < n; ++i) cptr[color[i] + 1]++;
        for (uint32_t c = 0; c < ncolor; ++c) cptr[c + 1] += cptr[c];
        std::vector<uint32_t> nid(n), pm(n);
        {
            std::vector<uint32_t> off(cptr.begin(), cptr.end() - 1);
            for (uint32_t i = 0; i < n; ++i) nid[i] = off[color[i]]++;
        }
        for (uint32_t i = 0; i < n; ++i) pm[nid[i]] = i;
        // A2 = P A P^T: permuted rows, renumbered columns sorted, zeros
        // dropped, duplicates summed (the COO constructor's semantics)
        std::vector<uint32_t> lp(n + 1, 0), up(n + 1, 0), lc, uc;
        std::vector<double> lv, uv, dg(n, 0.0);
        std::vector<std::pair<uint32_t, double>> row;
        for (uint32_t ni = 0; ni < n; ++ni) {
            const uint32_t i = pm[ni];
            row.clear();
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
                if (av[p] != 0.0) row.emplace_back(nid[ci[p]], av[p]);
            std::stable_sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
            bool found = false;
            for (size_t k = 0; k < row.size();) {
                const uint32_t j = row[k].first;
                double sum = row[k].second;
                size_t e = k + 1;
                while (e < row.size() && row[e].first == j) sum += row[e++].second;
                k = e;
                if (sum == 0.0) continue;
                if (j < ni) { lc.push_back(j); lv.push_back(sum); }
                else if (j > ni) { uc.push_back(j); uv.push_back(sum); }
                else { dg[ni] = sum; found = true; }
            }
            if (!found || std::fabs(dg[ni]) < 1e-300)
                fail(FEM_ERR_RUNTIME, "SSOR/red-black failed: missing nonzero diagonal entry");
            lp[ni + 1] = (uint32_t)lc.size();
            up[ni + 1] = (uint32_t)uc.size();
        }
        Lp.upload(lp, s); Lc.upload(lc.empty() ? std::vector<uint32_t>{0} : lc, s);
        Up.upload(up, s); Uc.upload(uc.empty() ? std::vector<uint32_t>{0} : uc, s);
        Lv.upload(lv.empty() ? std::vector<double>{0.0} : lv, s);
        Uv.upload(uv.empty() ? std::vector<double>{0.0} : uv, s);
        d.upload(dg, s);
        perm.upload(pm, s);
        w.alloc(n);
        z2.alloc(n);
        levels = (int)ncolor;
        FEM_HIP(hipStreamSynchronize(s));
    }
    int kind() const override { return FEM_PC_SSOR; }
    uint32_t size() const override { return n; }

    template <typename T>
    void run(const T* r, T* z, const T* lvv, const T* uvv, const T* dd, T* ww, T* zz2, hipStream_t s, Gate g) {
        const T om = (T)omega, f = (T)fw;
        for (uint32_t c = 0; c < ncolor; ++c) {
            const uint32_t lo = cptr[c], hi = cptr[c + 1];
            if (lo >= hi) continue;
            hipLaunchKernelGGL((k_ssor_fw<T>), dim3((hi - lo + 255) / 256), dim3(256), 0, s, lo, hi, Lp.get(), Lc.get(),
                               lvv, dd, perm.get(), r, ww, om, f, g);
        }
        for (int64_t c = (int64_t)ncolor - 1; c >= 0; --c) {
            const uint32_t lo = cptr[c], hi = cptr[c + 1];
            if (lo >= hi) continue;
            hipLaunchKernelGGL((k_ssor_bw<T>), dim3((hi - lo + 255) / 256), dim3(256), 0, s, lo, hi, Up.get(), Uc.get(),
                               uvv, dd, perm.get(), ww, zz2, z, om, g);
        }
        FEM_HIP(hipGetLastError());
    }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override {
        run<double>(r, z, Lv.get(), Uv.get(), d.get(), w.get(), z2.get(), s, g);
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override {
        run<float>(r, z, Lv32.get(), Uv32.get(), d32.get(), w32.get(), z232.get(), s, g);
    }
    bool has_fp32() const override { return w32.get() != nullptr; }
    void ensure_fp32(hipStream_t s) override {
        if (w32.get()) return;
        auto cv = [&](DevBuf<double>& a, DevBuf<float>& b) {
            b.alloc(a.n);
            vec_convert<float, double>(b.get(), a.get(), 1.0, a.n, s);
        };
        cv(Lv, Lv32);
        cv(Uv, Uv32);
        cv(d, d32);
        w32.alloc(n);
        z232.alloc(n);
    }
};

std::unique_ptr<Precond> make_ssor(CsrMat& A, double omega, hipStream_t s) {
    return std::unique_ptr<Precond>(new PcSSOR(A, omega, s));
}

}  // namespace fem

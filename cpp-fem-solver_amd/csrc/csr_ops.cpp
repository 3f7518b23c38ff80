// csr_ops.cpp — host-side CSR transpose and product used by setup paths.
//   SMatrix::transpose  ode-solver/src/math/array.cpp:2202-2235
//   SMatrix::operator*  ode-solver/src/math/array.cpp:2336-2447 (computed as
//   the exact product: every structurally reached column is kept)
#include "internal.hpp"

namespace fem {

void csr_transpose_host(const CsrMat& A, std::vector<uint32_t>& rp, std::vector<uint32_t>& ci, std::vector<double>& v) {
    A.sync_host();
    const uint32_t n = A.n, m = A.m;
    rp.assign(m + 1, 0);
    for (uint64_t p = 0; p < A.nnz; ++p) rp[A.h_ci[p] + 1]++;
    for (uint32_t j = 0; j < m; ++j) rp[j + 1] += rp[j];
    ci.resize(A.nnz);
    v.resize(A.nnz);
    std::vector<uint32_t> off(rp.begin(), rp.end() - 1);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t p = A.h_rp[i]; p < A.h_rp[i + 1]; ++p) {
            const uint32_t q = off[A.h_ci[p]]++;
            ci[q] = i;
            v[q] = A.h_v[p];
        }
}

void csr_spgemm_host(const CsrMat& A, const CsrMat& B, std::vector<uint32_t>& rp, std::vector<uint32_t>& ci,
                     std::vector<double>& v) {
    A.sync_host();
    B.sync_host();
    rp.assign(A.n + 1, 0);
    ci.clear();
    v.clear();
    std::vector<int64_t> mark(B.m, -1);
    std::vector<double> acc(B.m, 0.0);
    std::vector<uint32_t> cols;
    for (uint32_t i = 0; i < A.n; ++i) {
        cols.clear();
        for (uint32_t p = A.h_rp[i]; p < A.h_rp[i + 1]; ++p) {
            const uint32_t j = A.h_ci[p];
            const double a = A.h_v[p];
            for (uint32_t q = B.h_rp[j]; q < B.h_rp[j + 1]; ++q) {
                const uint32_t k = B.h_ci[q];
                if (mark[k] != (int64_t)i) { mark[k] = i; cols.push_back(k); acc[k] = 0.0; }
                acc[k] += a * B.h_v[q];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (uint32_t k : cols) { ci.push_back(k); v.push_back(acc[k]); }
        rp[i + 1] = (uint32_t)ci.size();
    }
}

}  // namespace fem

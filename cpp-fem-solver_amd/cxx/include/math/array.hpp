// array.hpp — drop-in replacement of the sparse part of the reference header
// ode-solver/include/math/array.hpp (namespace sparse, class SMatrix,
// array.hpp:264-373), backed by libfemgpu on an MI355X.
//
// Same public declarations, argument meanings, defaults and exception types
// as the reference. The matrix keeps its CSR arrays on the host (the accessors
// return them) and caches one device copy, created on first use; an SMatrix
// is immutable after construction, so the cache stays valid until the object
// is assigned. Dense array::Matrix/Tensor and the stationary sor/sor_rb/amg
// solvers of the reference header are not part of the GPU solve path and are
// not declared here.
#pragma once
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

struct fem_csr_s;

namespace sparse {

class SMatrix {
   private:
    uint32_t rows_, cols_;
    std::vector<uint32_t> row_ptr_;  // length rows_ + 1
    std::vector<uint32_t> col_idx_;  // length nnz
    std::vector<double> values_;     // nonzero values
    mutable std::shared_ptr<fem_csr_s> dev_;

   public:
    SMatrix();
    // COO constructor: drop zeros, sort (i,j), merge duplicates, drop merged zeros
    SMatrix(uint32_t rows, uint32_t cols, const std::vector<double>& values, const std::vector<uint32_t>& rowlist,
            const std::vector<uint32_t>& collist);
    // CSR constructor (no checks)
    SMatrix(uint32_t rows, uint32_t cols, std::vector<uint32_t> row_ptr, std::vector<uint32_t> col_idx,
            std::vector<double> values);

    static double dot(const std::vector<double>& a, const std::vector<double>& b);

    uint32_t rows() const;
    uint32_t cols() const;
    uint32_t nnz() const;

    const std::vector<double>& values() const;
    const std::vector<uint32_t>& row_ptr() const;
    const std::vector<uint32_t>& col_idx() const;

    SMatrix transpose() const;

    SMatrix operator*(const SMatrix& B) const;

    // preconditioner factories: z = M^{-1} r
    static std::function<void(const std::vector<double>&, std::vector<double>&)> make_diag_preconditioner(const SMatrix& A);
    static std::function<void(const std::vector<double>&, std::vector<double>&)> make_ic0_preconditioner(const SMatrix& A);
    static std::function<void(const std::vector<double>&, std::vector<double>&)> make_ilu0_preconditioner(const SMatrix& A);
    static std::function<void(const std::vector<double>&, std::vector<double>&)> make_ssor_preconditioner(const SMatrix& A, double omega = 1.2);
    static std::function<void(const std::vector<double>&, std::vector<double>&)> make_amg_preconditioner(const SMatrix& A);

    static std::vector<double> spmv(const SMatrix& A, const std::vector<double>& x);

    static std::tuple<std::vector<double>, int> pcg(const SMatrix& A, const std::vector<double>& b,
                                                    const std::string& precond = "none", double tol = 1e-10,
                                                    int max_iter = 2000);
    static std::tuple<std::vector<double>, int> bicgstab(const SMatrix& A, const std::vector<double>& b,
                                                         const std::string& precond, double tol = 1e-10,
                                                         int max_iter = 2000);
    static std::tuple<std::vector<double>, int> gmres(const SMatrix& A, const std::vector<double>& b,
                                                      const std::string& precond, int restart, double tol = 1e-10,
                                                      int max_iter = 2000);
    static std::tuple<std::vector<double>, int> gmres(
        const SMatrix& A, const std::vector<double>& b,
        const std::function<void(const std::vector<double>&, std::vector<double>&)>& solveM, int restart,
        double tol = 1e-10, int max_iter = 2000);

    // device handle of this matrix (created on first use); used by the
    // block facade and by callers that mix this API with the C ABI
    fem_csr_s* device() const;
};

}  // namespace sparse

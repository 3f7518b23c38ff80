// smatrix.cpp — sparse::SMatrix over the libfemgpu C ABI.
// Check order, early exits, alias tables and exception types follow
// ode-solver/src/math/array.cpp (pcg 3445-3601, bicgstab 3602-3842,
// gmres 4190-4533, to_lower 2238-2250).
#include <algorithm>
#include <cctype>
#include <cmath>

#include "femgpu.h"
#include "math/array.hpp"
#include "math/femgpu_errors.hpp"

namespace sparse {

namespace {

std::string to_lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

struct CsrDel {
    void operator()(fem_csr_s* h) const {
        if (h) fem_csr_destroy(h);
    }
};
struct PcDel {
    void operator()(fem_pc_s* h) const {
        if (h) fem_precond_destroy(h);
    }
};

using PcPtr = std::shared_ptr<fem_pc_s>;

// solver: 0 pcg, 1 bicgstab, 2 gmres (array.cpp:3484-3511, 3641-3669, 4215-4238)
int alias(const std::string& precond, int solver) {
    const std::string s = to_lower(precond);
    if (s == "none" || s == "identity" || s == "noprecond") return FEM_PC_NONE;
    if (solver == 0 && s == "cg") return FEM_PC_NONE;
    if (s == "diag" || s == "diagonal" || s == "jacobi") return FEM_PC_DIAG;
    if (s == "ic" || s == "ic0" || s == "ichol") return FEM_PC_IC0;
    if (solver != 2 && s == "incomplete_cholesky") return FEM_PC_IC0;
    if (s == "ilu" || s == "ilu0") return FEM_PC_ILU0;
    if (s == "sor" || s == "ssor") return FEM_PC_SSOR;
    if (s == "amg") return FEM_PC_AMG;
    throw std::invalid_argument("unknown preconditioner type: " + precond);
}

PcPtr make_pc(const SMatrix& A, int kind, double omega = 1.2) {
    fem_pc_params p;
    fem_pc_params_default(kind, &p);
    if (kind == FEM_PC_SSOR) p.omega = omega;
    fem_pc_s* h = nullptr;
    femgpu::check(fem_precond_create(A.device(), kind, &p, &h));
    return PcPtr(h, PcDel());
}

std::function<void(const std::vector<double>&, std::vector<double>&)> wrap(PcPtr pc, uint32_t n) {
    femgpu::DevicePrecond d;
    d.pc = std::move(pc);
    d.n = n;
    return d;
}

std::tuple<std::vector<double>, int> run(const SMatrix& A, const std::vector<double>& b, int solver,
                                         const std::string& precond, int restart, double tol, int max_iter,
                                         const char* name) {
    if (A.rows() != A.cols()) throw std::invalid_argument(std::string(name) + " needs a square matrix");
    if (A.cols() != b.size())
        throw std::invalid_argument(std::string(name) + ": matrix columns differ from the vector length");
    const uint32_t n = (uint32_t)b.size();
    std::vector<double> x(n, 0.0);
    const double normb = std::sqrt(SMatrix::dot(b, b));
    if (normb < 1e-300) return {x, 0};
    // pcg/bicgstab test ||r0||/||b|| before resolving the preconditioner
    if (solver != 2 && normb / normb < tol) return {x, 0};
    const int kind = alias(precond, solver);
    PcPtr pc = make_pc(A, kind, 1.2);
    int it = 0;
    if (solver == 0) femgpu::check(fem_pcg(A.device(), pc.get(), b.data(), tol, max_iter, x.data(), &it));
    else if (solver == 1) femgpu::check(fem_bicgstab(A.device(), pc.get(), b.data(), tol, max_iter, x.data(), &it));
    else femgpu::check(fem_gmres(A.device(), pc.get(), b.data(), restart, tol, max_iter, x.data(), &it));
    return {x, it};
}

}  // namespace

SMatrix::SMatrix() : rows_(0), cols_(0), row_ptr_{0} {}

SMatrix::SMatrix(uint32_t rows, uint32_t cols, const std::vector<double>& values, const std::vector<uint32_t>& rowlist,
                 const std::vector<uint32_t>& collist)
    : rows_(rows), cols_(cols) {
    if (values.size() != rowlist.size() || values.size() != collist.size())
        throw std::invalid_argument("COO input: values, rowlist and collist lengths differ");
    fem_csr_s* h = nullptr;
    femgpu::check(fem_csr_create_coo(rows, cols, values.size(), values.data(), rowlist.data(), collist.data(), &h));
    dev_.reset(h, CsrDel());
    uint64_t nnz = 0;
    femgpu::check(fem_csr_info(h, nullptr, nullptr, &nnz));
    row_ptr_.resize(rows + 1);
    col_idx_.resize(nnz);
    values_.resize(nnz);
    femgpu::check(fem_csr_export(h, row_ptr_.data(), col_idx_.data(), values_.data()));
}

SMatrix::SMatrix(uint32_t rows, uint32_t cols, std::vector<uint32_t> row_ptr, std::vector<uint32_t> col_idx,
                 std::vector<double> values)
    : rows_(rows), cols_(cols), row_ptr_(std::move(row_ptr)), col_idx_(std::move(col_idx)), values_(std::move(values)) {}

uint32_t SMatrix::rows() const { return rows_; }
uint32_t SMatrix::cols() const { return cols_; }
uint32_t SMatrix::nnz() const { return (uint32_t)values_.size(); }
const std::vector<double>& SMatrix::values() const { return values_; }
const std::vector<uint32_t>& SMatrix::row_ptr() const { return row_ptr_; }
const std::vector<uint32_t>& SMatrix::col_idx() const { return col_idx_; }

fem_csr_s* SMatrix::device() const {
    if (!dev_) {
        fem_csr_s* h = nullptr;
        femgpu::check(fem_csr_create(rows_, cols_, row_ptr_.data(), col_idx_.data(), values_.data(), &h));
        dev_.reset(h, CsrDel());
    }
    return dev_.get();
}

static SMatrix from_handle(fem_csr_s* h) {
    std::shared_ptr<fem_csr_s> hold(h, CsrDel());
    uint32_t n = 0, m = 0;
    uint64_t nnz = 0;
    femgpu::check(fem_csr_info(h, &n, &m, &nnz));
    std::vector<uint32_t> rp(n + 1), ci(nnz);
    std::vector<double> v(nnz);
    femgpu::check(fem_csr_export(h, rp.data(), ci.data(), v.data()));
    return SMatrix(n, m, std::move(rp), std::move(ci), std::move(v));
}

SMatrix SMatrix::transpose() const {
    fem_csr_s* h = nullptr;
    femgpu::check(fem_csr_transpose(device(), &h));
    return from_handle(h);
}

SMatrix SMatrix::operator*(const SMatrix& B) const {
    fem_csr_s* h = nullptr;
    femgpu::check(fem_csr_multiply(device(), B.device(), &h));
    return from_handle(h);
}

double SMatrix::dot(const std::vector<double>& a, const std::vector<double>& b) {
    if (a.size() != b.size()) throw std::invalid_argument("dot: vector lengths differ");
    double out = 0.0;
    femgpu::check(fem_dot(a.size(), a.data(), b.data(), &out));
    return out;
}

std::vector<double> SMatrix::spmv(const SMatrix& A, const std::vector<double>& x) {
    if (A.cols_ != x.size()) throw std::invalid_argument("spmv: matrix columns differ from the vector length");
    std::vector<double> y(A.rows_, 0.0);
    if (A.rows_ == 0) return y;
    femgpu::check(fem_spmv(A.device(), x.data(), y.data()));
    return y;
}

std::function<void(const std::vector<double>&, std::vector<double>&)> SMatrix::make_diag_preconditioner(const SMatrix& A) {
    if (A.rows() != A.cols()) throw std::invalid_argument("diag preconditioner needs a square matrix");
    return wrap(make_pc(A, FEM_PC_DIAG), A.rows());
}
std::function<void(const std::vector<double>&, std::vector<double>&)> SMatrix::make_ic0_preconditioner(const SMatrix& A) {
    if (A.rows() != A.cols()) throw std::invalid_argument("IC(0) preconditioner needs a square matrix");
    return wrap(make_pc(A, FEM_PC_IC0), A.rows());
}
std::function<void(const std::vector<double>&, std::vector<double>&)> SMatrix::make_ilu0_preconditioner(const SMatrix& A) {
    if (A.rows() != A.cols()) throw std::invalid_argument("ILU(0) preconditioner needs a square matrix");
    return wrap(make_pc(A, FEM_PC_ILU0), A.rows());
}
std::function<void(const std::vector<double>&, std::vector<double>&)> SMatrix::make_ssor_preconditioner(const SMatrix& A,
                                                                                                      double omega) {
    if (A.rows() != A.cols()) throw std::invalid_argument("SSOR preconditioner needs a square matrix");
    if (!(omega > 0.0 && omega < 2.0)) throw std::invalid_argument("SSOR: the relaxation factor must satisfy 0 < omega < 2");
    return wrap(make_pc(A, FEM_PC_SSOR, omega), A.rows());
}
std::function<void(const std::vector<double>&, std::vector<double>&)> SMatrix::make_amg_preconditioner(const SMatrix& A) {
    return wrap(make_pc(A, FEM_PC_AMG), A.rows());
}

std::tuple<std::vector<double>, int> SMatrix::pcg(const SMatrix& A, const std::vector<double>& b,
                                                  const std::string& precond, double tol, int max_iter) {
    return run(A, b, 0, precond, 0, tol, max_iter, "PCG");
}
std::tuple<std::vector<double>, int> SMatrix::bicgstab(const SMatrix& A, const std::vector<double>& b,
                                                       const std::string& precond, double tol, int max_iter) {
    return run(A, b, 1, precond, 0, tol, max_iter, "BiCGSTAB");
}
std::tuple<std::vector<double>, int> SMatrix::gmres(const SMatrix& A, const std::vector<double>& b,
                                                    const std::string& precond, int restart, double tol, int max_iter) {
    return run(A, b, 2, precond, restart, tol, max_iter, "GMRES");
}

namespace {
struct CbCtx {
    const std::function<void(const std::vector<double>&, std::vector<double>&)>* fn;
    std::vector<double> r, z;
};
void trampoline(const double* r, double* z, uint64_t n, void* user) {
    auto* c = static_cast<CbCtx*>(user);
    c->r.assign(r, r + n);
    (*c->fn)(c->r, c->z);
    std::copy(c->z.begin(), c->z.begin() + std::min<size_t>(n, c->z.size()), z);
}
}  // namespace

std::tuple<std::vector<double>, int> SMatrix::gmres(
    const SMatrix& A, const std::vector<double>& b,
    const std::function<void(const std::vector<double>&, std::vector<double>&)>& solveM, int restart, double tol,
    int max_iter) {
    const size_t n = b.size();
    std::vector<double> x(n, 0.0);
    const double normb = std::sqrt(SMatrix::dot(b, b));
    if (normb < 1e-300) return {x, 0};
    int it = 0;
    // a preconditioner made by this library runs on the device: the solve
    // needs no host call per step (same algorithm, array.cpp:4402-4533)
    if (const auto* dp = solveM.target<femgpu::DevicePrecond>()) {
        if (dp->n != n) throw std::invalid_argument("GMRES: preconditioner size differs from the vector length");
        femgpu::check(fem_gmres(A.device(), dp->pc.get(), b.data(), restart, tol, max_iter, x.data(), &it));
        return {x, it};
    }
    CbCtx ctx{&solveM, {}, {}};
    femgpu::check(fem_gmres_callback(A.device(), trampoline, &ctx, b.data(), restart, tol, max_iter, x.data(), &it));
    return {x, it};
}

}  // namespace sparse

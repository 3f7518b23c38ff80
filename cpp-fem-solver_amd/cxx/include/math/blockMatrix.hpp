// blockMatrix.hpp — drop-in replacement of ode-solver/include/math/blockMatrix.hpp
// (namespace block: BlockMatrix<T,BS>, block ILU(0), block SA-AMG, Schur
// preconditioner, BlockMatrix::gmres), backed by libfemgpu.
//
// Same layout and public members as the reference (blockMatrix.hpp:121-451):
// row-major BS x BS blocks stored contiguously, node-major vectors. The
// compute calls (matvec, preconditioners, gmres) run on the GPU through a
// cached device copy. T must be double (the library's boundary type). The
// MPI halo members (set_halo, distributed_matvec) are not provided: one GPU
// runs one rank's local block rows.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "femgpu.h"
#include "math/array.hpp"
#include "math/femgpu_errors.hpp"

using namespace sparse;

namespace block {

namespace detail {
struct BsrDel {
    void operator()(fem_bsr_s* h) const {
        if (h) fem_bsr_destroy(h);
    }
};
struct PcDel {
    void operator()(fem_pc_s* h) const {
        if (h) fem_precond_destroy(h);
    }
};
}  // namespace detail

template <typename T, int BS>
class BlockMatrix {
    static_assert(std::is_same<T, double>::value, "the GPU block matrix stores double values");
    static_assert(BS >= 1 && BS <= 4, "block size must be 1..4");

   private:
    uint32_t nbr_, nbc_;
    std::vector<uint32_t> row_ptr_;
    std::vector<uint32_t> col_idx_;
    std::vector<T> values_;
    std::vector<uint32_t> diag_ptr_;
    mutable std::shared_ptr<fem_bsr_s> dev_;

   public:
    static constexpr int kBlock = BS;
    static constexpr int kBlockVals = BS * BS;

    BlockMatrix() : nbr_(0), nbc_(0), row_ptr_{0} {}
    BlockMatrix(uint32_t nbr, uint32_t nbc, std::vector<uint32_t> row_ptr, std::vector<uint32_t> col_idx,
                std::vector<T> values)
        : nbr_(nbr), nbc_(nbc), row_ptr_(std::move(row_ptr)), col_idx_(std::move(col_idx)), values_(std::move(values)) {
        cache_diag();
    }

    uint32_t block_rows() const { return nbr_; }
    uint32_t block_cols() const { return nbc_; }
    uint32_t rows() const { return nbr_ * BS; }
    uint32_t cols() const { return nbc_ * BS; }
    size_t nnz_blocks() const { return col_idx_.size(); }
    size_t nnz() const { return col_idx_.size() * (size_t)kBlockVals; }

    const std::vector<uint32_t>& block_row_ptr() const { return row_ptr_; }
    const std::vector<uint32_t>& block_col_idx() const { return col_idx_; }
    const std::vector<T>& values() const { return values_; }

    const T* block(uint32_t p) const { return &values_[(size_t)p * kBlockVals]; }
    const T* diag_block(uint32_t bi) const { return &values_[(size_t)diag_ptr_[bi] * kBlockVals]; }

    void cache_diag() {
        diag_ptr_.assign(nbr_, UINT32_MAX);
        for (uint32_t bi = 0; bi < nbr_; ++bi)
            for (uint32_t p = row_ptr_[bi]; p < row_ptr_[bi + 1]; ++p)
                if (col_idx_[p] == bi) { diag_ptr_[bi] = p; break; }
    }

    fem_bsr_s* device() const {
        if (!dev_) {
            fem_bsr_s* h = nullptr;
            femgpu::check(fem_bsr_create(BS, nbr_, nbc_, row_ptr_.data(), col_idx_.data(), values_.data(), &h));
            dev_.reset(h, detail::BsrDel());
        }
        return dev_.get();
    }

    // y = A x (blockMatrix.hpp:208-234)
    void matvec(const std::vector<T>& x, std::vector<T>& y) const {
        if (x.size() != cols()) throw std::invalid_argument("matvec: vector length differs from the block matrix");
        y.assign(rows(), T(0));
        if (!rows()) return;
        femgpu::check(fem_bsr_matvec(device(), x.data(), y.data()));
    }

    BlockMatrix<T, BS> transpose() const {
        fem_bsr_s* h = nullptr;
        femgpu::check(fem_bsr_transpose(device(), &h));
        return from_device(h);
    }

    static BlockMatrix<T, BS> condense(const SMatrix& A, uint32_t nbr, uint32_t nbc) {
        fem_bsr_s* h = nullptr;
        femgpu::check(fem_bsr_condense(A.device(), BS, nbr, nbc, &h));
        return from_device(h);
    }

    static std::tuple<std::vector<T>, int> gmres(const BlockMatrix& A, const std::vector<T>& b,
                                                 const std::string& precond, double tol, int max_iter, int restart);

   private:
    static BlockMatrix<T, BS> from_device(fem_bsr_s* h) {
        std::shared_ptr<fem_bsr_s> hold(h, detail::BsrDel());
        int bs = 0;
        uint32_t nbr = 0, nbc = 0;
        uint64_t nz = 0;
        femgpu::check(fem_bsr_info(h, &bs, &nbr, &nbc, &nz));
        std::vector<uint32_t> rp(nbr + 1), ci(nz);
        std::vector<T> v(nz * kBlockVals);
        femgpu::check(fem_bsr_export(h, rp.data(), ci.data(), v.data()));
        BlockMatrix<T, BS> M(nbr, nbc, std::move(rp), std::move(ci), std::move(v));
        M.dev_ = hold;
        return M;
    }
};

namespace detail {
template <typename T, int BS>
std::function<void(const std::vector<T>&, std::vector<T>&)> block_pc(const BlockMatrix<T, BS>& A, int kind,
                                                                     const fem_pc_params* p = nullptr) {
    fem_pc_s* h = nullptr;
    femgpu::check(fem_bsr_precond_create(A.device(), kind, p, &h));
    femgpu::DevicePrecond d;
    d.pc = std::shared_ptr<fem_pc_s>(h, PcDel());
    d.n = A.rows();
    return d;
}
}  // namespace detail

// make_block_ilu0_preconditioner (blockMatrix.hpp:455-623)
template <typename T, int BS>
std::function<void(const std::vector<T>&, std::vector<T>&)> make_block_ilu0_preconditioner(const BlockMatrix<T, BS>& A) {
    if (A.block_rows() != A.block_cols()) throw std::invalid_argument("Block-ILU(0) needs a square block matrix");
    return detail::block_pc(A, FEM_PC_BILU0);
}

// make_block_amg_preconditioner (blockMatrix.hpp:1051-1179), reference
// parameters: theta 0.25, omega 2/3, one block-Jacobi sweep, aggregates <= 4,
// coarsest <= 50 block rows
template <typename T, int BS>
std::function<void(const std::vector<T>&, std::vector<T>&)> make_block_amg_preconditioner(const BlockMatrix<T, BS>& A) {
    if (A.block_rows() != A.block_cols()) throw std::invalid_argument("Block AMG needs a square matrix");
    fem_pc_params p;
    fem_pc_params_default(FEM_PC_BSAAMG, &p);
    return detail::block_pc(A, FEM_PC_BSAAMG, &p);
}

// BlockMatrix::gmres (blockMatrix.hpp:688-866); precond names are
// case-sensitive: "none", "identity", "blockjacobi", "blockilu", "blockilu0"
template <typename T, int BS>
std::tuple<std::vector<T>, int> BlockMatrix<T, BS>::gmres(const BlockMatrix& A, const std::vector<T>& b,
                                                          const std::string& precond, double tol, int max_iter,
                                                          int restart) {
    const size_t n = b.size();
    std::vector<T> x(n, T(0));
    T nb = T(0);
    for (size_t i = 0; i < n; ++i) nb += b[i] * b[i];
    if (std::sqrt(nb) < 1e-300) return {x, 0};
    int kind;
    if (precond == "none" || precond == "identity") kind = FEM_PC_NONE;
    else if (precond == "blockjacobi") kind = FEM_PC_BJACOBI;
    else if (precond == "blockilu" || precond == "blockilu0") kind = FEM_PC_BILU0;
    else throw std::invalid_argument("BlockMatrix::gmres unknown preconditioner type: " + precond);
    if (n != A.rows()) throw std::invalid_argument("BlockMatrix::gmres: vector length differs from the matrix");
    fem_pc_s* h = nullptr;
    femgpu::check(fem_bsr_precond_create(A.device(), kind, nullptr, &h));
    std::shared_ptr<fem_pc_s> pc(h, detail::PcDel());
    fem_solve_opts o;
    fem_solve_opts_default(FEM_SOLVER_GMRES, &o);
    o.tol = tol;
    o.max_iter = max_iter;
    o.restart = restart;
    fem_solve_stats st;
    femgpu::check(fem_bsr_solve(A.device(), pc.get(), b.data(), x.data(), &o, &st));
    return {x, st.iterations};
}

// Saddle-point system and block upper-triangular Schur preconditioner
// (blockMatrix.hpp:625-685): host-side composition of the GPU operators.
struct SaddlePointSystem {
    size_t n_vel;
    size_t n_pres;
    std::function<void(const std::vector<double>&, std::vector<double>&)> spmv_A;
    std::function<void(const std::vector<double>&, std::vector<double>&)> spmv_B;
    std::function<void(const std::vector<double>&, std::vector<double>&)> spmv_Bt;
};

inline std::function<void(const std::vector<double>&, std::vector<double>&)> make_schur_preconditioner(
    const SaddlePointSystem& sys, std::function<void(const std::vector<double>&, std::vector<double>&)> A_precond,
    std::function<void(const std::vector<double>&, std::vector<double>&)> S_precond) {
    struct Work {
        std::vector<double> r_u, r_p, w_u, w_p, r_p_mod, z_p, w_bt, r_u_mod, z_u;
    };
    auto wk = std::make_shared<Work>();
    return [sys, A_precond, S_precond, wk](const std::vector<double>& r, std::vector<double>& z) {
        const size_t nv = sys.n_vel, np = sys.n_pres;
        Work& w = *wk;
        w.r_u.assign(r.begin(), r.begin() + nv);
        w.r_p.assign(r.begin() + nv, r.begin() + nv + np);
        A_precond(w.r_u, w.w_u);                       // w_u = A^{-1} r_u
        sys.spmv_B(w.w_u, w.w_p);                      // w_p = B w_u
        w.r_p_mod.resize(np);
        for (size_t i = 0; i < np; ++i) w.r_p_mod[i] = w.w_p[i] - w.r_p[i];
        S_precond(w.r_p_mod, w.z_p);                   // z_p = S^{-1}(B w_u - r_p)
        sys.spmv_Bt(w.z_p, w.w_bt);                    // w_bt = B^T z_p
        w.r_u_mod.resize(nv);
        for (size_t i = 0; i < nv; ++i) w.r_u_mod[i] = w.r_u[i] - w.w_bt[i];
        A_precond(w.r_u_mod, w.z_u);                   // z_u = A^{-1}(r_u - B^T z_p)
        z.resize(nv + np);
        std::copy(w.z_u.begin(), w.z_u.end(), z.begin());
        std::copy(w.z_p.begin(), w.z_p.end(), z.begin() + nv);
    };
}

// The same preconditioner composed on the device (fem_precond_create_schur;
// no counterpart in the reference): from the matrices B (n_pres x n_vel) and
// B^T and preconditioners made by this library's factories (S_precond may
// be empty: the identity), with S_precond's output multiplied by s_scale (the
// reference test's S^{-1} ~ 6.1 L_p^{-1}). SMatrix::gmres(K, b, M, restart)
// then runs the whole solve on the device.
inline std::function<void(const std::vector<double>&, std::vector<double>&)> make_schur_preconditioner(
    const sparse::SMatrix& B, const sparse::SMatrix& Bt,
    const std::function<void(const std::vector<double>&, std::vector<double>&)>& A_precond,
    const std::function<void(const std::vector<double>&, std::vector<double>&)>& S_precond, double s_scale = 1.0) {
    const auto* a = A_precond.target<femgpu::DevicePrecond>();
    const auto* sp = S_precond ? S_precond.target<femgpu::DevicePrecond>() : nullptr;
    if (!a || (S_precond && !sp))
        throw std::invalid_argument("device Schur preconditioner: A_precond and S_precond must come from this library");
    fem_pc_s* h = nullptr;
    femgpu::check(fem_precond_create_schur(B.device(), Bt.device(), a->pc.get(), sp ? sp->pc.get() : nullptr, s_scale,
                                           &h));
    femgpu::DevicePrecond d;
    d.pc = std::shared_ptr<fem_pc_s>(h, detail::PcDel());
    d.n = B.rows() + B.cols();
    // keep what the device preconditioner points into alive with it
    d.keep = {std::make_shared<sparse::SMatrix>(B), std::make_shared<sparse::SMatrix>(Bt), a->pc};
    if (sp) d.keep.push_back(sp->pc);
    return d;
}

}  // namespace block

// precond.hpp — preconditioner constructors shared across translation units.
#pragma once
#include "internal.hpp"

namespace fem {

void extract_diag(const CsrMat& A, double* d, hipStream_t s);
template <typename T>
void launch_copy(const T* r, T* z, size_t n, hipStream_t s, Gate g);

std::unique_ptr<Precond> make_ic0(CsrMat& A, hipStream_t s);
std::unique_ptr<Precond> make_ilu0(CsrMat& A, hipStream_t s);
std::unique_ptr<Precond> make_ssor(CsrMat& A, double omega, hipStream_t s);
std::unique_ptr<Precond> make_amg_cf(CsrMat& A, const fem_pc_params& p, hipStream_t s);
std::unique_ptr<Precond> make_saamg(CsrMat& A, const fem_pc_params& p, hipStream_t s);
// Schur-complement preconditioner over B, B^T and two preconditioners (S_pc
// may be null); it keeps pointers to all four
std::unique_ptr<Precond> make_schur(const CsrMat& B, const CsrMat& Bt, Precond* A_pc, Precond* S_pc, double scale);

}  // namespace fem

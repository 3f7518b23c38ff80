// handles.hpp — definitions of the opaque C-ABI handle types.
#pragma once
#include "internal.hpp"
#include "solver.hpp"

namespace fem {
struct BsrMat;
hipStream_t new_stream();
}

struct fem_csr_s {
    hipStream_t stream = nullptr;
    fem::CsrMat A;
    fem::RedScratch red;
    fem::DevBuf<double> xs, ys;
    ~fem_csr_s() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct fem_pc_s {
    std::unique_ptr<fem::Precond> pc;
    hipStream_t stream = nullptr;  // borrowed from the matrix handle
    fem::DevBuf<double> r, z;
};

struct fem_solver_s {
    std::unique_ptr<fem::KrylovSolver> S;
    fem_csr A = nullptr;
    fem::LinOp* op = nullptr;
    hipStream_t stream = nullptr;
    fem::DevBuf<double> b, x;
};

void fem_set_error(const std::string& m);

// C-ABI wrapper: run f, map exceptions to status codes and fem_last_error()
template <class F>
inline int api_guard(F&& f) {
    try {
        f();
        return FEM_OK;
    } catch (const fem::Error& e) {
        fem_set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        fem_set_error("out of memory");
        return FEM_ERR_DEVICE;
    } catch (const std::exception& e) {
        fem_set_error(e.what());
        return FEM_ERR_RUNTIME;
    }
}

// run a solve through a solver handle with host or device pointers
void fem_solver_run(fem_solver_s* h, const double* b, double* x, fem_solve_stats* st);
void fem_solver_make(fem::LinOp* op, hipStream_t stream, fem_pc M, const fem_solve_opts* opts, fem_solver_s* h);

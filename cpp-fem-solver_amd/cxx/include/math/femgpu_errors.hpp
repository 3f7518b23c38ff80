// femgpu_errors.hpp — maps libfemgpu status codes onto the reference's
// exception types (std::invalid_argument / std::out_of_range /
// std::runtime_error, ode-solver/src/math/array.cpp), and the callable the
// facade's preconditioner factories return.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "femgpu.h"

namespace femgpu {

inline void check(int rc) {
    if (rc == FEM_OK) return;
    const std::string m = fem_last_error();
    switch (rc) {
        case FEM_ERR_INVALID_ARGUMENT: throw std::invalid_argument(m);
        case FEM_ERR_OUT_OF_RANGE: throw std::out_of_range(m);
        default: throw std::runtime_error(m);
    }
}

// z = M^{-1} r through a device preconditioner. The facade's factories wrap
// this in the std::function the reference returns; code that recognises it
// (std::function::target<DevicePrecond>()) passes the device handle on
// instead of calling back to the host: SMatrix::gmres with a function
// preconditioner, and the device form of make_schur_preconditioner.
struct DevicePrecond {
    std::shared_ptr<fem_pc_s> pc;
    uint32_t n = 0;
    std::vector<std::shared_ptr<const void>> keep;  // what pc points into
    void operator()(const std::vector<double>& r, std::vector<double>& z) const {
        if (r.size() != n) throw std::invalid_argument("preconditioner: vector length differs from the matrix");
        z.resize(n);
        check(fem_precond_apply(pc.get(), r.data(), z.data()));
    }
};

}  // namespace femgpu

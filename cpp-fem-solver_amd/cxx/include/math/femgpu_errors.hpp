// femgpu_errors.hpp — maps libfemgpu status codes onto the reference's
// exception types (std::invalid_argument / std::out_of_range /
// std::runtime_error, ode-solver/src/math/array.cpp).
#pragma once
#include <stdexcept>
#include <string>

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

}  // namespace femgpu

// precond_basic.hip — identity and diagonal (Jacobi) preconditioners and the
// preconditioner factory.
//   make_diag_preconditioner  ode-solver/src/math/array.cpp:2450-2508
#include <cmath>

#include "internal.hpp"
#include "precond.hpp"

namespace fem {

template <typename T>
__global__ void k_copy(const T* r, T* z, size_t n, Gate g) {
    if (g.closed()) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) z[i] = r[i];
}

template <typename T>
__global__ void k_div(const T* r, const T* d, T* z, size_t n, Gate g) {
    if (g.closed()) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        z[i] = r[i] / d[i];
}

// d[i] = first stored A(i,i) of row i, 0 if absent (array.cpp:2466-2480)
__global__ void k_extract_diag(const uint32_t* rp, const uint32_t* ci, const double* v, uint32_t n, double* d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = 0.0;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] == i) { x = v[p]; break; }
    d[i] = x;
}

void extract_diag(const CsrMat& A, double* d, hipStream_t s) {
    const uint32_t n = std::min(A.n, A.m);
    if (!n) return;
    hipLaunchKernelGGL(k_extract_diag, dim3((n + 255) / 256), dim3(256), 0, s, A.rp.get(), A.ci.get(), A.v.get(), n, d);
    FEM_HIP(hipGetLastError());
}

static unsigned ew_grid(size_t n) {
    size_t b = (n + 1023) / 1024;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(b, 8192));
}

template <typename T>
void launch_copy(const T* r, T* z, size_t n, hipStream_t s, Gate g) {
    hipLaunchKernelGGL((k_copy<T>), dim3(ew_grid(n)), dim3(256), 0, s, r, z, n, g);
    FEM_HIP(hipGetLastError());
}
template void launch_copy<double>(const double*, double*, size_t, hipStream_t, Gate);
template void launch_copy<float>(const float*, float*, size_t, hipStream_t, Gate);

struct PcNone : Precond {
    uint32_t n;
    explicit PcNone(uint32_t n_) : n(n_) {}
    int kind() const override { return FEM_PC_NONE; }
    uint32_t size() const override { return n; }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override { launch_copy(r, z, n, s, g); }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override { launch_copy(r, z, n, s, g); }
    bool has_fp32() const override { return true; }
};

struct PcDiag : Precond {
    uint32_t n;
    DevBuf<double> d;
    DevBuf<float> d32;
    PcDiag(const CsrMat& A, hipStream_t s) : n(A.n) {
        if (A.n != A.m) fail(FEM_ERR_INVALID_ARGUMENT, "diag preconditioner needs a square matrix");
        d.alloc(n);
        extract_diag(A, d.get(), s);
        std::vector<double> h(n);
        d.download(h.data(), n, s);
        FEM_HIP(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < n; ++i) {
            if (std::fabs(h[i]) < 1e-14)
                fail(FEM_ERR_RUNTIME, "diag preconditioner failed: the matrix has a tiny or zero diagonal entry");
            if (h[i] <= 0.0) fail(FEM_ERR_RUNTIME, "diag preconditioner failed: diag(A) must be positive for CG/PCG");
        }
    }
    int kind() const override { return FEM_PC_DIAG; }
    uint32_t size() const override { return n; }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override {
        hipLaunchKernelGGL((k_div<double>), dim3(ew_grid(n)), dim3(256), 0, s, r, d.get(), z, (size_t)n, g);
        FEM_HIP(hipGetLastError());
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override {
        hipLaunchKernelGGL((k_div<float>), dim3(ew_grid(n)), dim3(256), 0, s, r, d32.get(), z, (size_t)n, g);
        FEM_HIP(hipGetLastError());
    }
    double apply_bytes(int s) override { return 3.0 * n * s; }  // r, d, z
    bool has_fp32() const override { return d32.get() != nullptr; }
    void ensure_fp32(hipStream_t s) override {
        if (d32.get()) return;
        d32.alloc(n);
        vec_convert<float, double>(d32.get(), d.get(), 1.0, n, s);
    }
    const double* diag64() const override { return d.get(); }
    const float* diag32() const override { return d32.get(); }
};

std::unique_ptr<Precond> make_precond(CsrMat& A, int kind, const fem_pc_params& p, hipStream_t s) {
    Timer tm;
    std::unique_ptr<Precond> pc;
    switch (kind) {
        case FEM_PC_NONE: pc.reset(new PcNone(A.n)); break;
        case FEM_PC_DIAG: pc.reset(new PcDiag(A, s)); break;
        case FEM_PC_IC0: pc = make_ic0(A, s); break;
        case FEM_PC_ILU0: pc = make_ilu0(A, s); break;
        case FEM_PC_SSOR: pc = make_ssor(A, p.omega, s); break;
        case FEM_PC_AMG: pc = make_amg_cf(A, p, s); break;
        case FEM_PC_SAAMG: pc = make_saamg(A, p, s); break;
        default: fail(FEM_ERR_INVALID_ARGUMENT, "unknown preconditioner kind " + std::to_string(kind));
    }
    if (p.fp32) pc->ensure_fp32(s);
    FEM_HIP(hipStreamSynchronize(s));
    pc->setup_ms = tm.ms();
    return pc;
}

}  // namespace fem

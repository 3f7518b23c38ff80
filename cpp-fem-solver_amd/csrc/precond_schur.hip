// precond_schur.hip — the block upper-triangular Schur-complement
// preconditioner of a saddle-point matrix [[A, B^T], [B, 0]]
// (make_schur_preconditioner, ode-solver/include/math/blockMatrix.hpp:625-685),
// composed on the device from B, B^T and two device preconditioners:
//   w_u = A_pc(r_u)                 (r_u: the first n_vel entries of r)
//   z_p = c S_pc(B w_u - r_p)       (S_pc absent: the identity)
//   z_u = A_pc(r_u - B^T z_p)
// The reference composes host closures with host work vectors; here every
// step is a device launch on the caller's stream, so a Krylov solve with it
// makes no host round trip and can be captured as a graph. The scale c is
// the reference test's S^{-1} ~ 6.1 L_p^{-1} (test_BlockMatrix_schur3d.cpp:
// 149-160).
#include "internal.hpp"

namespace fem {

namespace {

// wp = wp - rp
template <typename T>
__global__ void k_sub_inplace(T* wp, const T* rp, uint32_t n, Gate g) {
    if (g.closed()) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wp[i] = wp[i] - rp[i];
}

// z = c w (z may be w)
template <typename T>
__global__ void k_scale_to(T* z, const T* w, double c, uint32_t n, Gate g) {
    if (g.closed()) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = (T)c * w[i];
}

inline dim3 g256(uint32_t n) { return dim3(std::max(1u, (n + 255) / 256)); }

struct PcSchur : Precond {
    const CsrMat* B;
    const CsrMat* Bt;
    Precond* Apc;
    Precond* Spc;  // null: identity
    double scale;
    uint32_t nv, np;
    DevBuf<double> wu, wp, ru;
    DevBuf<float> wu32, wp32, ru32;
    bool fp32 = false;

    PcSchur(const CsrMat* B_, const CsrMat* Bt_, Precond* A_, Precond* S_, double c)
        : B(B_), Bt(Bt_), Apc(A_), Spc(S_), scale(c), nv(B_->m), np(B_->n) {
        if (Bt->n != nv || Bt->m != np) fail(FEM_ERR_INVALID_ARGUMENT, "Schur: B^T must be n_vel x n_pres for B n_pres x n_vel");
        if (Apc->size() != nv) fail(FEM_ERR_INVALID_ARGUMENT, "Schur: the velocity preconditioner's size differs from n_vel");
        if (Spc && Spc->size() != np)
            fail(FEM_ERR_INVALID_ARGUMENT, "Schur: the Schur-complement preconditioner's size differs from n_pres");
        wu.alloc(nv);
        ru.alloc(nv);
        wp.alloc(np);
    }
    int kind() const override { return FEM_PC_SCHUR; }
    uint32_t size() const override { return nv + np; }
    bool graph_safe() const override { return Apc->graph_safe() && (!Spc || Spc->graph_safe()); }
    void prepare(bool plain64, bool plain32, hipStream_t s) override {
        Apc->prepare(plain64, plain32, s);
        if (Spc) Spc->prepare(plain64, plain32, s);
    }
    bool has_fp32() const override { return fp32; }
    void ensure_fp32(hipStream_t s) override {
        if (fp32) return;
        const_cast<CsrMat*>(B)->ensure_fp32(s);
        const_cast<CsrMat*>(Bt)->ensure_fp32(s);
        Apc->ensure_fp32(s);
        if (Spc) Spc->ensure_fp32(s);
        wu32.alloc(nv);
        ru32.alloc(nv);
        wp32.alloc(np);
        fp32 = true;
    }

    template <typename T>
    void run(const T* r, T* z, T* wu_, T* wp_, T* ru_, hipStream_t s, Gate g) {
        auto pc = [&](Precond* P, const T* in, T* out) {
            if constexpr (std::is_same<T, double>::value) P->apply(in, out, s, g);
            else P->apply32(in, out, s, g);
        };
        pc(Apc, r, wu_);                                                   // w_u = A_pc(r_u)
        csr_spmv<T>(*B, wu_, wp_, EPI_STORE, {}, s, g);                    // w_p = B w_u
        hipLaunchKernelGGL((k_sub_inplace<T>), g256(np), dim3(256), 0, s, wp_, r + nv, np, g);
        T* zp = z + nv;
        if (Spc) {
            pc(Spc, wp_, zp);                                              // z_p = S_pc(w_p - r_p)
            if (scale != 1.0) hipLaunchKernelGGL((k_scale_to<T>), g256(np), dim3(256), 0, s, zp, zp, scale, np, g);
        } else {
            hipLaunchKernelGGL((k_scale_to<T>), g256(np), dim3(256), 0, s, zp, wp_, scale, np, g);
        }
        SpmvArgs<T> ar;
        ar.f = r;
        csr_spmv<T>(*Bt, zp, ru_, EPI_RESID, ar, s, g);                    // r_u - B^T z_p
        pc(Apc, ru_, z);                                                   // z_u
        FEM_HIP(hipGetLastError());
    }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override {
        run<double>(r, z, wu.get(), wp.get(), ru.get(), s, g);
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override {
        if (!fp32) fail(FEM_ERR_RUNTIME, "Schur: the float32 copies were not built");
        run<float>(r, z, wu32.get(), wp32.get(), ru32.get(), s, g);
    }
};

}  // namespace

std::unique_ptr<Precond> make_schur(const CsrMat& B, const CsrMat& Bt, Precond* A_pc, Precond* S_pc, double scale) {
    return std::unique_ptr<Precond>(new PcSchur(&B, &Bt, A_pc, S_pc, scale));
}

}  // namespace fem

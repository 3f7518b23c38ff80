// solver.hpp — Krylov solver object (krylov.hip).
#pragma once
#include <map>
#include <memory>

#include "internal.hpp"

namespace fem {

template <typename T>
struct KWork {
    DevBuf<T> x, r, z, p, q;          // PCG (+ BiCGSTAB: q = v)
    DevBuf<T> rt, y, s, t, ax;        // BiCGSTAB
    DevBuf<T> V, w, wt, dz;           // GMRES basis (m+1)*n, work vectors
    DevBuf<double> gm;                // GMRES small dense state
    DevBuf<T> b;                      // right-hand side (internal copy)
    uint32_t n = 0;
    int m = 0;
    void ensure(int solver, uint32_t n_, int m_);
};

struct GraphCache {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int tag = -1;
    void reset();
    ~GraphCache() { reset(); }
};

// one captured graph per (solver, chunk length) in use
struct GraphSet {
    std::map<int, std::unique_ptr<GraphCache>> g;
    GraphCache& at(int tag) {
        auto& p = g[tag];
        if (!p) p.reset(new GraphCache);
        return *p;
    }
};

class KrylovSolver {
   public:
    KrylovSolver(LinOp* A, Precond* M, const fem_solve_opts& o, hipStream_t s);
    ~KrylovSolver();
    // b, x: device pointers (x is overwritten; every solve starts from x0 = 0)
    void solve(const double* b, double* x, fem_solve_stats* st);
    fem_solve_opts opts;

   private:
    void solve_once(const double* b, double* x, fem_solve_stats* st);
    template <typename T>
    int run(KWork<T>& W, double tol, int max_iter, bool soft, long long* spmv);
    template <typename T>
    void pcg_init(KWork<T>& W);
    template <typename T>
    void pcg_iter(KWork<T>& W);
    template <typename T>
    void bi_init(KWork<T>& W);
    template <typename T>
    void bi_iter(KWork<T>& W);
    template <typename T>
    void gm_cycle(KWork<T>& W);
    template <typename T>
    void precond(const T* r, T* z, Gate g);
    template <typename T>
    void matvec(const T* x, T* y, Gate g);
    // fused forms with the partial sums of x[i] * y[i] in dotp_ (0: none)
    template <typename T>
    unsigned precond_dot(const T* r, T* z, Gate g);
    template <typename T>
    unsigned matvec_dot(const T* x, T* y, Gate g);
    // captures body into gc (unless gc holds tag already), then launches it
    // unless `launch` is false
    void replay(GraphCache& gc, int tag, const std::function<void()>& body, bool launch = true);
    void read_state();
    double true_residual(const double* b, const double* x, double* work);

    LinOp* A_;
    Precond* M_;
    // the float32 inner solves' operator and preconditioner: A_ and M_, or
    // the preconditioner's permuted form Q A Q^T, Q M^{-1} Q^T with perm_[t]
    // the row of A at position t (vectors of the inner solves are then in
    // that order)
    LinOp* A32_ = nullptr;
    Precond* M32_ = nullptr;
    const uint32_t* perm_ = nullptr;
    hipStream_t s_;
    RedScratch red_;
    DevBuf<KState> st_;
    PinnedBuf<KState> hst_;
    KWork<double> w64_;
    KWork<float> w32_;
    DevBuf<double> res64_, x64_, ax64_;
    GraphSet g64_, g32_;
    // PCG: the iterations each Krylov run of the previous solve needed, in
    // call order (one run per refinement step); a run with an entry replays
    // chunks that end there, so a repeated solve of the same system
    // launches no iteration past convergence and no extra replay
    std::vector<int> hint_;
    int phase_ = 0;
    DevBuf<double> st_work_, st_zero_, st_out_;   // end-of-solve statistics
    DevBuf<double> tr_out_, nrm_out_;             // single scalars read back by the host
    DevBuf<double> dotp_;                         // partial sums of the fused product + dot kernels
    std::unique_ptr<KrylovSolver> fb_;            // solver over the fallback preconditioner
    int pc_mode64_ = 0, pc_mode32_ = 0;  // 0 general, 1 none, 2 diag
    // PCG with M = G^T G in three launches per iteration (pcg_iter_fac):
    // G, G^T, whether each precision uses it, and the two-value partial sums
    // of the t = G r kernel (one pair per block of G's SpMV blocking)
    const CsrMat* fG_ = nullptr;
    const CsrMat* fGt_ = nullptr;
    bool fac64_ = false, fac32_ = false;
    DevBuf<double> facp_;
    template <typename T>
    void pcg_iter_fac(KWork<T>& W, unsigned npq);
};

}  // namespace fem

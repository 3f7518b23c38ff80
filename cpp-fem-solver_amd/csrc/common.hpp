// common.hpp — shared internals of libfemgpu (gfx950 / MI355X).
//
// Error model: internal code throws fem::Error carrying one of the C-ABI
// status codes; the C-ABI layer (capi.cpp) catches it, stores the message in
// a thread-local buffer and returns the code. The codes map 1:1 onto the
// reference's exception types (std::invalid_argument, std::out_of_range,
// std::runtime_error; ode-solver/src/math/array.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "femgpu.h"

namespace fem {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

#define FEM_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            ::fem::fail(FEM_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + \
                                            " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

// Stream-ordered allocation scope. Inside an AllocScope(s) every DevBuf is
// allocated from the device's stream-ordered pool (hipMallocAsync on s), and a
// buffer released while the same scope stream is current goes back to the pool
// in stream order (hipFreeAsync on s): setup code that allocates and frees
// hundreds of temporaries then neither maps and unmaps memory nor waits for the
// device at every free. Such buffers are used on s only (the library runs all
// of a handle's work on its stream). A buffer released elsewhere (another
// scope, none, its stream gone) is freed with hipFree, which waits for the
// device. Allocations while s is being captured into a graph use hipMalloc.
inline hipStream_t& alloc_stream() {
    thread_local hipStream_t s = nullptr;
    return s;
}
// keep freed pool memory for reuse instead of returning it at every
// synchronisation (once per device)
void pool_keep_memory();
bool pool_enabled();  // false with FEM_POOL=0 in the environment (plain hipMalloc / hipFree)
struct AllocScope {
    hipStream_t prev;
    explicit AllocScope(hipStream_t s) : prev(alloc_stream()) {
        if (!pool_enabled()) return;
        pool_keep_memory();
        alloc_stream() = s;
    }
    ~AllocScope() { alloc_stream() = prev; }
    AllocScope(const AllocScope&) = delete;
    AllocScope& operator=(const AllocScope&) = delete;
};

// Owning device buffer.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipStream_t st = nullptr;  // pool memory allocated in a scope on this stream
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), st(o.st) { o.p = nullptr; o.n = 0; o.st = nullptr; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n; st = o.st;
            o.p = nullptr; o.n = 0; o.st = nullptr;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (!count) return;
        hipStream_t s = alloc_stream();
        hipStreamCaptureStatus cs = hipStreamCaptureStatusActive;
        if (s && hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
            FEM_HIP(hipMallocAsync(reinterpret_cast<void**>(&p), count * sizeof(T), s));
            st = s;
        } else {
            FEM_HIP(hipMalloc(&p, count * sizeof(T)));
        }
    }
    void release() {
        if (p) {
            if (st && st == alloc_stream()) (void)hipFreeAsync(p, st);
            else (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
        st = nullptr;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        if (count > n) alloc(count);
        if (count) FEM_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    template <typename A>
    void upload(const std::vector<T, A>& h, hipStream_t s) {
        upload(h.data(), h.size(), s);
    }
    void download(T* h, size_t count, hipStream_t s) const {
        if (count) FEM_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    T* get() const { return p; }
};

// Pinned host staging buffer (hipHostMalloc; never hipHostRegister).
template <typename T>
struct PinnedBuf {
    T* p = nullptr;
    size_t n = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
    void ensure(size_t count) {
        if (count <= n) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        FEM_HIP(hipHostMalloc(&p, count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
};

// Process-wide device selection: the library runs on the device that is
// current when the first handle is created.
void ensure_device();

constexpr int kBlock = 256;           // threads per block for streaming kernels
constexpr int kMaxRedBlocks = 2048;   // cap on blocks of a reduction kernel

inline unsigned red_blocks(size_t n, int per_thread = 4) {
    size_t b = (n + (size_t)kBlock * per_thread - 1) / ((size_t)kBlock * per_thread);
    if (b < 1) b = 1;
    if (b > (size_t)kMaxRedBlocks) b = kMaxRedBlocks;
    return (unsigned)b;
}

// Scratch for two-stage deterministic reductions: the per-block partial sums
// of stage 1, read by the stage-2 kernel (or by the blocks of a kernel that
// folds stage 2 in)
struct RedScratch {
    DevBuf<double> partials;  // kMaxRedBlocks * kMaxRedVals
    static constexpr int kMaxRedVals = 4;
    void init(hipStream_t s);
};

// Device-resident Krylov state shared by the kernels of one solve. It is
// written by one thread: the one-block tail kernel of a reduction, or thread
// 0 of block 0 of a kernel that folds the tail in.
struct KState {
    double normb;     // ||b||
    double tol;       // relative tolerance
    double rz;        // PCG rho = r^T z; BiCGSTAB rho_old
    double pap;       // p^T A p
    double alpha, beta, omega;
    double normr;     // ||r|| of the recursive residual
    double rho;       // BiCGSTAB rho_{k+1}
    double tt, ts, rv;
    double hnorm;     // GMRES h_{j+1,j} before rotation
    double beta_g;    // GMRES cycle residual norm
    double true_normr;
    int it;           // loop index (GMRES: total Arnoldi steps)
    int it_end;       // reference iteration semantics
    int max_iter;
    int done;         // 0 running, 1 converged/stopped, 2 error, 3 iteration cap
    int err;          // SolverErr code (negative: soft stop inside refinement)
    int soft;         // 1: breakdowns end the solve instead of failing it
    int check;        // BiCGSTAB: true-residual check requested
    int replace;      // BiCGSTAB: residual replacement requested
    int live;         // GMRES: inner Arnoldi loop continues
    int conv;         // GMRES: converged inside the current cycle
    int j_used;       // GMRES: Arnoldi steps in the current cycle
    int need_true;    // GMRES: cycle starts from the true residual
    int iters;        // iterations run (operator applications in the loop)
    int minicg;       // 1: mini_cg semantics (no breakdown tests, raw loop index)
    double rz_prev;   // PCG: rz before the current iteration's update (fused p update)
    int fac_go;       // PCG with M = G^T G: this iteration's t = G r ran (gate of the p update)
};

enum SolverErr : int {
    SE_NONE = 0,
    SE_PCG_RZ = 1,        // r^T M^{-1} r <= 0
    SE_PCG_PAP_SMALL = 2, // |p^T A p| < 1e-30
    SE_PCG_PAP_NEG = 3,   // p^T A p <= 0
    SE_BI_RHO = 4,
    SE_BI_RHO_OLD = 5,
    SE_BI_OMEGA = 6,
    SE_BI_RV = 7,
    SE_BI_TT = 8,
    SE_GM_DEN = 9,
    SE_GM_TRI = 10,
};

const char* solver_err_msg(int code);

// std::vector storage that resize and the size constructor leave
// uninitialized: host setup arrays that a (parallel) pass writes in full, so
// their pages are first touched by the threads that fill them
template <typename T>
struct NoInitAlloc : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <typename U>
    NoInitAlloc(const NoInitAlloc<U>&) {}
    template <typename U>
    void construct(U* p) noexcept {
        ::new ((void*)p) U;
    }
    template <typename U, typename... Args>
    void construct(U* p, Args&&... a) {
        ::new ((void*)p) U(std::forward<Args>(a)...);
    }
};
template <typename T>
using HostVec = std::vector<T, NoInitAlloc<T>>;

// Host-side setup loops: f(0), ..., f(count - 1) on up to 16 threads, items
// taken in order from a shared counter. f must not throw (workers would end
// the process); it reports failures through state the caller checks after.
template <typename F>
void host_parallel(uint32_t count, F&& f) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = std::min<unsigned>(std::min(hw, 16u), count);
    if (nt <= 1) {
        for (uint32_t i = 0; i < count; ++i) f(i);
        return;
    }
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (uint32_t i = next.fetch_add(1, std::memory_order_relaxed); i < count;
             i = next.fetch_add(1, std::memory_order_relaxed))
            f(i);
    };
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (unsigned t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

// Setup-phase timing: with FEM_TRACE_SETUP=1 in the environment, lap() waits
// for the stream and prints "[femgpu] <scope>: <phase> <ms>" to stderr, the
// host wall time since the previous lap. Without it, lap() does nothing.
bool trace_setup();
struct SetupTrace {
    const char* scope;
    hipStream_t s;
    bool on;
    std::chrono::steady_clock::time_point t0;
    SetupTrace(const char* sc, hipStream_t st) : scope(sc), s(st), on(trace_setup()) {
        if (on) {
            (void)hipStreamSynchronize(s);
            t0 = std::chrono::steady_clock::now();
        }
    }
    void lap(const char* phase) {
        if (!on) return;
        (void)hipStreamSynchronize(s);
        const auto t1 = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[femgpu] %s: %s %.3f ms\n", scope, phase,
                     std::chrono::duration<double, std::milli>(t1 - t0).count());
        t0 = t1;
    }
};

}  // namespace fem

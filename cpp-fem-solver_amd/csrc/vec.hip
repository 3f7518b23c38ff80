// vec.hip — elementwise vector kernels, deterministic dot product, and the
// library-wide device/bookkeeping helpers.
#include <cstdlib>
#include <mutex>

#include "internal.hpp"
#include "reduce.hpp"

namespace fem {

static std::once_flag g_dev_once;
static int g_dev = -1;

void ensure_device() {
    std::call_once(g_dev_once, [] {
        int cnt = 0;
        if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0) return;
        (void)hipGetDevice(&g_dev);
    });
    if (g_dev < 0) fail(FEM_ERR_DEVICE, "no HIP device is available (libfemgpu has no CPU fallback)");
    FEM_HIP(hipSetDevice(g_dev));
}

bool pool_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("FEM_POOL");
        return !(e && e[0] == '0');
    }();
    return on;
}

void pool_keep_memory() {
    // the library runs on one device (ensure_device), so one flag suffices
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        hipMemPool_t pool = nullptr;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) return;
        uint64_t keep = ~0ull;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    });
}

bool trace_setup() {
    static const bool on = [] {
        const char* e = std::getenv("FEM_TRACE_SETUP");
        return e && *e && std::string(e) != "0";
    }();
    return on;
}

const char* solver_err_msg(int code) {
    switch (code) {
        case SE_PCG_RZ: return "PCG failed: r^T M^{-1} r <= 0, the preconditioner may not be SPD";
        case SE_PCG_PAP_SMALL: return "PCG breakdown: p^T A p too small";
        case SE_PCG_PAP_NEG: return "PCG failed: p^T A p <= 0, the matrix A may not be SPD";
        case SE_BI_RHO: return "BiCGSTAB breakdown: rho_{k+1} = r0~^T r_k too small";
        case SE_BI_RHO_OLD: return "BiCGSTAB breakdown: rho_k too small";
        case SE_BI_OMEGA: return "BiCGSTAB breakdown: omega_k = 0";
        case SE_BI_RV: return "BiCGSTAB breakdown: r0~^T v_{k+1} too small";
        case SE_BI_TT: return "BiCGSTAB breakdown: t^T t too small";
        case SE_GM_DEN: return "GMRES breakdown: Hessenberg pivot too small";
        case SE_GM_TRI: return "GMRES breakdown: upper triangular pivot too small";
        default: return "solver error";
    }
}

void RedScratch::init(hipStream_t s) {
    if (partials.get()) return;
    partials.alloc((size_t)kMaxRedBlocks * kMaxRedVals);
}

template <typename T>
__global__ void k_fill(T* y, T val, size_t n, Gate g) {
    if (g.closed()) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] = val;
}

template <typename TD, typename TS>
__global__ void k_convert(TD* y, const TS* x, double scale, size_t n, Gate g) {
    if (g.closed()) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = (TD)(scale == 1.0 ? (double)x[i] : (double)x[i] * scale);
}

static inline unsigned ew_blocks(size_t n) {
    size_t b = (n + kBlock * 4 - 1) / (kBlock * 4);
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (unsigned)b;
}

template <typename T>
void vec_fill(T* y, T val, size_t n, hipStream_t s, Gate g) {
    if (!n) return;
    hipLaunchKernelGGL((k_fill<T>), dim3(ew_blocks(n)), dim3(kBlock), 0, s, y, val, n, g);
    FEM_HIP(hipGetLastError());
}
template <typename TD, typename TS>
void vec_convert(TD* y, const TS* x, double scale, size_t n, hipStream_t s, Gate g) {
    if (!n) return;
    hipLaunchKernelGGL((k_convert<TD, TS>), dim3(ew_blocks(n)), dim3(kBlock), 0, s, y, x, scale, n, g);
    FEM_HIP(hipGetLastError());
}
template void vec_fill<double>(double*, double, size_t, hipStream_t, Gate);
template void vec_fill<float>(float*, float, size_t, hipStream_t, Gate);
template void vec_fill<uint32_t>(uint32_t*, uint32_t, size_t, hipStream_t, Gate);
template void vec_convert<float, double>(float*, const double*, double, size_t, hipStream_t, Gate);
template void vec_convert<double, float>(double*, const float*, double, size_t, hipStream_t, Gate);
template void vec_convert<double, double>(double*, const double*, double, size_t, hipStream_t, Gate);
template void vec_convert<float, float>(float*, const float*, double, size_t, hipStream_t, Gate);

// reads n doubles and leaves one sum per block in sink (evicts the caches
// with clean lines: the next kernel pays no write-back)
__global__ void k_read_sweep(const double* a, size_t n, double* sink) {
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) s += a[i];
    double v[1] = {s};
    block_sum<1>(v);
    if (threadIdx.x == 0) sink[blockIdx.x] = v[0];
}

// streaming read of n 16-byte words (the device's read-bandwidth ceiling)
__global__ void k_stream_read(const float4* __restrict__ a, size_t n, float* sink) {
    float acc = 0.f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 v = a[i];
        acc += (v.x + v.y) + (v.z + v.w);
    }
    if (acc == 0.5f) sink[blockIdx.x] = acc;  // keeps the loads; never true for the filled data
}

void stream_read(const void* buf, size_t bytes, float* sink, unsigned blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_stream_read, dim3(blocks), dim3(kBlock), 0, s, (const float4*)buf, bytes / 16, sink);
    FEM_HIP(hipGetLastError());
}

void cache_flush_read(const double* buf, size_t n, double* sink, unsigned blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_read_sweep, dim3(blocks), dim3(kBlock), 0, s, buf, n, sink);
    FEM_HIP(hipGetLastError());
}

__global__ void k_dot(const double* a, const double* b, size_t n, double* partials) {
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += a[i] * b[i];
    double v[1] = {s};
    block_partials<1>(v, partials);
}

struct FinOut {
    double* out;
    __device__ void operator()(double (&t)[1]) const { *out = t[0]; }
};

double dot_sync(const double* a, const double* b, size_t n, double* partials, hipStream_t s) {
    DevBuf<double> out(1);
    const unsigned G = red_blocks(n);
    hipLaunchKernelGGL(k_dot, dim3(G), dim3(kBlock), 0, s, a, b, n, partials);
    launch_finalize<1>(partials, G, FinOut{out.get()}, s);
    FEM_HIP(hipGetLastError());
    double h = 0.0;
    out.download(&h, 1, s);
    FEM_HIP(hipStreamSynchronize(s));
    return h;
}

// ---- dot product without allocation ------------------------------------------
// The per-block partial sums of dot_noalloc live in this file-scope device
// array (one copy per device, kept until the process ends), not in memory
// the library allocates. Host code never passes the array's name to a
// kernel: the array's address on a device comes from hipGetSymbolAddress,
// looked up once per device with that device current.
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kMaxRedBlocks];

static double* scratch_on_current_device() {
    static std::mutex mu;
    static double* addr[kMaxDevices] = {};
    int dev = 0;
    FEM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(FEM_ERR_DEVICE, "dot_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(FEM_ERR_DEVICE, std::string("dot_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
        addr[dev] = static_cast<double*>(p);
    }
    return addr[dev];
}

// out = a . b on device pointers, enqueued on s: dot_sync's kernels and
// grid (the same sums in the same order), with the partial sums in the
// device's scratch array. Nothing is allocated.
void dot_noalloc(size_t n, const double* a, const double* b, double* out, hipStream_t s) {
    if (n >= (1ull << 32)) fail(FEM_ERR_INVALID_ARGUMENT, "dot_noalloc: vectors of 2^32 or more entries");
    if (!a || !b || !out) fail(FEM_ERR_INVALID_ARGUMENT, "dot_noalloc: null pointer");
    const unsigned G = red_blocks(n);
    if (G > (unsigned)kMaxRedBlocks) fail(FEM_ERR_INVALID_ARGUMENT, "dot_noalloc: more partial sums than the scratch holds");
    double* part = scratch_on_current_device();
    hipLaunchKernelGGL(k_dot, dim3(G), dim3(kBlock), 0, s, a, b, n, part);
    launch_finalize<1>(part, G, FinOut{out}, s);
    FEM_HIP(hipGetLastError());
}

}  // namespace fem

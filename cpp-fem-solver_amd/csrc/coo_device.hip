// coo_device.hip — the input step on the device:
//   SMatrix COO constructor   ode-solver/src/math/array.cpp:2072-2182
//   BlockMatrix::condense     ode-solver/include/math/blockMatrix.hpp:269-315
//
// COO -> CSR: validate the indices, drop zero values, sort by (row, col)
// keeping input order among equal keys (a stable LSD radix sort of the
// 64-bit keys row << 32 | col with the input position as payload), sum each
// run of equal keys sequentially in that order, drop sums equal to zero, and
// count rows for the row pointers. The arithmetic is the reference's: the
// first value of a run, then += each following value.
//
// condense: one thread per block row collects the block columns of its bs
// scalar rows, sorts and de-duplicates them, then adds every scalar entry, in
// scalar-row then stored order, into its zero-initialised block position.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "internal.hpp"

namespace fem {
namespace {

inline unsigned g1(uint64_t n) { return (unsigned)std::max<uint64_t>(1, (n + 255) / 256); }

template <typename T>
uint64_t scan_excl(const T* in, T* out, uint64_t n1, hipStream_t s) {
    size_t tmp = 0;
    FEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int64_t)n1, s));
    DevBuf<char> t(std::max<size_t>(tmp, 1));
    FEM_HIP(hipcub::DeviceScan::ExclusiveSum(t.get(), tmp, in, out, (int64_t)n1, s));
    T h = 0;
    FEM_HIP(hipMemcpyAsync(&h, out + n1 - 1, sizeof(T), hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    return (uint64_t)h;
}

// first input position with an index out of range (atomicMin), keep flags
__global__ void k_coo_check(const double* v, const uint32_t* r, const uint32_t* c, uint64_t nin, uint32_t n,
                            uint32_t m, unsigned long long* bad, uint32_t* keep) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k > nin) return;
    if (k == nin) { keep[nin] = 0; return; }
    if (r[k] >= n || c[k] >= m) atomicMin(bad, (unsigned long long)k);
    keep[k] = v[k] != 0.0 ? 1u : 0u;
}

__global__ void k_coo_compact(const uint32_t* r, const uint32_t* c, const uint32_t* keep, const uint32_t* pos,
                              uint64_t nin, unsigned long long* key, uint32_t* idx) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= nin || !keep[k]) return;
    const uint32_t p = pos[k];
    key[p] = ((unsigned long long)r[k] << 32) | c[k];
    idx[p] = (uint32_t)k;
}

__global__ void k_run_heads(const unsigned long long* key, uint64_t nk, uint32_t* head) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t > nk) return;
    if (t == nk) { head[nk] = 0; return; }
    head[t] = (t == 0 || key[t] != key[t - 1]) ? 1u : 0u;
}

// one thread per run: sequential sum in input order; keep2 = sum != 0
__global__ void k_run_sum(const unsigned long long* key, const uint32_t* idx, const uint32_t* head,
                          const uint32_t* runpos, uint64_t nk, const double* v, double* rsum, uint32_t* rrow,
                          uint32_t* rcol, uint32_t* keep2) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= nk || !head[t]) return;
    const uint32_t g = runpos[t];
    double s = v[idx[t]];
    uint64_t u = t + 1;
    while (u < nk && key[u] == key[t]) s += v[idx[u++]];
    rsum[g] = s;
    rrow[g] = (uint32_t)(key[t] >> 32);
    rcol[g] = (uint32_t)(key[t] & 0xffffffffULL);
    keep2[g] = s != 0.0 ? 1u : 0u;
}

__global__ void k_run_emit(const double* rsum, const uint32_t* rrow, const uint32_t* rcol, const uint32_t* keep2,
                           const uint32_t* pos2, uint64_t nr, uint32_t* ci, double* vv, uint32_t* rowcnt) {
    const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (g >= nr || !keep2[g]) return;
    const uint32_t p = pos2[g];
    ci[p] = rcol[g];
    vv[p] = rsum[g];
    atomicAdd(&rowcnt[rrow[g]], 1u);
}

// ---- condense --------------------------------------------------------------
constexpr int kMaxBlockCols = 256;  // distinct block columns per block row

__device__ int blk_cols(const uint32_t* rp, const uint32_t* ci, uint32_t b, int bs, uint32_t (&cols)[kMaxBlockCols],
                        uint32_t nbc, int* err) {
    int k = 0;
    for (uint32_t i = b * bs; i < (b + 1) * (uint32_t)bs; ++i)
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
            const uint32_t bj = ci[p] / bs;
            if (bj >= nbc) { atomicOr(err, 1); continue; }
            int q = 0;
            while (q < k && cols[q] < bj) ++q;
            if (q < k && cols[q] == bj) continue;
            if (k == kMaxBlockCols) { atomicOr(err, 2); continue; }
            for (int t = k; t > q; --t) cols[t] = cols[t - 1];
            cols[q] = bj;
            ++k;
        }
    return k;
}

__global__ void k_cond_count(const uint32_t* rp, const uint32_t* ci, uint32_t nbr, int bs, uint32_t nbc,
                             uint32_t* cnt, int* err) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nbr) return;
    if (b == nbr) { cnt[nbr] = 0; return; }
    uint32_t cols[kMaxBlockCols];
    cnt[b] = (uint32_t)blk_cols(rp, ci, b, bs, cols, nbc, err);
}

__global__ void k_cond_fill(const uint32_t* rp, const uint32_t* ci, const double* v, uint32_t nbr, int bs,
                            uint32_t nbc, const uint32_t* brp, uint32_t* bci, double* bv, int* err) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbr) return;
    uint32_t cols[kMaxBlockCols];
    const int k = blk_cols(rp, ci, b, bs, cols, nbc, err);
    const uint32_t o = brp[b];
    const int BV = bs * bs;
    for (int q = 0; q < k; ++q) {
        bci[o + q] = cols[q];
        for (int t = 0; t < BV; ++t) bv[(size_t)(o + q) * BV + t] = 0.0;
    }
    for (uint32_t i = b * bs; i < (b + 1) * (uint32_t)bs; ++i) {
        const int r = (int)(i % bs);
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
            const uint32_t bj = ci[p] / bs, c = ci[p] % bs;
            if (bj >= nbc) continue;
            int q = 0;
            while (q < k && cols[q] != bj) ++q;
            bv[(size_t)(o + q) * BV + r * bs + c] += v[p];
        }
    }
}

}  // namespace

void coo_to_csr_device(uint32_t n, uint32_t m, uint64_t nin, const double* hv, const uint32_t* hr, const uint32_t* hc,
                       CsrMat& out, hipStream_t s) {
    if (nin >= (1ull << 32)) fail(FEM_ERR_INVALID_ARGUMENT, "COO input: too many entries");
    DevBuf<double> v(std::max<uint64_t>(nin, 1));
    DevBuf<uint32_t> r(std::max<uint64_t>(nin, 1)), c(std::max<uint64_t>(nin, 1)), keep(nin + 1), pos(nin + 1);
    v.upload(hv, nin, s);
    r.upload(hr, nin, s);
    c.upload(hc, nin, s);
    DevBuf<unsigned long long> bad(1);
    const unsigned long long none = ~0ull;
    FEM_HIP(hipMemcpyAsync(bad.get(), &none, 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_coo_check, dim3(g1(nin + 1)), dim3(256), 0, s, v.get(), r.get(), c.get(), nin, n, m,
                       bad.get(), keep.get());
    unsigned long long hb = 0;
    FEM_HIP(hipMemcpyAsync(&hb, bad.get(), 8, hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    if (hb != none) fail(FEM_ERR_OUT_OF_RANGE, "COO input: row or column index out of range");
    const uint64_t nk = scan_excl<uint32_t>(keep.get(), pos.get(), nin + 1, s);
    DevBuf<unsigned long long> key(std::max<uint64_t>(nk, 1)), key2(std::max<uint64_t>(nk, 1));
    DevBuf<uint32_t> idx(std::max<uint64_t>(nk, 1)), idx2(std::max<uint64_t>(nk, 1));
    hipLaunchKernelGGL(k_coo_compact, dim3(g1(nin)), dim3(256), 0, s, r.get(), c.get(), keep.get(), pos.get(), nin,
                       key.get(), idx.get());
    if (nk) {
        size_t tmp = 0;
        FEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key.get(), key2.get(), idx.get(), idx2.get(),
                                                   (int64_t)nk, 0, 64, s));
        DevBuf<char> t(std::max<size_t>(tmp, 1));
        FEM_HIP(hipcub::DeviceRadixSort::SortPairs(t.get(), tmp, key.get(), key2.get(), idx.get(), idx2.get(),
                                                   (int64_t)nk, 0, 64, s));
    }
    DevBuf<uint32_t> head(nk + 1), runpos(nk + 1);
    hipLaunchKernelGGL(k_run_heads, dim3(g1(nk + 1)), dim3(256), 0, s, key2.get(), nk, head.get());
    const uint64_t nr = scan_excl<uint32_t>(head.get(), runpos.get(), nk + 1, s);
    DevBuf<double> rsum(std::max<uint64_t>(nr, 1));
    DevBuf<uint32_t> rrow(std::max<uint64_t>(nr, 1)), rcol(std::max<uint64_t>(nr, 1)), keep2(nr + 1), pos2(nr + 1);
    FEM_HIP(hipMemsetAsync(keep2.get() + nr, 0, 4, s));
    hipLaunchKernelGGL(k_run_sum, dim3(g1(nk)), dim3(256), 0, s, key2.get(), idx2.get(), head.get(), runpos.get(), nk,
                       v.get(), rsum.get(), rrow.get(), rcol.get(), keep2.get());
    const uint64_t nnz = scan_excl<uint32_t>(keep2.get(), pos2.get(), nr + 1, s);
    DevBuf<uint32_t> rowcnt(n + 1), rp(n + 1), ci(nnz + 4);
    DevBuf<double> vv(nnz + 4);
    FEM_HIP(hipMemsetAsync(rowcnt.get(), 0, (size_t)(n + 1) * 4, s));
    FEM_HIP(hipMemsetAsync(ci.get() + nnz, 0, 16, s));
    FEM_HIP(hipMemsetAsync(vv.get() + nnz, 0, 32, s));
    hipLaunchKernelGGL(k_run_emit, dim3(g1(nr)), dim3(256), 0, s, rsum.get(), rrow.get(), rcol.get(), keep2.get(),
                       pos2.get(), nr, ci.get(), vv.get(), rowcnt.get());
    scan_excl<uint32_t>(rowcnt.get(), rp.get(), (uint64_t)n + 1, s);
    FEM_HIP(hipGetLastError());
    out.adopt(n, m, nnz, std::move(rp), std::move(ci), std::move(vv), s);
}

void condense_device(const CsrMat& A, int bs, uint32_t nbr, uint32_t nbc, std::vector<uint32_t>& hrp,
                     std::vector<uint32_t>& hci, std::vector<double>& hv, hipStream_t s) {
    if ((uint64_t)nbr * bs > A.n) fail(FEM_ERR_OUT_OF_RANGE, "condense: nbr * BS exceeds the matrix rows");
    DevBuf<int> err(1);
    FEM_HIP(hipMemsetAsync(err.get(), 0, 4, s));
    DevBuf<uint32_t> cnt(nbr + 1), brp(nbr + 1);
    hipLaunchKernelGGL(k_cond_count, dim3(g1(nbr + 1)), dim3(256), 0, s, A.rp.get(), A.ci.get(), nbr, bs, nbc,
                       cnt.get(), err.get());
    const uint64_t nnzb = scan_excl<uint32_t>(cnt.get(), brp.get(), (uint64_t)nbr + 1, s);
    int he = 0;
    FEM_HIP(hipMemcpyAsync(&he, err.get(), 4, hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    if (he & 1) fail(FEM_ERR_OUT_OF_RANGE, "condense: block column out of range");
    if (he & 2) fail(FEM_ERR_RUNTIME, "condense: more than 256 block columns in a block row");
    const int BV = bs * bs;
    DevBuf<uint32_t> bci(std::max<uint64_t>(nnzb, 1));
    DevBuf<double> bv(std::max<uint64_t>(nnzb * BV, 1));
    hipLaunchKernelGGL(k_cond_fill, dim3(g1(nbr)), dim3(256), 0, s, A.rp.get(), A.ci.get(), A.v.get(), nbr, bs, nbc,
                       brp.get(), bci.get(), bv.get(), err.get());
    FEM_HIP(hipGetLastError());
    hrp.resize(nbr + 1);
    hci.resize(nnzb);
    hv.resize(nnzb * BV);
    brp.download(hrp.data(), nbr + 1, s);
    bci.download(hci.data(), nnzb, s);
    bv.download(hv.data(), nnzb * BV, s);
    FEM_HIP(hipStreamSynchronize(s));
}

}  // namespace fem

// reduce.hpp — deterministic wavefront/block reductions (device code).
//
// Sums are always carried in float64. A reduction kernel runs a fixed grid
// for a given length, each thread accumulates a fixed strided subset in
// index order, waves combine with xor-butterflies (64 lanes), blocks combine
// wave partials in wave order, and the last block to arrive (atomic ticket)
// adds the per-block partials in block order. The result therefore depends
// only on the vector length, never on scheduling.
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace fem {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// s[i] += partials[b * NV + i] for b = t, t + blockDim, t + 2 blockDim, ...
// (t = threadIdx.x) in that order. Eight are loaded before they are added;
// past the end a load gives 0.0, whose addition changes no sum.
template <int NV>
__device__ __forceinline__ void add_partials(const double* __restrict__ partials, unsigned nparts, double (&s)[NV]) {
    constexpr unsigned U = 8;
    for (unsigned b = threadIdx.x; b < nparts; b += U * blockDim.x) {
        double v[U][NV];
#pragma unroll
        for (unsigned u = 0; u < U; ++u) {
            const unsigned k = b + u * blockDim.x;
#pragma unroll
            for (int i = 0; i < NV; ++i) v[u][i] = k < nparts ? partials[k * NV + i] : 0.0;
        }
#pragma unroll
        for (unsigned u = 0; u < U; ++u)
#pragma unroll
            for (int i = 0; i < NV; ++i) s[i] += v[u][i];
    }
}

// Block-wide sum of NV values; the result is valid in every thread.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[NV][kBlock / 64];
    __shared__ double res[NV];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[i][w] = v[i];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[threadIdx.x][k];
        res[threadIdx.x] = s;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = res[i];
    __syncthreads();
}

// Stage 1 of a reduction: the block's NV sums go to partials[block * NV + i]
// with plain stores. No device-scope fence is used: on gfx950 a fence at
// agent scope writes back and invalidates the whole L2 (buffer_wbl2 /
// buffer_inv), which stalls every other block that is streaming.
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* partials) {
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) partials[blockIdx.x * NV + i] = v[i];
    }
}

// Stage 2: one block adds the per-block partials in block order (the kernel
// boundary makes stage 1's stores visible) and calls fin(total) on thread 0.
template <int NV, class Fin>
__global__ __launch_bounds__(kBlock) void k_finalize(const double* __restrict__ partials, unsigned nparts, Fin fin) {
    double s[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) s[i] = 0.0;
    add_partials<NV>(partials, nparts, s);
    block_sum<NV>(s);
    if (threadIdx.x == 0) fin(s);
}

template <int NV, class Fin>
inline void launch_finalize(const double* partials, unsigned nparts, Fin fin, hipStream_t s) {
    hipLaunchKernelGGL((k_finalize<NV, Fin>), dim3(1), dim3(kBlock), 0, s, partials, nparts, fin);
}

}  // namespace fem

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

// Each block contributes NV partial sums; the last block to arrive reduces
// them in block order and calls fin(total) on thread 0. Returns true in the
// finalizing block (all threads). `counter` must be zero on entry; the last
// block resets it.
template <int NV, class Fin>
__device__ __forceinline__ void reduce_finalize(double (&v)[NV], double* partials, unsigned* counter, Fin&& fin) {
    block_sum<NV>(v);
    __shared__ int is_last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
            __hip_atomic_store(&partials[blockIdx.x * NV + i], v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        is_last = (prev == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double s[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) s[i] = 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
            s[i] += __hip_atomic_load(&partials[b * NV + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum<NV>(s);
    if (threadIdx.x == 0) {
        *counter = 0u;
        fin(s);
    }
}

}  // namespace fem

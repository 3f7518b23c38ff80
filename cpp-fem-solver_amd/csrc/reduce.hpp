// reduce.hpp — deterministic wavefront/block reductions (device code).
//
// Sums are carried in float64. A reduction kernel runs a fixed grid for a
// given length; each thread accumulates a fixed strided subset in index
// order, wavefronts combine their lanes with a fixed butterfly (group_sum),
// blocks combine wavefront partials in wavefront order and store one partial
// per block (block_partials); a second one-block kernel adds the partials in
// block order (k_finalize). The result depends only on the vector length,
// never on scheduling.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "common.hpp"

namespace fem {

// Lane exchanges without LDS: DPP row permutations for partners 1, 2
// (quad_perm [1,0,3,2], [2,3,0,1]), 4 (row_half_mirror) and 8 (row_mirror),
// and gfx950's v_permlane16_swap / v_permlane32_swap for 16 and 32. The
// mirrors equal xor 4 / xor 8 here because every quad (every 8 lanes) already
// holds one value when they are used. Each step adds the same two operands in
// both lanes of a pair, so all lanes of a group end with bit-identical sums.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_partner(T v) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t b = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = dpp_u32<CTRL>((uint32_t)b), hi = dpp_u32<CTRL>((uint32_t)(b >> 32));
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, dpp_u32<CTRL>(__builtin_bit_cast(uint32_t, v)));
    }
}
// v_permlane{16,32}_swap with both operands v: the results hold the lower and
// the upper lane of each pair (lanes l and l ^ 16 / l ^ 32); their sum, in
// that order, is the same in both lanes
template <bool X32>
__device__ __forceinline__ void swap_pair(uint32_t v, uint32_t& a, uint32_t& b) {
    if constexpr (X32) {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        a = r[0];
        b = r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        a = r[0];
        b = r[1];
    }
}
template <bool X32, typename T>
__device__ __forceinline__ T swap_sum(T v) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t bits = __builtin_bit_cast(uint64_t, v);
        uint32_t la, lb, ha, hb;
        swap_pair<X32>((uint32_t)bits, la, lb);
        swap_pair<X32>((uint32_t)(bits >> 32), ha, hb);
        return __builtin_bit_cast(T, ((uint64_t)ha << 32) | la) + __builtin_bit_cast(T, ((uint64_t)hb << 32) | lb);
    } else {
        uint32_t a, b;
        swap_pair<X32>(__builtin_bit_cast(uint32_t, v), a, b);
        return __builtin_bit_cast(T, a) + __builtin_bit_cast(T, b);
    }
}

// Sum over each aligned group of G lanes (G = 1, 2, 4, ..., 64); the result
// is valid, and identical, in every lane of the group.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
    static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "group of 2^k lanes");
    if constexpr (G >= 2) v += dpp_partner<0xB1>(v);   // quad_perm [1,0,3,2]
    if constexpr (G >= 4) v += dpp_partner<0x4E>(v);   // quad_perm [2,3,0,1]
    if constexpr (G >= 8) v += dpp_partner<0x141>(v);  // row_half_mirror
    if constexpr (G >= 16) v += dpp_partner<0x140>(v); // row_mirror
    if constexpr (G >= 32) v = swap_sum<false>(v);
    if constexpr (G >= 64) v = swap_sum<true>(v);
    return v;
}

__device__ __forceinline__ double wave_sum(double v) { return group_sum<64>(v); }

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

// learned.hip — learned sparse approximate inverse (no counterpart in the
// reference): a small MLP reads each row's local stencil couplings and emits
// the row's entries of a lower-triangular G on A's own lower pattern; the
// preconditioner is M = G^T G (factorized sparse approximate inverse), SPD
// for any network output, applied as two sparse products instead of a
// serial triangular solve.
//
//   features  f_i[d] = -6 a_ij / sqrt(a_ii a_jj), d = 0..5 the directions
//             -nx*ny, -nx, -1, +1, +nx, +nx*ny (0 where row i has no entry)
//   network   6 -> 64 -> 64 -> 7, ReLU, float32 master weights, bfloat16
//             operands on the matrix cores (gfx950's v_mfma_f32_16x16x32_bf16
//             for the K = 64 products, v_mfma_f32_16x16x16_bf16 for K = 16), float32
//             accumulation; flat parameter layout W1[6][64], b1[64], W2[64][64],
//             b2[64], W3[64][7], b3[7] (W[in][out]), 5063 floats
//   G         G_ii = softplus(o_i[0]) dis_i, G_ij = o_i[1+d] dis_j for the
//             lower neighbours (d = 0, 1, 2), dis = a^-1/2; outputs 4..6 unused
//   training  L = 1/(n k) sum_c ||v_c - G^T G A v_c||^2 over k probe vectors
//             (random signs and solver error vectors), Adam; with R = v - M A v
//             and W = A v, dL/dG_ij = -2/(n k) sum_c ((G W)_i R_j + (G R)_i W_j)
//
// MFMA tiling: one wavefront owns a tile of 16 rows. For
// v_mfma_f32_16x16x16_bf16 lane l holds A[l%16][4(l/16)+v], B[4(l/16)+v][l%16]
// and D[4(l/16)+v][l%16], v = 0..3; for v_mfma_f32_16x16x32_bf16
// A[l%16][8(l/16)+v] and B[8(l/16)+v][l%16], v = 0..7, D the same. Layers 2
// and 3 and the back-propagation through W2 (K = 64) take two K = 32
// products; layer 1 (K = 6, padded to 16), the back-propagation through W3
// (K = 7) and the weight gradients (K = the tile's 16 rows) take K = 16. Hidden activations go through LDS
// (16 x 64 bfloat16 per wave) to turn the D layout into the next layer's A
// layout. Weights are packed once per update into B-operand fragments:
// 64 lanes x 4 bfloat16 for W1 (4) and W3^T (4), 64 lanes x 8 for W2 (8),
// W3 (2) and W2^T (8).
// The backward pass recomputes the forward pass per tile, accumulates weight
// gradients per wave in registers (fixed tile order), sums the 4 waves of a
// block in LDS in wave order and writes one partial per block; a second
// kernel adds the partials in block order, so gradients are deterministic.
#include <algorithm>
#include <cmath>
#include <random>

#include "amg.hpp"
#include "handles.hpp"
#include "learned.hpp"
#include "reduce.hpp"

namespace fem {

namespace {

constexpr int NI = 6, NH = 64, NO = 7;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NI * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NO;
// B-operand fragments of the K = 16 products (layer 1, the W3^T back-propagation)
constexpr int FR_W1 = 0, FR_W3T = 4, NFRAG = 8;
// B-operand fragments of the K = 64 products, as two K = 32 chunks each
// (gfx950's v_mfma_f32_16x16x32_bf16): W2 (2 x 4), W3 (2), W2^T (2 x 4)
constexpr int F8_W2 = 0, F8_W3 = 8, F8_W2T = 10, NFRAG8 = 18;
constexpr int kLds = 68;  // LDS row stride (bfloat16) of a 16 x 64 activation tile

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (short)(u >> 16);
}

__device__ __forceinline__ f32x4 mma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// 16 x 16 x 32: lane l holds A[l%16][8(l/16)+v] and B[8(l/16)+v][l%16],
// v = 0..7; D as for 16 x 16 x 16
__device__ __forceinline__ f32x4 mma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
                                                   0, 0, 0);
}
// eight consecutive bfloat16 of an LDS tile row (two 8-byte reads: rows are
// 8-byte aligned)
__device__ __forceinline__ bf16x8 lds8(const short* p) {
    const bf16x4 lo = *reinterpret_cast<const bf16x4*>(p), hi = *reinterpret_cast<const bf16x4*>(p + 4);
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

// element (k, n) of the B operand of fragment family `fam`
__device__ __forceinline__ float frag_elem(const float* P, int fam, int k, int n) {
    switch (fam) {
        case 0: return k < NI ? P[OFF_W1 + k * NH + n] : 0.f;       // W1, K padded to 16
        case 1: return P[OFF_W2 + k * NH + n];                      // W2
        case 2: return n < NO ? P[OFF_W3 + k * NO + n] : 0.f;       // W3, N padded to 16
        case 3: return k < NO ? P[OFF_W3 + n * NO + k] : 0.f;       // W3^T, K padded to 16
        default: return P[OFF_W2 + n * NH + k];                     // W2^T
    }
}

// fragments: NFRAG x 64 lanes x 4 bfloat16, then NFRAG8 x 64 lanes x 8
__global__ void k_pack(const float* __restrict__ P, bf16x4* __restrict__ frag, bf16x8* __restrict__ frag8) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < NFRAG * 64) {
        const int f = t / 64, l = t % 64;
        const int fam = f < FR_W3T ? 0 : 3, nt = f < FR_W3T ? f - FR_W1 : f - FR_W3T;
        bf16x4 r;
        for (int v = 0; v < 4; ++v) r[v] = f2bf(frag_elem(P, fam, 4 * (l / 16) + v, 16 * nt + l % 16));
        frag[t] = r;
        return;
    }
    const int u = t - NFRAG * 64;
    if (u >= NFRAG8 * 64) return;
    const int f = u / 64, l = u % 64;
    int fam, kt, nt;
    if (f < F8_W3) { fam = 1; kt = (f - F8_W2) / 4; nt = (f - F8_W2) % 4; }
    else if (f < F8_W2T) { fam = 2; kt = f - F8_W3; nt = 0; }
    else { fam = 4; kt = (f - F8_W2T) / 4; nt = (f - F8_W2T) % 4; }
    bf16x8 r;
    for (int v = 0; v < 8; ++v) r[v] = f2bf(frag_elem(P, fam, 32 * kt + 8 * (l / 16) + v, 16 * nt + l % 16));
    frag8[u] = r;
}

struct MlpArgs {
    const short* feat;    // n x 8 bfloat16 bits (features 0..5, 6..7 zero)
    const bf16x4* frag;   // NFRAG x 64
    const bf16x8* frag8;  // NFRAG8 x 64
    const float* P;       // master parameters (biases)
    uint32_t n;
    float* out;           // forward: n x 8 (outputs 0..6)
    const float* dout;    // backward: dL/do, n x 8
    float* partials;      // backward: gridDim.x x NPARAM
};

template <bool BWD>
__global__ __launch_bounds__(256) void k_mlp(MlpArgs a) {
    // tiles are read as 8-byte bfloat16 x 4 vectors
    __shared__ __attribute__((aligned(16))) short H1[4][16][kLds];
    __shared__ __attribute__((aligned(16))) short H2[4][16][kLds];
    __shared__ __attribute__((aligned(16))) short G2[BWD ? 4 : 1][16][kLds];
    __shared__ __attribute__((aligned(16))) short G1[BWD ? 4 : 1][16][kLds];
    __shared__ float red[BWD ? kMlpParams : 1];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int lr = l % 16, lg = l / 16;
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 aW1[4], aW2[4][4], aW3[4];
    float ab1[4], ab2[4], ab3 = 0.f;
    if constexpr (BWD) {
        for (int i = 0; i < 4; ++i) {
            aW1[i] = zero;
            aW3[i] = zero;
            ab1[i] = ab2[i] = 0.f;
            for (int j = 0; j < 4; ++j) aW2[i][j] = zero;
        }
    }
    const uint32_t ntiles = (a.n + 15) / 16;
    // every wave of a block runs the same number of iterations (barriers)
    for (uint32_t base = blockIdx.x * 4; base < ntiles; base += gridDim.x * 4) {
        const uint32_t r0 = (base + w) * 16;
        // ---- layer 1: z1 = X W1 + b1 ------------------------------------
        bf16x4 xa = {0, 0, 0, 0};
        if (lg < 2 && r0 + lr < a.n) xa = *reinterpret_cast<const bf16x4*>(a.feat + (size_t)(r0 + lr) * 8 + 4 * lg);
        f32x4 z1[4], z2[4];
        for (int nt = 0; nt < 4; ++nt) {
            z1[nt] = mma(xa, a.frag[(FR_W1 + nt) * 64 + l], zero);
            const float b = a.P[OFF_B1 + 16 * nt + lr];
            for (int v = 0; v < 4; ++v) {
                z1[nt][v] += b;
                H1[w][4 * lg + v][16 * nt + lr] = f2bf(z1[nt][v] > 0.f ? z1[nt][v] : 0.f);
            }
        }
        __syncthreads();
        // ---- layer 2: z2 = h1 W2 + b2 -----------------------------------
        bf16x8 ha[2];
        for (int kt = 0; kt < 2; ++kt) ha[kt] = lds8(&H1[w][lr][32 * kt + 8 * lg]);
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = zero;
            for (int kt = 0; kt < 2; ++kt) acc = mma32(ha[kt], a.frag8[(F8_W2 + kt * 4 + nt) * 64 + l], acc);
            const float b = a.P[OFF_B2 + 16 * nt + lr];
            for (int v = 0; v < 4; ++v) {
                acc[v] += b;
                H2[w][4 * lg + v][16 * nt + lr] = f2bf(acc[v] > 0.f ? acc[v] : 0.f);
            }
            z2[nt] = acc;
        }
        __syncthreads();
        // ---- layer 3: o = h2 W3 + b3 ------------------------------------
        if constexpr (!BWD) {
            f32x4 o = zero;
            for (int kt = 0; kt < 2; ++kt)
                o = mma32(lds8(&H2[w][lr][32 * kt + 8 * lg]), a.frag8[(F8_W3 + kt) * 64 + l], o);
            if (lr < 8) {
                const float b = lr < NO ? a.P[OFF_B3 + lr] : 0.f;
                for (int v = 0; v < 4; ++v) {
                    const uint32_t row = r0 + 4 * lg + v;
                    if (row < a.n) a.out[(size_t)row * 8 + lr] = lr < NO ? o[v] + b : 0.f;
                }
            }
        } else {
            // dL/do in the B layout (rows 4lg+v, column lr) and the A layout
            // (row lr, columns 4lg+v)
            float gb[4];
            bf16x4 gbb, gab = {0, 0, 0, 0};
            for (int v = 0; v < 4; ++v) {
                const uint32_t row = r0 + 4 * lg + v;
                gb[v] = (lr < NO && row < a.n) ? a.dout[(size_t)row * 8 + lr] : 0.f;
                gbb[v] = f2bf(gb[v]);
            }
            if (lg < 2 && r0 + lr < a.n)
                for (int v = 0; v < 4; ++v) {
                    const int c = 4 * lg + v;
                    gab[v] = f2bf(c < NO ? a.dout[(size_t)(r0 + lr) * 8 + c] : 0.f);
                }
            {
                float s = gb[0] + gb[1] + gb[2] + gb[3];
                s = swap_sum<true>(swap_sum<false>(s));
                ab3 += s;
            }
            // dW3 += h2^T g
            for (int mt = 0; mt < 4; ++mt) {
                bf16x4 at;
                for (int v = 0; v < 4; ++v) at[v] = H2[w][4 * lg + v][16 * mt + lr];
                aW3[mt] = mma(at, gbb, aW3[mt]);
            }
            // dz2 = (g W3^T) * [z2 > 0]
            for (int nt = 0; nt < 4; ++nt) {
                f32x4 d = mma(gab, a.frag[(FR_W3T + nt) * 64 + l], zero);
                float s = 0.f;
                for (int v = 0; v < 4; ++v) {
                    const float g = z2[nt][v] > 0.f ? d[v] : 0.f;
                    s += g;
                    G2[w][4 * lg + v][16 * nt + lr] = f2bf(g);
                }
                s = swap_sum<true>(swap_sum<false>(s));
                ab2[nt] += s;
            }
            __syncthreads();
            // dW2 += h1^T dz2
            {
                bf16x4 at[4], bt[4];
                for (int t = 0; t < 4; ++t)
                    for (int v = 0; v < 4; ++v) {
                        at[t][v] = H1[w][4 * lg + v][16 * t + lr];
                        bt[t][v] = G2[w][4 * lg + v][16 * t + lr];
                    }
                for (int mt = 0; mt < 4; ++mt)
                    for (int nt = 0; nt < 4; ++nt) aW2[mt][nt] = mma(at[mt], bt[nt], aW2[mt][nt]);
            }
            // dz1 = (dz2 W2^T) * [z1 > 0]
            {
                bf16x8 ga[2];
                for (int kt = 0; kt < 2; ++kt) ga[kt] = lds8(&G2[w][lr][32 * kt + 8 * lg]);
                for (int nt = 0; nt < 4; ++nt) {
                    f32x4 d = zero;
                    for (int kt = 0; kt < 2; ++kt) d = mma32(ga[kt], a.frag8[(F8_W2T + kt * 4 + nt) * 64 + l], d);
                    float s = 0.f;
                    for (int v = 0; v < 4; ++v) {
                        const float g = z1[nt][v] > 0.f ? d[v] : 0.f;
                        s += g;
                        G1[w][4 * lg + v][16 * nt + lr] = f2bf(g);
                    }
                    s = swap_sum<true>(swap_sum<false>(s));
                    ab1[nt] += s;
                }
            }
            __syncthreads();
            // dW1 += X^T dz1
            {
                bf16x4 xt;
                for (int v = 0; v < 4; ++v) {
                    const uint32_t row = r0 + 4 * lg + v;
                    xt[v] = (lr < NI && row < a.n) ? a.feat[(size_t)row * 8 + lr] : (short)0;
                }
                for (int nt = 0; nt < 4; ++nt) {
                    bf16x4 bt;
                    for (int v = 0; v < 4; ++v) bt[v] = G1[w][4 * lg + v][16 * nt + lr];
                    aW1[nt] = mma(xt, bt, aW1[nt]);
                }
            }
        }
        __syncthreads();
    }
    if constexpr (BWD) {
        // sum the four waves in wave order, then write this block's partial
        for (int ww = 0; ww < 4; ++ww) {
            if (w == ww) {
                auto put = [&](int idx, float val) {
                    if (ww == 0) red[idx] = val;
                    else red[idx] += val;
                };
                for (int v = 0; v < 4; ++v) {
                    const int r = 4 * lg + v;
                    for (int nt = 0; nt < 4; ++nt)
                        if (r < NI) put(OFF_W1 + r * NH + 16 * nt + lr, aW1[nt][v]);
                    for (int mt = 0; mt < 4; ++mt) {
                        for (int nt = 0; nt < 4; ++nt) put(OFF_W2 + (16 * mt + r) * NH + 16 * nt + lr, aW2[mt][nt][v]);
                        if (lr < NO) put(OFF_W3 + (16 * mt + r) * NO + lr, aW3[mt][v]);
                    }
                }
                if (lg == 0) {
                    for (int nt = 0; nt < 4; ++nt) {
                        put(OFF_B1 + 16 * nt + lr, ab1[nt]);
                        put(OFF_B2 + 16 * nt + lr, ab2[nt]);
                    }
                    if (lr < NO) put(OFF_B3 + lr, ab3);
                }
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) a.partials[(size_t)blockIdx.x * kMlpParams + i] = red[i];
    }
}

// grad[p] = scale * sum over blocks (in block order) of partials[b][p]
__global__ void k_grad_reduce(const float* __restrict__ partials, unsigned nblk, double scale, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    double s = 0.0;
    for (unsigned b = 0; b < nblk; ++b) s += (double)partials[(size_t)b * kMlpParams + p];
    grad[p] = (float)(s * scale);
}

__global__ void k_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float gi = g[p];
    const float mi = b1 * m[p] + (1.f - b1) * gi;
    const float vi = b2 * v[p] + (1.f - b2) * gi * gi;
    m[p] = mi;
    v[p] = vi;
    P[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// float32 features (host-supplied, n x 6) -> bfloat16 n x 8
__global__ void k_feat_pack(const float* __restrict__ f, uint32_t n, short* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < 8; ++d) out[(size_t)i * 8 + d] = d < NI ? f2bf(f[(size_t)i * NI + d]) : (short)0;
}

// ---- stencil analysis --------------------------------------------------
struct Offsets {
    long long o[6];
};

__device__ __forceinline__ int dir_of(const Offsets& G, long long delta) {
    for (int d = 0; d < 6; ++d)
        if (G.o[d] == delta) return d;
    return -1;
}

// err bits: 1 = entry outside the stencil, 2 = diagonal missing or <= 0,
// 4 = pattern not symmetric
__global__ void k_spai_rows(uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                            const double* __restrict__ v, Offsets G, int* __restrict__ epos, int* __restrict__ dpos,
                            double* __restrict__ dis, int* err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int e[6] = {-1, -1, -1, -1, -1, -1};
    int dp = -1;
    int bad = 0;
    for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
        const uint32_t j = ci[p];
        if (j == i) { dp = (int)p; continue; }
        const int d = dir_of(G, (long long)j - (long long)i);
        if (d < 0) bad |= 1;
        else e[d] = (int)p;
    }
    if (dp < 0 || !(v[dp] > 0.0)) {
        bad |= 2;
        dis[i] = 0.0;
    } else {
        dis[i] = 1.0 / sqrt(v[dp]);
    }
    dpos[i] = dp;
    for (int d = 0; d < 6; ++d) epos[(size_t)i * 6 + d] = e[d];
    if (bad) atomicOr(err, bad);
}

__global__ void k_spai_features(uint32_t n, const uint32_t* __restrict__ ci, const double* __restrict__ v,
                                 const int* __restrict__ epos, const double* __restrict__ dis, short* __restrict__ feat,
                                 float* __restrict__ feat32, int* err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int bad = 0;
    for (int d = 0; d < 8; ++d) {
        float f = 0.f;
        if (d < 6) {
            const int p = epos[(size_t)i * 6 + d];
            if (p >= 0) {
                const uint32_t j = ci[p];
                if (epos[(size_t)j * 6 + (5 - d)] < 0) bad |= 4;
                f = (float)(-6.0 * v[p] * dis[i] * dis[j]);
            }
            if (feat32) feat32[(size_t)i * 6 + d] = f;
        }
        feat[(size_t)i * 8 + d] = d < 6 ? f2bf(f) : (short)0;
    }
    if (bad) atomicOr(err, bad);
}

__device__ __forceinline__ double softplus(double x) { return x > 30.0 ? x : log1p(exp(x)); }

// G (lower, A's pattern) and G^T (upper) values from the network outputs;
// gdir[i*4 + 0] is the diagonal's position in G, gdir[i*4 + 1 + d] that of
// the lower neighbour in direction d (-1: none); tpos maps a G position to
// the position of the same entry in G^T
__global__ void k_fsai_assemble(uint32_t n, const uint32_t* __restrict__ gci, const int* __restrict__ gdir,
                                const int* __restrict__ tpos, const double* __restrict__ dis,
                                const float* __restrict__ o, double* __restrict__ gv, double* __restrict__ tv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int pd = gdir[(size_t)i * 4];
    const double gd = softplus((double)o[(size_t)i * 8]) * dis[i];
    gv[pd] = gd;
    tv[tpos[pd]] = gd;
    for (int d = 0; d < 3; ++d) {
        const int p = gdir[(size_t)i * 4 + 1 + d];
        if (p < 0) continue;
        const double g = (double)o[(size_t)i * 8 + 1 + d] * dis[gci[p]];
        gv[p] = g;
        tv[tpos[p]] = g;
    }
}

// g_p = scale * sum_c (GW_c[i] R_c[j] + GR_c[i] W_c[j]) for the entries p =
// (i, j) of G (probe c at offset c*n), plus per-block partial sums of sum_c R^2
__global__ __launch_bounds__(kBlock) void k_fsai_gm(uint32_t n, int k, const uint32_t* __restrict__ rp,
                                                    const uint32_t* __restrict__ ci, const double* __restrict__ R,
                                                    const double* __restrict__ W, const double* __restrict__ GR,
                                                    const double* __restrict__ GW, double scale,
                                                    double* __restrict__ g, double* partials) {
    double ss = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int c = 0; c < k; ++c) ss += R[(size_t)c * n + i] * R[(size_t)c * n + i];
        for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
            const uint32_t j = ci[p];
            double s = 0.0;
            for (int c = 0; c < k; ++c)
                s += GW[(size_t)c * n + i] * R[(size_t)c * n + j] + GR[(size_t)c * n + i] * W[(size_t)c * n + j];
            g[p] = scale * s;
        }
    }
    double t[1] = {ss};
    block_partials<1>(t, partials);
}

struct FinLoss {
    double* out;
    double scale;
    __device__ void operator()(double (&t)[1]) const { *out = t[0] * scale; }
};

// dL/do from dL/dG (o[0] through softplus, o[1..3] linear; o[4..7] unused)
__global__ void k_fsai_dout(uint32_t n, const uint32_t* __restrict__ gci, const int* __restrict__ gdir,
                            const double* __restrict__ dis, const float* __restrict__ o, const double* __restrict__ g,
                            float* __restrict__ dout) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double o0 = (double)o[(size_t)i * 8];
    dout[(size_t)i * 8] = (float)(g[gdir[(size_t)i * 4]] * dis[i] / (1.0 + exp(-o0)));
    for (int d = 0; d < 3; ++d) {
        const int p = gdir[(size_t)i * 4 + 1 + d];
        dout[(size_t)i * 8 + 1 + d] = p < 0 ? 0.f : (float)(g[p] * dis[gci[p]]);
    }
    for (int d = 4; d < 8; ++d) dout[(size_t)i * 8 + d] = 0.f;
}

inline unsigned g1(size_t n) { return (unsigned)((n + 255) / 256); }

Offsets offsets_of(const uint32_t* grid) {
    if (!grid || grid[0] == 0 || grid[1] == 0 || grid[2] == 0) fail(FEM_ERR_INVALID_ARGUMENT, "grid dimensions must be positive");
    const long long nx = grid[0], ny = grid[1];
    return Offsets{{-nx * ny, -nx, -1, 1, nx, nx * ny}};
}

}  // namespace

// ---------------------------------------------------------------------------
// network state
// ---------------------------------------------------------------------------
MlpNet::MlpNet(hipStream_t s) : stream(s) {
    P.alloc(kMlpParams);
    m.alloc(kMlpParams);
    v.alloc(kMlpParams);
    g.alloc(kMlpParams);
    frag.alloc((size_t)NFRAG * 64 * 4 + (size_t)NFRAG8 * 64 * 8);
    vec_fill<float>(m.get(), 0.f, kMlpParams, s);
    vec_fill<float>(v.get(), 0.f, kMlpParams, s);
}

void MlpNet::set(const float* host) {
    P.upload(host, kMlpParams, stream);
    FEM_HIP(hipStreamSynchronize(stream));
    packed = false;
}

void MlpNet::get(float* host) const {
    P.download(host, kMlpParams, stream);
    FEM_HIP(hipStreamSynchronize(stream));
}

void MlpNet::pack() {
    if (packed) return;
    hipLaunchKernelGGL(k_pack, dim3(g1((NFRAG + NFRAG8) * 64)), dim3(256), 0, stream, P.get(), (bf16x4*)frag.get(),
                       (bf16x8*)(frag.get() + (size_t)NFRAG * 64 * 4));
    FEM_HIP(hipGetLastError());
    packed = true;
}

void MlpNet::forward(const short* feat, uint32_t n, float* out) {
    pack();
    MlpArgs a{feat, (const bf16x4*)frag.get(), (const bf16x8*)(frag.get() + (size_t)NFRAG * 64 * 4), P.get(), n, out,
              nullptr, nullptr};
    const unsigned ntiles = (n + 15) / 16;
    const unsigned blocks = std::max(1u, std::min(4096u, (ntiles + 3) / 4));
    hipLaunchKernelGGL((k_mlp<false>), dim3(blocks), dim3(256), 0, stream, a);
    FEM_HIP(hipGetLastError());
}

void MlpNet::backward(const short* feat, uint32_t n, const float* dout) {
    pack();
    const unsigned ntiles = (n + 15) / 16;
    // at least 4 tiles per wave, at most 256 block partials
    const unsigned blocks = std::max(1u, std::min(256u, (ntiles + 15) / 16));
    if (part.n < (size_t)blocks * kMlpParams) part.alloc((size_t)blocks * kMlpParams);
    MlpArgs a{feat, (const bf16x4*)frag.get(), (const bf16x8*)(frag.get() + (size_t)NFRAG * 64 * 4), P.get(), n, nullptr,
              dout, part.get()};
    hipLaunchKernelGGL((k_mlp<true>), dim3(blocks), dim3(256), 0, stream, a);
    hipLaunchKernelGGL(k_grad_reduce, dim3(g1(kMlpParams)), dim3(256), 0, stream, part.get(), blocks, 1.0, g.get());
    FEM_HIP(hipGetLastError());
}

void MlpNet::adam(double lr, double b1, double b2, double eps) {
    ++t;
    const float c1 = (float)(1.0 - std::pow(b1, (double)t)), c2 = (float)(1.0 - std::pow(b2, (double)t));
    hipLaunchKernelGGL(k_adam, dim3(g1(kMlpParams)), dim3(256), 0, stream, P.get(), m.get(), v.get(), g.get(),
                       (float)lr, (float)b1, (float)b2, (float)eps, c1, c2);
    FEM_HIP(hipGetLastError());
    packed = false;
}

// ---------------------------------------------------------------------------
// stencil analysis of A and the assembled M
// ---------------------------------------------------------------------------
void SpaiPattern::build(const CsrMat& A, const uint32_t* grid, hipStream_t s, bool keep_f32) {
    if (A.n != A.m) fail(FEM_ERR_INVALID_ARGUMENT, "learned SPAI needs a square matrix");
    const uint32_t n = A.n;
    if ((uint64_t)grid[0] * grid[1] * grid[2] != n)
        fail(FEM_ERR_INVALID_ARGUMENT, "grid dimensions do not match the matrix size");
    const Offsets G = offsets_of(grid);
    this->n = n;
    epos.alloc((size_t)n * 6);
    dpos.alloc(n);
    dis.alloc(n);
    feat.alloc((size_t)n * 8);
    if (keep_f32) feat32.alloc((size_t)n * 6);
    DevBuf<int> err(1);
    vec_fill<uint32_t>((uint32_t*)err.get(), 0u, 1, s);
    hipLaunchKernelGGL(k_spai_rows, dim3(g1(n)), dim3(256), 0, s, n, A.rp.get(), A.ci.get(), A.v.get(), G, epos.get(),
                       dpos.get(), dis.get(), err.get());
    hipLaunchKernelGGL(k_spai_features, dim3(g1(n)), dim3(256), 0, s, n, A.ci.get(), A.v.get(), epos.get(), dis.get(),
                       feat.get(), keep_f32 ? feat32.get() : nullptr, err.get());
    FEM_HIP(hipGetLastError());
    int e = 0;
    err.download(&e, 1, s);
    FEM_HIP(hipStreamSynchronize(s));
    if (e & 1) fail(FEM_ERR_INVALID_ARGUMENT, "learned SPAI: the matrix has entries outside the 7-point stencil of the grid");
    if (e & 2) fail(FEM_ERR_INVALID_ARGUMENT, "learned SPAI: every row needs a positive diagonal entry");
    if (e & 4) fail(FEM_ERR_INVALID_ARGUMENT, "learned SPAI: the sparsity pattern must be symmetric");
    // G on A's lower pattern (with the diagonal), G^T on the upper one
    A.sync_host();
    std::vector<int> hd(n), he((size_t)n * 6);
    dpos.download(hd.data(), n, s);
    epos.download(he.data(), (size_t)n * 6, s);
    FEM_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> grp(n + 1, 0), gci, trp(n + 1, 0), tci;
    std::vector<int> a2g(A.nnz, -1), a2t(A.nnz, -1);
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t p = A.h_rp[i]; p < A.h_rp[i + 1]; ++p) {
            const uint32_t j = A.h_ci[p];
            if (j <= i) { a2g[p] = (int)gci.size(); gci.push_back(j); }
            if (j >= i) { a2t[p] = (int)tci.size(); tci.push_back(j); }
        }
        grp[i + 1] = (uint32_t)gci.size();
        trp[i + 1] = (uint32_t)tci.size();
    }
    std::vector<int> gdir((size_t)n * 4, -1), tp(gci.size(), -1);
    for (uint32_t i = 0; i < n; ++i) {
        gdir[(size_t)i * 4] = a2g[hd[i]];
        tp[a2g[hd[i]]] = a2t[hd[i]];
        for (int d = 0; d < 3; ++d) {
            const int p = he[(size_t)i * 6 + d];
            if (p < 0) continue;
            // the mirrored entry (j, i) is row j's neighbour in direction 5 - d
            const uint32_t j = A.h_ci[p];
            const int q = he[(size_t)j * 6 + (5 - d)];
            gdir[(size_t)i * 4 + 1 + d] = a2g[p];
            tp[a2g[p]] = a2t[q];
        }
    }
    const size_t gnz = gci.size();
    this->G.set_host(n, n, std::move(grp), std::move(gci), std::vector<double>(gnz, 0.0), s);
    this->Gt.set_host(n, n, std::move(trp), std::move(tci), std::vector<double>(gnz, 0.0), s);
    this->gdir.upload(gdir, s);
    tpos.upload(tp, s);
    out.alloc((size_t)n * 8);
    FEM_HIP(hipStreamSynchronize(s));
}

void SpaiPattern::assemble(MlpNet& net, hipStream_t s) {
    net.forward(feat.get(), n, out.get());
    hipLaunchKernelGGL(k_fsai_assemble, dim3(g1(n)), dim3(256), 0, s, n, G.ci.get(), gdir.get(), tpos.get(), dis.get(),
                       out.get(), G.v.get(), Gt.v.get());
    FEM_HIP(hipGetLastError());
    if (G.v32.get()) vec_convert<float, double>(G.v32.get(), G.v.get(), 1.0, G.nnz, s);
    if (Gt.v32.get()) vec_convert<float, double>(Gt.v32.get(), Gt.v.get(), 1.0, Gt.nnz, s);
}

// loss and gradient (left in net.g) for k probe vectors V (device, k x n)
double spai_loss_grad(MlpNet& net, const CsrMat& A, SpaiPattern& S, const double* V, int k, RedScratch& red,
                      hipStream_t s) {
    const uint32_t n = S.n;
    if (k < 1) fail(FEM_ERR_INVALID_ARGUMENT, "at least one probe vector is needed");
    S.assemble(net, s);
    const size_t nk = (size_t)n * k;
    DevBuf<double> W(nk), R(nk), GW(nk), GR(nk), g(S.G.nnz), out(1);
    DevBuf<float> dout((size_t)n * 8);
    for (int c = 0; c < k; ++c) {
        const size_t o = (size_t)c * n;
        csr_spmv<double>(A, V + o, W.get() + o, EPI_STORE, {}, s, Gate{});
        csr_spmv<double>(S.G, W.get() + o, GW.get() + o, EPI_STORE, {}, s, Gate{});
        SpmvArgs<double> ar;
        ar.f = V + o;
        csr_spmv<double>(S.Gt, GW.get() + o, R.get() + o, EPI_RESID, ar, s, Gate{});  // R = v - G^T G A v
        csr_spmv<double>(S.G, R.get() + o, GR.get() + o, EPI_STORE, {}, s, Gate{});
    }
    const double sc = 1.0 / ((double)n * k);
    const unsigned G = red_blocks(n);
    hipLaunchKernelGGL(k_fsai_gm, dim3(G), dim3(kBlock), 0, s, n, k, S.G.rp.get(), S.G.ci.get(), R.get(), W.get(),
                       GR.get(), GW.get(), -2.0 * sc, g.get(), red.partials.get());
    launch_finalize<1>(red.partials.get(), G, FinLoss{out.get(), sc}, s);
    hipLaunchKernelGGL(k_fsai_dout, dim3(g1(n)), dim3(256), 0, s, n, S.G.ci.get(), S.gdir.get(), S.dis.get(),
                       S.out.get(), g.get(), dout.get());
    FEM_HIP(hipGetLastError());
    net.backward(S.feat.get(), n, dout.get());
    double L = 0.0;
    out.download(&L, 1, s);
    FEM_HIP(hipStreamSynchronize(s));
    return L;
}

// ---------------------------------------------------------------------------
// the preconditioner
// ---------------------------------------------------------------------------
// The apply of a learned G (PcLearned) and of a given sparse approximate
// inverse (PcExplicitSpai: M, or G with M = G^T G).
struct PcFactorized : Precond {
    const CsrMat* Gm = nullptr;   // G (lower), or M when !factorized
    const CsrMat* Gtm = nullptr;  // G^T
    bool factorized = true;
    uint32_t n = 0;
    DevBuf<double> t64;
    DevBuf<float> t32;  // G r
    bool fp32 = false;
    uint32_t size() const override { return n; }
    // z = G^T (G r) as two CSR-stream products through t; with DOT the first
    // also leaves the partial sums of t_i^2 (||G r||^2 = r^T M r = r.z)
    template <typename T, bool DOT>
    unsigned two_products(const T* r, T* z, double* part, hipStream_t s, Gate g) {
        T* t;
        if constexpr (std::is_same<T, double>::value) t = t64.get();
        else t = t32.get();
        unsigned np = 0;
        if constexpr (DOT) np = csr_spmv_selfdot<T>(*Gm, r, t, part, s, g);
        else csr_spmv<T>(*Gm, r, t, EPI_STORE, {}, s, g);
        csr_spmv<T>(*Gtm, t, z, EPI_STORE, {}, s, g);
        return np;
    }
    void apply(const double* r, double* z, hipStream_t s, Gate g) override {
        if (!factorized) { csr_spmv<double>(*Gm, r, z, EPI_STORE, {}, s, g); return; }
        two_products<double, false>(r, z, nullptr, s, g);
    }
    void apply32(const float* r, float* z, hipStream_t s, Gate g) override {
        if (!factorized) { csr_spmv<float>(*Gm, r, z, EPI_STORE, {}, s, g); return; }
        two_products<float, false>(r, z, nullptr, s, g);
    }
    // r.z from the apply's own partial sums
    unsigned apply_dot(const double* r, double* z, double* part, hipStream_t s, Gate g) override {
        return factorized ? two_products<double, true>(r, z, part, s, g) : csr_spmv_dot<double>(*Gm, r, z, part, s, g);
    }
    unsigned apply32_dot(const float* r, float* z, double* part, hipStream_t s, Gate g) override {
        return factorized ? two_products<float, true>(r, z, part, s, g) : csr_spmv_dot<float>(*Gm, r, z, part, s, g);
    }
    unsigned dot_parts() const override { return Gm->dot_parts(); }
    bool factors(const CsrMat*& G, const CsrMat*& Gt) const override {
        if (!factorized) return false;
        G = Gm;
        Gt = Gtm;
        return true;
    }
    bool has_fp32() const override { return fp32; }
    void ensure_fp32(hipStream_t s) override {
        if (fp32) return;
        const_cast<CsrMat*>(Gm)->ensure_fp32(s);
        if (factorized) {
            const_cast<CsrMat*>(Gtm)->ensure_fp32(s);
            t32.alloc(n);
        }
        FEM_HIP(hipStreamSynchronize(s));
        fp32 = true;
    }
    double apply_bytes(int sz) override {
        return factorized ? csr_bytes(*Gm, sz, 0) + csr_bytes(*Gtm, sz, 0) : csr_bytes(*Gm, sz, 0);
    }
};

struct PcLearned : PcFactorized {
    SpaiPattern S;
    CsrMat* A = nullptr;
    int fb_kind = FEM_PC_IC0;
    int budget_ = 0;
    int ref_iters = 0;  // the fallback's iteration count on this matrix (0: unknown)
    std::unique_ptr<Precond> fb;
    void bind() {
        Gm = &S.G;
        Gtm = &S.Gt;
        factorized = true;
        n = S.n;
        t64.alloc(n);
    }
    int kind() const override { return FEM_PC_LEARNED; }
    // a learned solve stops once it needs more iterations than the explicit
    // budget or than the fallback needed (an iteration-count regression)
    int budget() const override {
        if (budget_ > 0 && ref_iters > 0) return std::min(budget_, ref_iters);
        return budget_ > 0 ? budget_ : ref_iters;
    }
    void fallback_ran(int iterations) override {
        if (iterations > 0) ref_iters = iterations;
    }
    Precond* fallback(hipStream_t s) override {
        if (!fb) {
            fem_pc_params p;
            fem_pc_params_default(fb_kind, &p);
            fb = make_precond(*A, fb_kind, p, s);
        }
        return fb.get();
    }
};

struct PcExplicitSpai : PcFactorized {
    CsrMat Gt_own;  // G^T (factorized)
    int kind() const override { return FEM_PC_SPAI; }
};

}  // namespace fem

struct fem_mlp_s {
    hipStream_t stream = nullptr;
    std::unique_ptr<fem::MlpNet> net;
    fem::RedScratch red;
    ~fem_mlp_s() {
        net.reset();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

using namespace fem;

extern "C" {

int fem_mlp_create(const float* params, uint64_t seed, fem_mlp* out) {
    return api_guard([&] {
        if (!out) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        ensure_device();
        auto h = std::make_unique<fem_mlp_s>();
        h->stream = new_stream();
        h->net.reset(new MlpNet(h->stream));
        h->red.init(h->stream);
        std::vector<float> P(kMlpParams, 0.f);
        if (params) {
            std::copy(params, params + kMlpParams, P.begin());
        } else {
            // He initialisation of the weights, zero biases, from seed
            std::mt19937_64 rng(seed);
            std::normal_distribution<double> nd(0.0, 1.0);
            for (int i = 0; i < NI * NH; ++i) P[OFF_W1 + i] = (float)(nd(rng) * std::sqrt(2.0 / NI));
            for (int i = 0; i < NH * NH; ++i) P[OFF_W2 + i] = (float)(nd(rng) * std::sqrt(2.0 / NH));
            for (int i = 0; i < NH * NO; ++i) P[OFF_W3 + i] = (float)(nd(rng) * std::sqrt(1.0 / NH) * 0.1);
        }
        h->net->set(P.data());
        *out = h.release();
    });
}

int fem_mlp_destroy(fem_mlp h) {
    return api_guard([&] { delete h; });
}

int fem_mlp_get_params(fem_mlp h, float* params) {
    return api_guard([&] {
        if (!h || !params) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        h->net->get(params);
    });
}

int fem_mlp_set_params(fem_mlp h, const float* params) {
    return api_guard([&] {
        if (!h || !params) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        h->net->set(params);
    });
}

int fem_mlp_forward(fem_mlp h, const float* feat, uint32_t n, float* out) {
    return api_guard([&] {
        if (!h || (!feat && n) || (!out && n)) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        if (!n) return;
        hipStream_t s = h->stream;
        DevBuf<float> f((size_t)n * NI), o((size_t)n * 8);
        DevBuf<short> fb((size_t)n * 8);
        f.upload(feat, (size_t)n * NI, s);
        hipLaunchKernelGGL(k_feat_pack, dim3(g1(n)), dim3(256), 0, s, f.get(), n, fb.get());
        h->net->forward(fb.get(), n, o.get());
        std::vector<float> ho((size_t)n * 8);
        o.download(ho.data(), ho.size(), s);
        FEM_HIP(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < n; ++i)
            for (int c = 0; c < NO; ++c) out[(size_t)i * NO + c] = ho[(size_t)i * 8 + c];
    });
}

int fem_spai_features(fem_csr A, const uint32_t* grid, float* feat) {
    return api_guard([&] {
        if (!A || !grid || !feat) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        SpaiPattern S;
        S.build(A->A, grid, A->stream, true);
        S.feat32.download(feat, (size_t)S.n * NI, A->stream);
        FEM_HIP(hipStreamSynchronize(A->stream));
    });
}

static const double* probes_on_device(fem_csr A, const double* V, int k, int device_ptrs, DevBuf<double>& tmp) {
    if (!V) fail(FEM_ERR_INVALID_ARGUMENT, "null probe vectors");
    if (k < 1) fail(FEM_ERR_INVALID_ARGUMENT, "at least one probe vector is needed");
    if (device_ptrs) return V;
    const size_t len = (size_t)A->A.n * k;
    tmp.alloc(len);
    tmp.upload(V, len, A->stream);
    return tmp.get();
}

// The network's stream is used for everything below: the matrix's stream
// only ran work that finished before its handle returned.
int fem_spai_loss_grad(fem_mlp h, fem_csr A, const uint32_t* grid, const double* V, int k, int device_ptrs,
                       double* loss, float* grad) {
    return api_guard([&] {
        if (!h || !A || !grid) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        hipStream_t s = h->stream;
        DevBuf<double> tmp;
        const double* Vd = probes_on_device(A, V, k, device_ptrs, tmp);
        SpaiPattern S;
        S.build(A->A, grid, s, false);
        const double L = spai_loss_grad(*h->net, A->A, S, Vd, k, h->red, s);
        if (loss) *loss = L;
        if (grad) {
            h->net->g.download(grad, kMlpParams, s);
            FEM_HIP(hipStreamSynchronize(s));
        }
    });
}

int fem_spai_train_step(fem_mlp h, fem_csr A, const uint32_t* grid, const double* V, int k, int device_ptrs,
                        const fem_adam_opts* opt, double* loss) {
    return api_guard([&] {
        if (!h || !A || !grid || !opt) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        if (!(opt->lr > 0) || !(opt->beta1 >= 0 && opt->beta1 < 1) || !(opt->beta2 >= 0 && opt->beta2 < 1) ||
            !(opt->eps > 0))
            fail(FEM_ERR_INVALID_ARGUMENT, "invalid Adam parameters");
        hipStream_t s = h->stream;
        DevBuf<double> tmp;
        const double* Vd = probes_on_device(A, V, k, device_ptrs, tmp);
        SpaiPattern S;
        S.build(A->A, grid, s, false);
        const double L = spai_loss_grad(*h->net, A->A, S, Vd, k, h->red, s);
        h->net->adam(opt->lr, opt->beta1, opt->beta2, opt->eps);
        FEM_HIP(hipStreamSynchronize(s));
        if (loss) *loss = L;
    });
}

int fem_precond_create_learned(fem_csr A, const uint32_t* grid, fem_mlp net, int fallback_kind, int budget,
                               fem_pc* out) {
    return api_guard([&] {
        if (!A || !grid || !net || !out) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        if (fallback_kind != FEM_PC_IC0 && fallback_kind != FEM_PC_AMG && fallback_kind != FEM_PC_SAAMG &&
            fallback_kind != FEM_PC_DIAG && fallback_kind != FEM_PC_ILU0)
            fail(FEM_ERR_INVALID_ARGUMENT, "fallback must be diag, ic0, ilu0, amg or saamg");
        if (budget < 0) fail(FEM_ERR_INVALID_ARGUMENT, "budget must be >= 0");
        Timer tm;
        hipStream_t s = A->stream;
        auto pc = std::make_unique<PcLearned>();
        pc->A = &A->A;
        pc->fb_kind = fallback_kind;
        pc->budget_ = budget;
        pc->S.build(A->A, grid, s, false);
        pc->bind();
        // the network runs on its own stream: order it after the analysis
        FEM_HIP(hipStreamSynchronize(s));
        pc->S.assemble(*net->net, net->stream);
        FEM_HIP(hipStreamSynchronize(net->stream));
        pc->setup_ms = tm.ms();
        auto h = std::make_unique<fem_pc_s>();
        h->stream = s;
        h->pc = std::move(pc);
        *out = h.release();
    });
}

int fem_precond_learned_reference(fem_pc M, int iterations) {
    return api_guard([&] {
        auto* p = M ? dynamic_cast<PcLearned*>(M->pc.get()) : nullptr;
        if (!p) fail(FEM_ERR_INVALID_ARGUMENT, "not a learned preconditioner");
        if (iterations < 0) fail(FEM_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
        p->ref_iters = iterations;
    });
}

int fem_precond_create_spai(fem_csr A, fem_csr Mat, int factorized, fem_pc* out) {
    return api_guard([&] {
        if (!A || !Mat || !out) fail(FEM_ERR_INVALID_ARGUMENT, "null argument");
        if (Mat->A.n != A->A.n || Mat->A.m != A->A.n)
            fail(FEM_ERR_INVALID_ARGUMENT, "the approximate inverse must be square and sized as A");
        Timer tm;
        hipStream_t s = A->stream;
        auto pc = std::make_unique<PcExplicitSpai>();
        pc->n = A->A.n;
        pc->Gm = &Mat->A;
        pc->factorized = factorized != 0;
        if (pc->factorized) {
            transpose_device(Mat->A, pc->Gt_own, s);
            pc->Gtm = &pc->Gt_own;
            pc->t64.alloc(pc->n);
        }
        FEM_HIP(hipStreamSynchronize(s));
        pc->setup_ms = tm.ms();
        auto h = std::make_unique<fem_pc_s>();
        h->stream = s;
        h->pc = std::move(pc);
        *out = h.release();
    });
}

}  // extern "C"

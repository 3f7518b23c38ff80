// spai_oracle.cpp — CPU restatement (test infrastructure only) of the
// learned sparse approximate inverse: the local features, the MLP forward
// and backward pass, the factorized SPD approximate inverse M = G^T G with G
// on the lower triangle of A's pattern, and the training loss with its
// gradient. No counterpart in the reference; the product implements the same
// definitions in cpp-fem-solver_amd/csrc/learned.hip.
//
// Definitions (7-point grid matrix, rows (z*ny + y)*nx + x):
//   directions d = 0..5 <-> column offsets -nx*ny, -nx, -1, +1, +nx, +nx*ny;
//   the opposite of d is 5 - d; d = 0, 1, 2 are the lower neighbours.
//   dis_i = 1 / sqrt(a_ii).
//   features  f_i[d] = -6 a_ij dis_i dis_j for the neighbour j in direction d
//             (0 when row i has no entry there)
//   network   o_i = W3^T relu(W2^T relu(W1^T f_i + b1) + b2) + b3, 6 -> 64 ->
//             64 -> 7, parameters float32, flat layout W1[6][64], b1[64],
//             W2[64][64], b2[64], W3[64][7], b3[7] (W[in][out])
//   G         lower triangular on A's pattern:
//             G_ii = softplus(o_i[0]) dis_i, G_ij = o_i[1+d] dis_j for the
//             lower neighbour j in direction d (outputs 4..6 are unused);
//             M = G^T G (an approximation of A^{-1} = D^{-1/2} Ahat^{-1}
//             D^{-1/2} through Ghat = G D^{1/2}, Ghat^T Ghat ~ Ahat^{-1}); G
//             has a positive diagonal, so M is SPD for any network output
//   loss      L = 1/(n k) sum_c || v_c - G^T G A v_c ||^2 over k probe vectors
//   gradient  with R_c = v_c - M A v_c and W_c = A v_c,
//             dL/dG_ij = -2/(n k) sum_c ((G W_c)_i R_c,j + (G R_c)_i W_c,j)
//             on G's pattern (from dM = dG^T G + G^T dG).
// With bf16 = 1 the matrix-product operands are rounded to bfloat16 (round to
// nearest even) exactly where the device kernels round them (features,
// weights, hidden activations, and in the backward pass dL/do and the masked
// hidden gradients); sums stay in double.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr int NI = 6, NH = 64, NO = 7;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NI * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NO, NPARAM = OFF_B3 + NO;

double bf(double x, bool on) {
    if (!on) return x;
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

struct Grid {
    int64_t off[6];
    explicit Grid(const uint32_t* g) {
        const int64_t nx = g[0], ny = g[1];
        const int64_t o[6] = {-nx * ny, -nx, -1, 1, nx, nx * ny};
        for (int d = 0; d < 6; ++d) off[d] = o[d];
    }
    int dir(int64_t delta) const {
        for (int d = 0; d < 6; ++d)
            if (off[d] == delta) return d;
        return -1;
    }
};

struct Mat {
    uint32_t n;
    const uint32_t *rp, *ci;
    const double* v;
};

// dis_i, features, diagonal positions and the per-direction entry positions
void analyse(const Mat& A, const Grid& G, std::vector<double>& dis, std::vector<double>& feat,
             std::vector<int64_t>& dpos, std::vector<int64_t>& epos) {
    const uint32_t n = A.n;
    dis.assign(n, 0.0);
    dpos.assign(n, -1);
    epos.assign((size_t)n * 6, -1);
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            const uint32_t j = A.ci[p];
            if (j == i) {
                dpos[i] = p;
                continue;
            }
            const int d = G.dir((int64_t)j - (int64_t)i);
            if (d < 0) throw std::invalid_argument("learned SPAI: row " + std::to_string(i) + " has an entry outside the 7-point stencil");
            epos[(size_t)i * 6 + d] = p;
        }
        if (dpos[i] < 0 || !(A.v[dpos[i]] > 0.0))
            throw std::invalid_argument("learned SPAI: row " + std::to_string(i) + " needs a positive diagonal");
        dis[i] = 1.0 / std::sqrt(A.v[dpos[i]]);
    }
    feat.assign((size_t)n * NI, 0.0);
    for (uint32_t i = 0; i < n; ++i)
        for (int d = 0; d < 6; ++d) {
            const int64_t p = epos[(size_t)i * 6 + d];
            if (p < 0) continue;
            const uint32_t j = A.ci[p];
            if (epos[(size_t)j * 6 + (5 - d)] < 0) throw std::invalid_argument("learned SPAI: the pattern must be symmetric");
            feat[(size_t)i * NI + d] = -6.0 * A.v[p] * dis[i] * dis[j];
        }
}

struct Fwd {
    std::vector<double> z1, h1, z2, h2, o;  // per row
};

void forward(const float* P, const std::vector<double>& feat, uint32_t n, bool b16, Fwd& F) {
    F.z1.assign((size_t)n * NH, 0.0);
    F.h1.assign((size_t)n * NH, 0.0);
    F.z2.assign((size_t)n * NH, 0.0);
    F.h2.assign((size_t)n * NH, 0.0);
    F.o.assign((size_t)n * NO, 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        double x[NI];
        for (int f = 0; f < NI; ++f) x[f] = bf(feat[(size_t)i * NI + f], b16);
        for (int h = 0; h < NH; ++h) {
            double s = 0.0;
            for (int f = 0; f < NI; ++f) s += x[f] * bf(P[OFF_W1 + f * NH + h], b16);
            s += P[OFF_B1 + h];
            F.z1[(size_t)i * NH + h] = s;
            F.h1[(size_t)i * NH + h] = s > 0.0 ? s : 0.0;
        }
        for (int h = 0; h < NH; ++h) {
            double s = 0.0;
            for (int q = 0; q < NH; ++q) s += bf(F.h1[(size_t)i * NH + q], b16) * bf(P[OFF_W2 + q * NH + h], b16);
            s += P[OFF_B2 + h];
            F.z2[(size_t)i * NH + h] = s;
            F.h2[(size_t)i * NH + h] = s > 0.0 ? s : 0.0;
        }
        for (int c = 0; c < NO; ++c) {
            double s = 0.0;
            for (int q = 0; q < NH; ++q) s += bf(F.h2[(size_t)i * NH + q], b16) * bf(P[OFF_W3 + q * NO + c], b16);
            F.o[(size_t)i * NO + c] = s + P[OFF_B3 + c];
        }
    }
}

double softplus(double x) { return x > 30.0 ? x : std::log1p(std::exp(x)); }
double sigmoid(double x) { return 1.0 / (1.0 + std::exp(-x)); }

// G's values on A's pattern (entries above the diagonal stay 0)
void assemble(const Mat& A, const std::vector<double>& dis, const std::vector<int64_t>& dpos,
              const std::vector<int64_t>& epos, const std::vector<double>& o, std::vector<double>& g) {
    g.assign(A.rp[A.n], 0.0);
    for (uint32_t i = 0; i < A.n; ++i) {
        for (int d = 0; d < 3; ++d) {
            const int64_t p = epos[(size_t)i * 6 + d];
            if (p >= 0) g[p] = o[(size_t)i * NO + 1 + d] * dis[A.ci[p]];
        }
        g[dpos[i]] = softplus(o[(size_t)i * NO]) * dis[i];
    }
}

void spmv(const Mat& A, const double* vals, const double* x, double* y) {
    for (uint32_t i = 0; i < A.n; ++i) {
        double s = 0.0;
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) s += vals[p] * x[A.ci[p]];
        y[i] = s;
    }
}

// y = G^T x for G stored on A's pattern
void spmv_t(const Mat& A, const double* vals, const double* x, double* y) {
    for (uint32_t j = 0; j < A.n; ++j) y[j] = 0.0;
    for (uint32_t i = 0; i < A.n; ++i)
        for (uint32_t p = A.rp[i]; p < A.rp[i + 1]; ++p) y[A.ci[p]] += vals[p] * x[i];
}

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument&) {
        return 1;
    } catch (...) {
        return 4;
    }
}

}  // namespace

extern "C" {

int or_spai_nparam(void) { return NPARAM; }

// features (n x 6, row-major) and dis (n)
int or_spai_features(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* v, const uint32_t* grid,
                     double* feat, double* dis) {
    return guard([&] {
        Mat A{n, rp, ci, v};
        std::vector<double> d, f;
        std::vector<int64_t> dp, ep;
        analyse(A, Grid(grid), d, f, dp, ep);
        std::copy(f.begin(), f.end(), feat);
        std::copy(d.begin(), d.end(), dis);
    });
}

// network outputs (n x 7, row-major) for the given features
int or_mlp_forward(const float* params, uint32_t n, const double* feat, int bf16, double* out) {
    return guard([&] {
        std::vector<double> f(feat, feat + (size_t)n * NI);
        Fwd F;
        forward(params, f, n, bf16 != 0, F);
        std::copy(F.o.begin(), F.o.end(), out);
    });
}

// values of G on A's pattern (nnz; zero above the diagonal)
int or_spai_assemble(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* v, const uint32_t* grid,
                     const float* params, int bf16, double* mvals) {
    return guard([&] {
        Mat A{n, rp, ci, v};
        std::vector<double> dis, feat, m;
        std::vector<int64_t> dp, ep;
        analyse(A, Grid(grid), dis, feat, dp, ep);
        Fwd F;
        forward(params, feat, n, bf16 != 0, F);
        assemble(A, dis, dp, ep, F.o, m);
        std::copy(m.begin(), m.end(), mvals);
    });
}

// loss and its gradient with respect to the NPARAM parameters; V is n x k
// column-major (probe c at V + c*n)
int or_spai_loss_grad(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* v, const uint32_t* grid,
                      const float* params, const double* V, int k, int bf16, double* loss, double* grad) {
    return guard([&] {
        const bool b16 = bf16 != 0;
        Mat A{n, rp, ci, v};
        std::vector<double> dis, feat, m;
        std::vector<int64_t> dp, ep;
        analyse(A, Grid(grid), dis, feat, dp, ep);
        Fwd F;
        forward(params, feat, n, b16, F);
        assemble(A, dis, dp, ep, F.o, m);
        const uint64_t nnz = rp[n];
        std::vector<double> W((size_t)n * k), R((size_t)n * k), GW((size_t)n * k), GR((size_t)n * k), g(nnz, 0.0),
            t(n);
        double L = 0.0;
        for (int c = 0; c < k; ++c) {
            double* Wc = &W[(size_t)c * n];
            double* Rc = &R[(size_t)c * n];
            spmv(A, v, V + (size_t)c * n, Wc);
            spmv(A, m.data(), Wc, &GW[(size_t)c * n]);
            spmv_t(A, m.data(), &GW[(size_t)c * n], t.data());
            for (uint32_t i = 0; i < n; ++i) {
                Rc[i] = V[(size_t)c * n + i] - t[i];
                L += Rc[i] * Rc[i];
            }
            spmv(A, m.data(), Rc, &GR[(size_t)c * n]);
        }
        const double sc = 1.0 / ((double)n * k);
        *loss = L * sc;
        // dL/dG_ij on the lower pattern
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t p = rp[i]; p < rp[i + 1]; ++p) {
                const uint32_t j = ci[p];
                if (j > i) continue;
                double s = 0.0;
                for (int c = 0; c < k; ++c)
                    s += GW[(size_t)c * n + i] * R[(size_t)c * n + j] + GR[(size_t)c * n + i] * W[(size_t)c * n + j];
                g[p] = -2.0 * sc * s;
            }
        // dL/do
        std::vector<double> dO((size_t)n * NO, 0.0);
        for (uint32_t i = 0; i < n; ++i) {
            dO[(size_t)i * NO] = g[dp[i]] * dis[i] * sigmoid(F.o[(size_t)i * NO]);
            for (int d = 0; d < 3; ++d) {
                const int64_t p = ep[(size_t)i * 6 + d];
                if (p >= 0) dO[(size_t)i * NO + 1 + d] = g[p] * dis[ci[p]];
            }
        }
        // backward through the network
        std::vector<double> G(NPARAM, 0.0);
        for (uint32_t i = 0; i < n; ++i) {
            double go[NO], dh2[NH], dz2[NH], dh1[NH], dz1[NH];
            for (int c = 0; c < NO; ++c) {
                G[OFF_B3 + c] += dO[(size_t)i * NO + c];
                go[c] = bf(dO[(size_t)i * NO + c], b16);
            }
            for (int q = 0; q < NH; ++q) {
                const double h2 = bf(F.h2[(size_t)i * NH + q], b16);
                double s = 0.0;
                for (int c = 0; c < NO; ++c) {
                    G[OFF_W3 + q * NO + c] += h2 * go[c];
                    s += go[c] * bf(params[OFF_W3 + q * NO + c], b16);
                }
                dh2[q] = s;
                dz2[q] = F.z2[(size_t)i * NH + q] > 0.0 ? s : 0.0;
                G[OFF_B2 + q] += dz2[q];
            }
            for (int q = 0; q < NH; ++q) {
                const double h1 = bf(F.h1[(size_t)i * NH + q], b16);
                double s = 0.0;
                for (int h = 0; h < NH; ++h) {
                    const double gz = bf(dz2[h], b16);
                    G[OFF_W2 + q * NH + h] += h1 * gz;
                    s += gz * bf(params[OFF_W2 + q * NH + h], b16);
                }
                dh1[q] = s;
                dz1[q] = F.z1[(size_t)i * NH + q] > 0.0 ? s : 0.0;
                G[OFF_B1 + q] += dz1[q];
            }
            for (int f = 0; f < NI; ++f) {
                const double x = bf(feat[(size_t)i * NI + f], b16);
                for (int h = 0; h < NH; ++h) G[OFF_W1 + f * NH + h] += x * bf(dz1[h], b16);
            }
            (void)dh2;
            (void)dh1;
        }
        std::copy(G.begin(), G.end(), grad);
    });
}

}  // extern "C"

// learned.hpp — learned sparse approximate inverse (learned.hip).
#pragma once
#include "internal.hpp"

namespace fem {

constexpr int kMlpParams = 5063;  // 6*64 + 64 + 64*64 + 64 + 64*7 + 7

// MLP parameters (float32 master copy), Adam moments, the last gradient and
// the packed bfloat16 MFMA fragments of the weights
struct MlpNet {
    hipStream_t stream;
    DevBuf<float> P, m, v, g, part;
    DevBuf<short> frag;
    long long t = 0;  // Adam step count
    bool packed = false;
    explicit MlpNet(hipStream_t s);
    void set(const float* host);
    void get(float* host) const;
    void pack();
    // out: n x 8 float (outputs 0..6) for bfloat16 features feat (n x 8)
    void forward(const short* feat, uint32_t n, float* out);
    // g = d/dparams of sum_i dout_i . o_i
    void backward(const short* feat, uint32_t n, const float* dout);
    void adam(double lr, double b1, double b2, double eps);
};

// per-matrix data of the learned SPAI: entry positions per direction, the
// diagonal scaling, the network's features and outputs, G on A's lower
// pattern and G^T on its upper one (gdir: per row the G positions of the
// diagonal and of the three lower neighbours; tpos: G position -> G^T
// position of the same entry)
struct SpaiPattern {
    uint32_t n = 0;
    DevBuf<int> epos, dpos, gdir, tpos;
    DevBuf<double> dis;
    DevBuf<short> feat;
    DevBuf<float> feat32, out;
    CsrMat G, Gt;
    void build(const CsrMat& A, const uint32_t* grid, hipStream_t s, bool keep_f32);
    void assemble(MlpNet& net, hipStream_t s);
};

double spai_loss_grad(MlpNet& net, const CsrMat& A, SpaiPattern& S, const double* V, int k, RedScratch& red,
                      hipStream_t s);

}  // namespace fem

// amg.hpp — AMG hierarchies: the host-built SA description (amg_setup.cpp),
// the device-built C/F and SA hierarchies (amg_device.hip), and the level
// data the device V-cycle (amg.hip) runs on.
#pragma once
#include <memory>

#include "internal.hpp"

namespace fem {

struct HCsr {
    uint32_t n = 0, m = 0;
    std::vector<uint32_t> rp{0}, ci;
    std::vector<double> v;
};

HCsr transpose(const HCsr& A);
HCsr spgemm(const HCsr& A, const HCsr& B);

// Multicolour Gauss-Seidel runs one launch per colour; a level needing more
// colours than this is smoothed by damped Jacobi instead (coarse SA operators
// with 27+ entries per row colour greedily into 15-40 colours).
constexpr int kMaxGsColours = 8;

// A hierarchy as the host setups build it.
struct AmgHierarchy {
    std::vector<HCsr> A;                  // level operators, A[0] = input
    std::vector<HCsr> P;                  // prolongators, P[k]: n_k x n_{k+1}
    std::vector<std::vector<double>> d;   // smoother diagonal (C/F: a_ii, SA: 1/a_ii)
    std::vector<double> coarse;           // dense coarse solution operator (row-major)
    std::vector<double> omega;            // smoother weight per level
};

// One level on the device: the operators, the smoother data and the V-cycle's
// work vectors.
struct AmgLevelDev {
    CsrMat A, P, R;  // P, R: none on the coarsest level
    uint32_t n = 0, nc = 0;
    double omega = 0.0;  // smoother weight
    DevBuf<double> d;      // C/F: diagonal; SA: inverse diagonal
    DevBuf<float> d32;
    DevBuf<double> f, u, u2, res;
    DevBuf<float> f32, u32, u232, res32;
    // SGS smoother: greedy multicolour ordering (colour of row i = the least
    // colour that no lower-numbered neighbour in pattern(A + A^T) holds)
    int ncol = 0;         // colours; kMaxGsColours + 1 stands for "more"
    bool use_gs = false;  // colour-by-colour GS on this level
    double jw = 0.0;      // Jacobi weight (4/3) / max_i sum_j |a_ij / a_ii| when !use_gs
    std::vector<uint32_t> h_cptr;
    DevBuf<uint32_t> cptr, crows;
    DevBuf<double> diag;  // a_ii (natural order)
    DevBuf<float> diag32;
    // colour-grouped copy of A for the GS sweeps; gblk: CSR-stream blocks of
    // the grouped rows, colour c owning blocks [h_cblk[c], h_cblk[c+1]);
    // h_cblk32/gblk32 the same for the float32 sweeps (larger blocks)
    int lanes = 1;
    std::vector<uint32_t> h_cblk, h_cblk32;
    DevBuf<uint2> gblk, gblk32;
    DevBuf<uint32_t> gp, gci;
    std::vector<uint32_t> h_gp;  // gp on the host (for the float32 blocking)
    DevBuf<uint16_t> gcd;     // packed gci (empty: not packable)
    // R without the columns of the last colour, whose residual is zero after
    // a forward sweep of weight 1 (rows 0 when that shortcut does not apply)
    CsrMat Rz;
    DevBuf<uint32_t> gcbase;
    DevBuf<double> gv, gd;
    DevBuf<float> gv32, gd32;
    // Colour-ordered form of the level (float32 inner solves on the finest
    // level): position t holds row crows[t]; O holds the off-diagonal entries
    // of A in that row order with columns given as positions (the diagonal
    // is gd); Pp the rows of P at positions [0, cptr[ncol-1]) (the rows the
    // post-smoothing reads before it rewrites them); Rzp is Rz with its
    // columns given as positions. O's float32 CSR-stream blocks per colour:
    // colour c owns [h_oblk32[c], h_oblk32[c+1]) of oblk32.
    bool perm = false;
    CsrMat O, Pp, Rzp;
    std::vector<uint32_t> h_oblk32;
    std::vector<uint64_t> h_onnz;  // entries of O per colour
    DevBuf<uint2> oblk32;
};

// A hierarchy on the device, ready for the V-cycle (smoother data aside).
struct AmgDevHierarchy {
    std::vector<std::unique_ptr<AmgLevelDev>> lv;
    std::vector<double> coarse;  // dense coarsest solution operator (row-major)
};

// the C/F hierarchy built on the device (amg_device.hip); `splits` (may be
// null) receives each coarsened level's split (1 C, 2 F)
AmgDevHierarchy build_cf_hierarchy_device(const CsrMat& A, const fem_pc_params& p, hipStream_t s,
                                          std::vector<std::vector<uint8_t>>* splits);
AmgHierarchy build_sa_hierarchy(const CsrMat& A, const fem_pc_params& p);
int mis2_aggregate(const HCsr& A, double theta, std::vector<int>& agg);
// the robust SA hierarchy (p_omega <= 0) built on the device (amg_device.hip);
// identical to build_sa_hierarchy's
AmgDevHierarchy build_sa_hierarchy_device(const CsrMat& A, const fem_pc_params& p, hipStream_t s);
// a host-built hierarchy moved to the device (R_k = P_k^T formed there)
AmgDevHierarchy upload_hierarchy(AmgHierarchy&& H, hipStream_t s);
std::vector<double> sa_coarse_operator_host(const HCsr& A, bool exact);
// R = P^T on the device; each row of R lists its columns in ascending order
void transpose_device(const CsrMat& P, CsrMat& R, hipStream_t s);
// C = A B on the device, the exact product with each row's columns ascending;
// B's rows must list strictly increasing columns
void spgemm_device(const CsrMat& A, const CsrMat& B, CsrMat& C, hipStream_t s);
// SGS smoother data of level D (colouring, diagonal, Jacobi weight and, with
// at most kMaxGsColours colours, the colour-grouped copy of D.A, or with
// grouped = false only the diagonal in colour order, gd)
void sgs_setup_device(AmgLevelDev& D, hipStream_t s, bool grouped = true);
// the colour-grouped copy of a level coloured by sgs_setup_device
void sgs_grouped_device(AmgLevelDev& D, hipStream_t s);
// out = R without the columns rows[0..count) (entries in stored order)
void drop_columns_device(const CsrMat& R, const uint32_t* rows, uint32_t count, CsrMat& out, hipStream_t s);
// the float32 sweep blocks (gblk32, h_cblk32) of a level set up by sgs_setup_device
void sgs_blocks32(AmgLevelDev& D, hipStream_t s);
// the colour-ordered form (O, Pp, Rzp, oblk32) of a coloured level with R
// without the last colour's columns and the float32 diagonal gd32
void perm_setup_device(AmgLevelDev& D, hipStream_t s);
// aggregation by a distance-2 maximal independent set of the strength graph
// on the device: v holds the entries (norms = false) or one norm per entry
// (norms = true); agg (n entries) receives the aggregates, their count is
// returned
uint32_t mis2_aggregate_device(const uint32_t* rp, const uint32_t* ci, const double* v, bool norms, uint32_t n,
                               uint64_t nnz, double theta, int* agg, hipStream_t s);
// true if every row of A lists strictly increasing columns
bool rows_sorted_device(const CsrMat& A, hipStream_t s);

}  // namespace fem

// amg.hpp — host-side AMG hierarchy description (amg_setup.cpp) consumed by
// the device V-cycle (amg.hip).
#pragma once
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

struct AmgHierarchy {
    std::vector<HCsr> A;                  // level operators, A[0] = input
    std::vector<HCsr> P;                  // prolongators, P[k]: n_k x n_{k+1}
    std::vector<std::vector<double>> d;   // smoother diagonal (C/F: a_ii, SA: 1/a_ii)
    std::vector<double> coarse;           // dense coarse solution operator (row-major)
    std::vector<double> omega;            // smoother weight per level
    // multicolour ordering per level (SGS smoother)
    std::vector<std::vector<uint32_t>> cptr, crows;
    std::vector<std::vector<double>> diag;
    // SGS smoother: damped-Jacobi weight (4/3) / rho(D^{-1} A) (Gershgorin)
    // for the levels whose colouring has more than kMaxGsColours colours
    std::vector<double> jw;
};

AmgHierarchy build_cf_hierarchy(const CsrMat& A, const fem_pc_params& p);
AmgHierarchy build_sa_hierarchy(const CsrMat& A, const fem_pc_params& p);
int mis2_aggregate(const HCsr& A, double theta, std::vector<int>& agg);
// the robust SA hierarchy (p_omega <= 0) built on the device (amg_device.hip);
// identical to build_sa_hierarchy's. Throws DeviceSetupTooLarge when a level
// has more columns than the product's LDS bitmap holds (524288).
struct DeviceSetupTooLarge {};
AmgHierarchy build_sa_hierarchy_device(const CsrMat& A, const fem_pc_params& p, hipStream_t s);
// host parts shared by both setups
std::vector<double> sa_coarse_operator_host(const HCsr& A, bool exact);
void sgs_levels_host(AmgHierarchy& H);

}  // namespace fem

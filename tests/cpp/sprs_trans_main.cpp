// Command-line driver for the SprsTrans tests: reads a matrix as text
// (rows cols nnz, then row col re im per entry, then rows flags), applies one
// transform and writes the result as text (rows cols nnz, row offsets,
// column indices, values as re im).
#include <SparseSolv/SprsTrans.hpp>

#include <cstdio>
#include <iostream>
#include <string>
#include <type_traits>
#include <vector>

using namespace JPMRspace;
using namespace JPMRspace::SparseSolv;

template <typename D> D make(double re, double im);
template <> double make<double>(double re, double) { return re; }
template <> dcomplex make<dcomplex>(double re, double im) { return dcomplex(re, im); }
static double re_of(double v) { return v; }
static double im_of(double) { return 0.0; }
static double re_of(dcomplex v) { return v.real(); }
static double im_of(dcomplex v) { return v.imag(); }

template <typename D> static void put(const SparseMatTMPL<D>& A) {
    const slv_int n = A.getSize();
    std::printf("%d %d %lld\n", (int)n, (int)A.getColSize(), (long long)A.nonZeros());
    for (slv_int i = 0; i <= n; ++i) std::printf("%d ", (int)A.getRowPtr()[i]);
    std::printf("\n");
    for (int64_t k = 0; k < A.nonZeros(); ++k) std::printf("%d ", (int)A.getColPtr()[k]);
    std::printf("\n");
    for (int64_t k = 0; k < A.nonZeros(); ++k)
        std::printf("%.17g %.17g\n", re_of(A.getValuePtr()[k]), im_of(A.getValuePtr()[k]));
}

template <typename D> static int run(const std::string& op, const std::vector<long>& args) {
    int rows, cols;
    long nnz;
    std::cin >> rows >> cols >> nnz;
    std::vector<slv_int> r(nnz), c(nnz);
    std::vector<D> v(nnz);
    for (long k = 0; k < nnz; ++k) {
        double re, im;
        std::cin >> r[k] >> c[k] >> re >> im;
        v[k] = make<D>(re, im);
    }
    std::vector<char> fl(rows);
    for (int i = 0; i < rows; ++i) {
        int f;
        std::cin >> f;
        fl[i] = (char)f;
    }
    bool* flag = reinterpret_cast<bool*>(fl.data());
    // a builder of the declared size, as an application assembles
    SparseBuilderTMPL<D> b(rows);
    for (long k = 0; k < nnz; ++k) b.add(r[k], c[k], v[k]);
    SparseMatTMPL<D> A(b);
    if (op == "trans") put(SprsTrans::trans(A));
    else if (op == "lower") put(SprsTrans::MatLower(A));
    else if (op == "upper") put(SprsTrans::MatUpper(A));
    else if (op == "del") { SprsTrans::DelFlagPosition(A, flag); put(A); }
    else if (op == "diag") { SprsTrans::DiagFlagPosition(A, flag); put(A); }
    else if (op == "sub") put(SprsTrans::makeSubMat(A, args[0], args[1], args[2], args[3]));
    else if (op == "div") {
        SparseMatTMPL<D> k11, k12;
        SprsTrans::MatDiv(k11, k12, A, args[0], args[1]);
        put(k11);
        put(k12);
    } else if (op == "selfdot") { SprsTrans::SelfDot(A, 2.5); put(A); }
    else if (op == "round") {
        if constexpr (std::is_same<D, double>::value) {
            SprsTrans::round(A);
            put(A);
        } else {
            return 2;
        }
    }
    else return 2;
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string kind = argv[1], op = argv[2];
    std::vector<long> args;
    for (int i = 3; i < argc; ++i) args.push_back(std::stol(argv[i]));
    return kind == "real" ? run<double>(op, args) : run<dcomplex>(op, args);
}

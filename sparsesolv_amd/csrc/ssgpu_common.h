// Shared device/host definitions of the gfx950 solver library.
//
// Scalar types: double and `zc` (complex double as an (re, im) pair).  The
// complex operations are written out so their rounding is fixed: product
// (ar*br - ai*bi, ar*bi + ai*br) with no fused multiply-add (the library is
// compiled with -ffp-contract=off), which is what the reference's g++ build
// computes for std::complex<double> multiplication on x86-64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace ssgpu {

struct zc {
    double re, im;
};
struct fc {   // complex single, storage type of FP32 complex factors
    float re, im;
};

__host__ __device__ inline zc make_zc(double r, double i) { return zc{r, i}; }
__host__ __device__ inline zc operator+(zc a, zc b) { return zc{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline zc operator-(zc a, zc b) { return zc{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline zc operator*(zc a, zc b) {
    return zc{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline zc operator*(zc a, double s) { return zc{a.re * s, a.im * s}; }

// Complex division: Smith's algorithm (Smith 1962, "Algorithm 116").
__host__ __device__ inline zc zdiv(zc a, zc b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re;
        const double den = b.re + b.im * r;
        return zc{(a.re + a.im * r) / den, (a.im - a.re * r) / den};
    }
    const double r = b.re / b.im;
    const double den = b.re * r + b.im;
    return zc{(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}

// Generic helpers usable for double and zc.
template <typename T> struct Num;
template <> struct Num<double> {
    __host__ __device__ static double zero() { return 0.0; }
    __host__ __device__ static double mul(double a, double b) { return a * b; }
    __host__ __device__ static double sub(double a, double b) { return a - b; }
    __host__ __device__ static double add(double a, double b) { return a + b; }
    __host__ __device__ static double div(double a, double b) { return a / b; }
    __host__ __device__ static double abs2(double a) { return a * a; }
};
template <> struct Num<zc> {
    __host__ __device__ static zc zero() { return zc{0.0, 0.0}; }
    __host__ __device__ static zc mul(zc a, zc b) { return a * b; }
    __host__ __device__ static zc sub(zc a, zc b) { return a - b; }
    __host__ __device__ static zc add(zc a, zc b) { return a + b; }
    __host__ __device__ static zc div(zc a, zc b) { return zdiv(a, b); }
    __host__ __device__ static double abs2(zc a) { return a.re * a.re + a.im * a.im; }
};

// Status codes mirrored from the C-ABI.
enum Status { OK = 0, NOT_CONVERGED = 1, NO_DIAGONAL = 2, BAD_ARG = 3, OOM = 4, HIP_ERROR = 5,
              NOT_SQUARE = 6, UNSUPPORTED = 7 };

struct Error {
    int code;
    std::string msg;
};

void set_last_error(const std::string& s);

#define SSGPU_HIP(call)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::ssgpu::Error{e_ == hipErrorOutOfMemory ? ::ssgpu::OOM : ::ssgpu::HIP_ERROR, \
                                 std::string(#call) + ": " + hipGetErrorString(e_)};      \
    } while (0)

}  // namespace ssgpu

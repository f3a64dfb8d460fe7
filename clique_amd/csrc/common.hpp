// Scalar types shared by host and device code of libcliqg.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace cliqg {

struct alignas(8) cfloat {
    float x, y;
};
struct alignas(16) cdouble {
    double x, y;
};

// --- traits -----------------------------------------------------------------
template <class T> struct Scalar;
template <> struct Scalar<float> {
    using real = float;
    using wide = double;   // FP64 counterpart (pivots, refinement)
    static constexpr bool complex = false;
};
template <> struct Scalar<double> {
    using real = double;
    using wide = double;
    static constexpr bool complex = false;
};
template <> struct Scalar<cfloat> {
    using real = float;
    using wide = cdouble;
    static constexpr bool complex = true;
};
template <> struct Scalar<cdouble> {
    using real = double;
    using wide = cdouble;
    static constexpr bool complex = true;
};

#define CQ_HD __host__ __device__ __forceinline__

// --- arithmetic ---------------------------------------------------------------
CQ_HD float re(float a) { return a; }
CQ_HD double re(double a) { return a; }
CQ_HD float re(cfloat a) { return a.x; }
CQ_HD double re(cdouble a) { return a.x; }
CQ_HD float im(float) { return 0.f; }
CQ_HD double im(double) { return 0.0; }
CQ_HD float im(cfloat a) { return a.y; }
CQ_HD double im(cdouble a) { return a.y; }

CQ_HD cfloat operator+(cfloat a, cfloat b) { return {a.x + b.x, a.y + b.y}; }
CQ_HD cfloat operator-(cfloat a, cfloat b) { return {a.x - b.x, a.y - b.y}; }
CQ_HD cfloat operator-(cfloat a) { return {-a.x, -a.y}; }
CQ_HD cfloat operator*(cfloat a, cfloat b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
CQ_HD cdouble operator+(cdouble a, cdouble b) { return {a.x + b.x, a.y + b.y}; }
CQ_HD cdouble operator-(cdouble a, cdouble b) { return {a.x - b.x, a.y - b.y}; }
CQ_HD cdouble operator-(cdouble a) { return {-a.x, -a.y}; }
CQ_HD cdouble operator*(cdouble a, cdouble b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
CQ_HD cfloat& operator+=(cfloat& a, cfloat b) { a = a + b; return a; }
CQ_HD cfloat& operator-=(cfloat& a, cfloat b) { a = a - b; return a; }
CQ_HD cdouble& operator+=(cdouble& a, cdouble b) { a = a + b; return a; }
CQ_HD cdouble& operator-=(cdouble& a, cdouble b) { a = a - b; return a; }

// division by a complex number (Smith's algorithm, scaled to avoid overflow)
CQ_HD cdouble operator/(cdouble a, cdouble b) {
    double ar = a.x, ai = a.y, br = b.x, bi = b.y;
    if (fabs(br) >= fabs(bi)) {
        double r = bi / br, den = br + bi * r;
        return {(ar + ai * r) / den, (ai - ar * r) / den};
    } else {
        double r = br / bi, den = bi + br * r;
        return {(ar * r + ai) / den, (ai * r - ar) / den};
    }
}
CQ_HD cfloat operator/(cfloat a, cfloat b) {
    cdouble q = cdouble{a.x, a.y} / cdouble{b.x, b.y};
    return {(float)q.x, (float)q.y};
}

CQ_HD float conj_(float a) { return a; }
CQ_HD double conj_(double a) { return a; }
CQ_HD cfloat conj_(cfloat a) { return {a.x, -a.y}; }
CQ_HD cdouble conj_(cdouble a) { return {a.x, -a.y}; }

template <class T> CQ_HD T cj(T a, bool c) { return c ? conj_(a) : a; }

CQ_HD double sq_mod(double a) { return a * a; }
CQ_HD double sq_mod(cdouble a) { return a.x * a.x + a.y * a.y; }
CQ_HD double abs1(double a) { return fabs(a); }
CQ_HD double abs1(float a) { return fabs((double)a); }
CQ_HD double abs1(cdouble a) { return hypot(a.x, a.y); }
CQ_HD double abs1(cfloat a) { return hypot((double)a.x, (double)a.y); }

CQ_HD bool is_zero_or_nonfinite(double a) { return !(a != 0.0) || !isfinite(a); }
CQ_HD bool is_zero_or_nonfinite(cdouble a) {
    return (a.x == 0.0 && a.y == 0.0) || !isfinite(a.x) || !isfinite(a.y);
}

// conversions between the FP64 field type and the working type
template <class To, class From> CQ_HD To cvt(From a);
template <> CQ_HD float cvt<float, double>(double a) { return (float)a; }
template <> CQ_HD double cvt<double, double>(double a) { return a; }
template <> CQ_HD double cvt<double, float>(float a) { return (double)a; }
template <> CQ_HD float cvt<float, float>(float a) { return a; }
template <> CQ_HD cfloat cvt<cfloat, cdouble>(cdouble a) { return {(float)a.x, (float)a.y}; }
template <> CQ_HD cdouble cvt<cdouble, cdouble>(cdouble a) { return a; }
template <> CQ_HD cdouble cvt<cdouble, cfloat>(cfloat a) { return {(double)a.x, (double)a.y}; }
template <> CQ_HD cfloat cvt<cfloat, cfloat>(cfloat a) { return a; }

CQ_HD double scale(double a, double s) { return a * s; }
CQ_HD cdouble scale(cdouble a, double s) { return {a.x * s, a.y * s}; }

template <class T> CQ_HD T zero() { return T{}; }
template <class T> CQ_HD T one();
template <> CQ_HD float one<float>() { return 1.f; }
template <> CQ_HD double one<double>() { return 1.0; }
template <> CQ_HD cfloat one<cfloat>() { return {1.f, 0.f}; }
template <> CQ_HD cdouble one<cdouble>() { return {1.0, 0.0}; }

// --- errors -------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// a number in scientific notation for messages
inline std::string sci(double v) {
    char b[32];
    std::snprintf(b, sizeof(b), "%.3e", v);
    return b;
}

#define CQ_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::cliqg::Error(e_ == hipErrorOutOfMemory ? 4 : 5,                     \
                                 std::string("HIP error in " #call ": ") +              \
                                     hipGetErrorString(e_));                            \
    } while (0)

}  // namespace cliqg

// Device helpers shared by the kernel translation units.
#pragma once

#include "kernels.hpp"

namespace pf {

// 2x2 block [[a, b], [c, d]] (row-major), stored as 4 consecutive T.
template <typename T>
struct B4 {
    T a, b, c, d;
};

__device__ __forceinline__ B4<double> load_b4(const double *p, int s) {
    const double2 *q = reinterpret_cast<const double2 *>(p) + 2 * (size_t)s;
    double2 u = q[0], v = q[1];
    return B4<double>{u.x, u.y, v.x, v.y};
}
__device__ __forceinline__ B4<float> load_b4(const float *p, int s) {
    float4 u = reinterpret_cast<const float4 *>(p)[s];
    return B4<float>{u.x, u.y, u.z, u.w};
}
__device__ __forceinline__ void store_b4(double *p, int s, B4<double> v) {
    double2 *q = reinterpret_cast<double2 *>(p) + 2 * (size_t)s;
    q[0] = make_double2(v.a, v.b);
    q[1] = make_double2(v.c, v.d);
}
__device__ __forceinline__ void store_b4(float *p, int s, B4<float> v) {
    reinterpret_cast<float4 *>(p)[s] = make_float4(v.a, v.b, v.c, v.d);
}

// c -= a * b
template <typename T>
__device__ __forceinline__ void mms(B4<T> &c, const B4<T> &a, const B4<T> &b) {
    c.a -= a.a * b.a + a.b * b.c;
    c.b -= a.a * b.b + a.b * b.d;
    c.c -= a.c * b.a + a.d * b.c;
    c.d -= a.c * b.b + a.d * b.d;
}

template <typename T>
__device__ __forceinline__ B4<T> mul(const B4<T> &a, const B4<T> &b) {
    return B4<T>{a.a * b.a + a.b * b.c, a.a * b.b + a.b * b.d, a.c * b.a + a.d * b.c,
                 a.c * b.b + a.d * b.d};
}

// Smith's complex division (CACM Algorithm 116, 1962): (ar + j ai) / (br + j bi)
__device__ __forceinline__ void smith_div(double ar, double ai, double br, double bi, double &cr,
                                          double &ci) {
#pragma clang fp contract(off)
    if (fabs(br) >= fabs(bi)) {
        double r = bi / br, d = br + bi * r;
        cr = (ar + ai * r) / d;
        ci = (ai - ar * r) / d;
    } else {
        double r = br / bi, d = bi + br * r;
        cr = (ar * r + ai) / d;
        ci = (ai * r - ar) / d;
    }
}

__device__ __forceinline__ unsigned long long double_as_u64(double v) {
    return (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ double u64_as_double(unsigned long long v) {
    return __longlong_as_double((long long)v);
}

// max that keeps a NaN (a non-finite mismatch must never look converged)
__device__ __forceinline__ double dmax_nan(double a, double b) { return (a != a || a > b) ? a : b; }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dmax_nan(v, __shfl_xor(v, off));
    return v;
}

}  // namespace pf

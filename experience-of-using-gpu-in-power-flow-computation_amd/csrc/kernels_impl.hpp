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

// c += a * b
template <typename T>
__device__ __forceinline__ void mma2(B4<T> &c, const B4<T> &a, const B4<T> &b) {
    c.a += a.a * b.a + a.b * b.c;
    c.b += a.a * b.b + a.b * b.d;
    c.c += a.c * b.a + a.d * b.c;
    c.d += a.c * b.b + a.d * b.d;
}

template <typename T, int G>
__device__ __forceinline__ B4<T> group_sum(B4<T> v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        v.a += __shfl_xor(v.a, off, G);
        v.b += __shfl_xor(v.b, off, G);
        v.c += __shfl_xor(v.c, off, G);
        v.d += __shfl_xor(v.d, off, G);
    }
    return v;
}

// One elimination-tree level factored by a whole workgroup: groups of G
// lanes take the level's destination blocks (every node's pivot and column
// entries), the lanes of a group split that destination's update pairs and
// reduce them in a fixed order; then the pivots are inverted and the column
// scaled.  Three barriers per level whatever its width, so the narrow top of
// the tree (one node per level) keeps every lane of the workgroup busy.
// The caller synchronises after the call.
template <typename T, int G>
__device__ void coop_factor_level(const DevSym &Y, T *blk, int l, int *bad) {
    const int L0 = Y.n, U0 = Y.n + Y.nnzL;
    const int e0 = Y.ld_ptr[l], ne = Y.ld_ptr[l + 1] - e0;
    const int grp = (int)threadIdx.x / G, lane = (int)threadIdx.x % G, ng = (int)blockDim.x / G;
    for (int e = grp; e < ne; e += ng) {
        const int j = Y.ld_node[e0 + e], d = Y.ld_d[e0 + e];
        const int t = d - Y.fd_ptr[j];
        B4<T> a{0, 0, 0, 0}, u{0, 0, 0, 0};
        if (t == 0) {
            for (int q = Y.fc_ptr[d] + lane; q < Y.fc_ptr[d + 1]; q += G)
                mma2(a, load_b4(blk, L0 + Y.fc_a[q]), load_b4(blk, U0 + Y.fc_b[q]));
        } else {
            for (int q = Y.fc_ptr[d] + lane; q < Y.fc_ptr[d + 1]; q += G) {
                const int pik = Y.fc_a[q], pjk = Y.fc_b[q];
                mma2(a, load_b4(blk, L0 + pik), load_b4(blk, U0 + pjk));
                mma2(u, load_b4(blk, L0 + pjk), load_b4(blk, U0 + pik));
            }
        }
        a = group_sum<T, G>(a);
        u = group_sum<T, G>(u);
        if (lane == 0) {
            if (t == 0) {
                B4<T> D = load_b4(blk, j);
                store_b4(blk, j, B4<T>{D.a - a.a, D.b - a.b, D.c - a.c, D.d - a.d});
            } else {
                const int p = Y.colptr[j] + t - 1;
                B4<T> Lb = load_b4(blk, L0 + p), Ub = load_b4(blk, U0 + p);
                store_b4(blk, L0 + p, B4<T>{Lb.a - a.a, Lb.b - a.b, Lb.c - a.c, Lb.d - a.d});
                store_b4(blk, U0 + p, B4<T>{Ub.a - u.a, Ub.b - u.b, Ub.c - u.c, Ub.d - u.d});
            }
        }
    }
    __syncthreads();
    const int n0 = Y.lvl_ptr[l], nn = Y.lvl_ptr[l + 1] - n0;
    for (int i = threadIdx.x; i < nn; i += blockDim.x) {
        const int j = Y.lvl_nodes[n0 + i];
        B4<T> D = load_b4(blk, j);
        T det = D.a * D.d - D.b * D.c;
        if (!(det != (T)0) || !isfinite((double)det)) {
            atomicOr(bad, 1);
            det = (T)1;
        }
        store_b4(blk, j, B4<T>{D.d / det, -D.b / det, -D.c / det, D.a / det});
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        const int j = Y.ld_node[e0 + e], d = Y.ld_d[e0 + e];
        const int t = d - Y.fd_ptr[j];
        if (t == 0) continue;
        const int p = Y.colptr[j] + t - 1;
        store_b4(blk, L0 + p, mul(load_b4(blk, L0 + p), load_b4(blk, j)));
    }
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

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dmax_nan(v, __shfl_xor(v, off));
    return v;
}

}  // namespace pf

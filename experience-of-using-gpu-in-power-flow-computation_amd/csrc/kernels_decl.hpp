// Declarations of the kernels defined in kernels.hip (launched from solver.hip).
#pragma once

#include "kernels.hpp"

namespace pf {

__global__ void k_ybus(DevNet N, int nnz);
__global__ void k_init_E(DevNet N, DevState S);
template <int LPR> __global__ void k_pq(DevNet N, DevState S, int gate);
__global__ void k_limits(DevNet N, DevState S, int mode, int lim);
__global__ void k_mismatch(DevNet N, DevState S, int gate);
__global__ void k_control(DevState S, int lim);
template <typename T> __global__ void k_jfill(DevNet N, DevState S, DevSym Y, T *blk, int gate);
template <typename T> __global__ void k_factor_level(DevSym Y, T *blk, Ctl *ctl, int l0, int count, int gate);
template <typename T> __global__ void k_factor_tail(DevSym Y, T *blk, Ctl *ctl, int lv0, int lv1, int gate);
template <typename T> __global__ void k_tail_factor_ext(DevSym Y, DevFlat F, T *blk, Ctl *ctl, int ne, int gate);
template <typename T> __global__ void k_factor_tail_flat(DevSym Y, DevFlat F, T *blk, Ctl *ctl, int nsteps, int gate, int cap);
template <typename T> __global__ void k_factor_parts_flat(DevSym Y, DevFlat F, T *blk, Ctl *ctl, int gate);
template <typename T> __global__ void k_factor_parts(DevSym Y, T *blk, Ctl *ctl, int gate);
template <typename T> __global__ void k_chain_schur(DevSym Y, T *blk, Ctl *ctl, int gate);
template <typename T> __global__ void k_chain_invert(DevSym Y, T *blk, Ctl *ctl, int gate);
__global__ void k_flows(DevNet N, DevState S, const FlowBranch *br, int nL, double *out);
template <typename T> __global__ void k_fwd_parts(DevSym Y, const T *blk, double *x, Ctl *ctl, int gate);
template <typename T> __global__ void k_fwd_parts_lds(DevSym Y, DevParts P, const T *blk, double *x, Ctl *ctl, int gate);
template <typename T> __global__ void k_bwd_parts_lds(DevSym Y, DevParts P, const T *blk, double *x, Ctl *ctl, int gate);
template <typename T> __global__ void k_bwd_parts(DevSym Y, const T *blk, double *x, Ctl *ctl, int gate);
template <typename T> __global__ void k_fwd_level(DevSym Y, const T *blk, double *x, Ctl *ctl, int l0, int count, int gate);
template <typename T> __global__ void k_bwd_level(DevSym Y, const T *blk, double *x, Ctl *ctl, int l0, int count, int gate);
template <typename T> __global__ void k_fwd_tail(DevSym Y, const T *blk, double *x, Ctl *ctl, int lv0, int lv1, int gate);
template <typename T> __global__ void k_bwd_tail(DevSym Y, const T *blk, double *x, Ctl *ctl, int lv0, int lv1, int gate);
template <typename T> __global__ void k_tail_fwd_ext(DevSym Y, DevTail D, const T *blk, double *x, Ctl *ctl, int gate);
template <typename T> __global__ void k_tail_bwd_ext(DevSym Y, DevTail D, const T *blk, double *x, Ctl *ctl, int gate);
template <typename T, bool STAGED>
__global__ void k_fwd_tail_lds(DevSym Y, DevTail D, const T *blk, double *x, Ctl *ctl, int lv0, int lv1, int gate);
template <typename T, bool STAGED>
__global__ void k_bwd_tail_lds(DevSym Y, DevTail D, const T *blk, double *x, Ctl *ctl, int lv0, int lv1, int gate);
__global__ void k_copy(double *dst, const double *src, int n, Ctl *ctl, int gate);
__global__ void k_axpy(double *x, const double *d, int n, Ctl *ctl, int gate);
__global__ void k_fill(double *x, size_t n, double v);
template <int LPR> __global__ void k_residual(DevNet N, DevState S, int gate);
__global__ void k_update(DevNet N, DevState S, int gate);
__global__ void k_dense_scatter(DevSym Y, const double *blk, double *A, int ld, Ctl *ctl, int gate);
__global__ void k_dense(double *A, double *b, int n, Ctl *ctl, int gate);

}  // namespace pf

// Device-side wave64 / block primitives shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>

namespace apc {

__device__ inline int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Exclusive scan over a 256-thread block; *total receives the block sum.
__device__ inline int block_excl_scan256(int v, int *ws, int *total) {
  int inc = wave_incl_scan(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  int off = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) off += (q < w) ? ws[q] : 0;
  *total = ws[0] + ws[1] + ws[2] + ws[3];
  __syncthreads();
  return off + inc - v;
}

// set bits of the wave mask m in the lanes below this one (v_mbcnt_lo /
// v_mbcnt_hi: two vector instructions where popcount(m & lanes-below mask)
// takes four)
__device__ inline int below(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Whole-wave reductions by DPP row operations (register to register, no LDS
// crossbar): an inclusive scan in each 16-lane row by shifts of 1, 2, 4 and
// 8 lanes, then row 0's total into row 1 and row 2's into row 3
// (row_bcast:15), then rows 0-1's total into rows 2-3 (row_bcast:31); lane 63
// ends with the reduction of all 64 lanes, read back as a wave-uniform value.
// Lanes with no source in a step take the identity (bound_ctrl off, disabled
// rows keep the `old` operand).  Every lane must be active.
template <int CTRL, int ROWS>
__device__ inline int dpp_or(int id, int v) {
  return __builtin_amdgcn_update_dpp(id, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ inline double dpp_or(double id, double v) {
  const int lo = dpp_or<CTRL, ROWS>(__double2loint(id), __double2loint(v));
  const int hi = dpp_or<CTRL, ROWS>(__double2hiint(id), __double2hiint(v));
  return __hiloint2double(hi, lo);
}
template <class T, class OP>
__device__ inline T wave_reduce_dpp(T v, T id, OP op) {
  v = op(v, dpp_or<0x111, 0xf>(id, v));  // row_shr:1
  v = op(v, dpp_or<0x112, 0xf>(id, v));  // row_shr:2
  v = op(v, dpp_or<0x114, 0xf>(id, v));  // row_shr:4
  v = op(v, dpp_or<0x118, 0xf>(id, v));  // row_shr:8
  v = op(v, dpp_or<0x142, 0xa>(id, v));  // row_bcast:15 into rows 1 and 3
  v = op(v, dpp_or<0x143, 0xc>(id, v));  // row_bcast:31 into rows 2 and 3
  return v;
}
__device__ inline int wave_min_int_dpp(int v) {
  v = wave_reduce_dpp(v, INT_MAX, [](int a, int b) { return min(a, b); });
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ inline double wave_sum_dpp(double v) {
  v = wave_reduce_dpp(v, 0.0, [](double a, double b) { return a + b; });
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                          __builtin_amdgcn_readlane(__double2loint(v), 63));
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Sum over a 256-thread block in a fixed order (deterministic); all threads
// get the result.
__device__ inline double block_sum256(double v, double *ws) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) ws[w] = v;
  __syncthreads();
  double r = (ws[0] + ws[1]) + (ws[2] + ws[3]);
  __syncthreads();
  return r;
}

__device__ inline double block_max256(double v, double *ws) {
  v = wave_max(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) ws[w] = v;
  __syncthreads();
  double r = fmax(fmax(ws[0], ws[1]), fmax(ws[2], ws[3]));
  __syncthreads();
  return r;
}

}  // namespace apc

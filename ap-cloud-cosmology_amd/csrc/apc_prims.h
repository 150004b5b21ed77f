// Device-side wave64 / block primitives shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>

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

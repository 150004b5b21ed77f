// GFD least-squares helpers shared by the stencil / gradient kernels and
// their adjoints (solver.cpp:77-100, 295-362).
#pragma once
#include "apc_internal.h"
#include "apc_prims.h"

namespace apc {

__device__ inline double pwrap_s(double x, double len, int bc) {
  if (bc == 1 && isfinite(x)) {  // a non-finite value would never leave the loop
    while (x >= (0.5 * len)) x -= len;
    while (x < -(0.5 * len)) x += len;
  }
  return x;
}

constexpr int RI(int k, int m) { return 9 * k - (k * (k - 1)) / 2 + (m - k); }

struct Geo {
  double len[3];
  int bc;
};

// Row basis (solver.cpp:87-100): [p1,p2,p3,p1^2/2,p2^2/2,p3^2/2,p1p2,p1p3,p2p3]
__device__ inline void basis_row(double p1, double p2, double p3, double *r) {
  r[0] = p1; r[1] = p2; r[2] = p3;
  r[3] = p1 * p1 / 2.0; r[4] = p2 * p2 / 2.0; r[5] = p3 * p3 / 2.0;
  r[6] = p1 * p2; r[7] = p1 * p3; r[8] = p2 * p3;
}

// Add the outer product r r^T to the Gram matrix R (upper, packed as RI).
__device__ inline void gram_add(double *R, const double *r) {
#pragma unroll
  for (int k = 0; k < 9; ++k)
#pragma unroll
    for (int q = k; q < 9; ++q) R[RI(k, q)] += r[k] * r[q];
}

// Cholesky factorisation in place: the Gram matrix B^T B becomes the upper
// triangular R with R^T R = B^T B.  Returns true if a pivot is not positive
// (B has no full column rank).
__device__ inline bool chol_inplace(double *R) {
  bool deficient = false;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    double d = R[RI(k, k)];
#pragma unroll
    for (int q = 0; q < k; ++q) d -= R[RI(q, k)] * R[RI(q, k)];
    deficient |= !(d > 0.0);
    const double rk = sqrt(fmax(d, 0.0));
    R[RI(k, k)] = rk;
    const double inv = 1.0 / rk;
#pragma unroll
    for (int m = k + 1; m < 9; ++m) {
      double e = R[RI(k, m)];
#pragma unroll
      for (int q = 0; q < k; ++q) e -= R[RI(q, k)] * R[RI(q, m)];
      R[RI(k, m)] = e * inv;
    }
  }
  return deficient;
}

// z = R^-1 R^-T u (the semi-normal solve (B^T B) z = u)
__device__ inline void seminormal_solve(const double *R, const double *u, double *z) {
  double y[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    double s = u[k];
#pragma unroll
    for (int q = 0; q < k; ++q) s -= R[RI(q, k)] * y[q];
    y[k] = s / R[RI(k, k)];
  }
#pragma unroll
  for (int k = 8; k >= 0; --k) {
    double s = y[k];
#pragma unroll
    for (int q = k + 1; q < 9; ++q) s -= R[RI(k, q)] * z[q];
    z[k] = s / R[RI(k, k)];
  }
}

// Gaussian weights of solver.cpp:77-85 (maxdis = 0.7 * max distance).
__device__ inline double row_maxdis(const double *__restrict__ dist, int M, int i, int lm) {
  double maxdis = 0;
  for (int j = 0; j < lm; ++j) {
    double d = dist[(size_t)j * M + i];
    if (d > maxdis) maxdis = d;
  }
  return maxdis * 0.7;
}
__device__ inline double weight_of(double d, double maxdis) {
  return (exp(-d * d / maxdis / maxdis) - exp(-4.0)) / (1 - exp(-4.0));
}

inline Geo geo_of(apc_ctx *c) {
  Geo G;
  const double *b = c->prm.bbox;
  G.len[0] = b[1] - b[0];
  G.len[1] = b[3] - b[2];
  G.len[2] = b[5] - b[4];
  G.bc = c->prm.bc;
  return G;
}

}  // namespace apc

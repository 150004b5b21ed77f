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

// Rotate row r (with optional RHS rb) into R (and Qtb).
template <bool RHS>
__device__ inline void givens_add(double *R, double *Qtb, double *r, double rb) {
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    double a = R[RI(k, k)], b = r[k];
    if (b != 0.0) {
      double rr = sqrt(a * a + b * b);
      double cs = a / rr, sn = b / rr;
      R[RI(k, k)] = rr;
#pragma unroll
      for (int m = k + 1; m < 9; ++m) {
        double u = R[RI(k, m)], v = r[m];
        R[RI(k, m)] = cs * u + sn * v;
        r[m] = cs * v - sn * u;
      }
      if (RHS) {
        double u = Qtb[k];
        Qtb[k] = cs * u + sn * rb;
        rb = cs * rb - sn * u;
      }
    }
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

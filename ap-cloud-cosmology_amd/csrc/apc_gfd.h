// GFD least-squares helpers shared by the stencil / gradient kernels and
// their adjoints (solver.cpp:77-100, 295-362).
#pragma once
#include "apc_internal.h"
#include "apc_prims.h"

namespace apc {

// Periodic minimum image of a coordinate difference: subtract len while x >=
// len/2, then add it while x < -len/2.  A difference of two positions inside
// the box lies in (-len, len), where one conditional step of each kind is the
// whole loop; the loop runs only for anything farther out.
__device__ inline double pwrap_s(double x, double len, int bc) {
  if (bc == 1) {
    const double hl = 0.5 * len;
    x = (x >= hl) ? x - len : x;
    x = (x < -hl) ? x + len : x;
    if (!(x >= -hl && x < hl) && isfinite(x)) {  // a non-finite value would never leave the loop
      while (x >= hl) x -= len;
      while (x < -hl) x += len;
    }
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

// The Gram matrix of a fit from moments.  Every basis entry is a monomial of
// degree 1 or 2 in the scaled offset p, so every Gram entry sum over slots of
// W b_k b_l (W = w^2) is a weighted moment sum W p1^a p2^b p3^c of degree 2
// to 4, times 1, 1/2 or 1/4: 31 distinct sums instead of 45 (fewer registers
// held across the slot loop, fewer products per slot).
// Moment order: degree 2 (200 110 101 020 011 002), degree 3 (300 210 201 120
// 111 102 030 021 012 003), degree 4 (400 310 301 220 211 202 130 121 112
// 103 040 031 022 013 004).
constexpr int kMoments = 31;

__device__ inline void moments_add(double *m, double W, double p1, double p2, double p3) {
  const double q11 = p1 * p1, q12 = p1 * p2, q13 = p1 * p3;
  const double q22 = p2 * p2, q23 = p2 * p3, q33 = p3 * p3;
  const double w11 = W * q11, w12 = W * q12, w13 = W * q13;
  const double w22 = W * q22, w23 = W * q23, w33 = W * q33;
  m[0] += w11; m[1] += w12; m[2] += w13; m[3] += w22; m[4] += w23; m[5] += w33;
  m[6] += w11 * p1; m[7] += w11 * p2; m[8] += w11 * p3; m[9] += w22 * p1;
  m[10] += w12 * p3; m[11] += w33 * p1; m[12] += w22 * p2; m[13] += w22 * p3;
  m[14] += w33 * p2; m[15] += w33 * p3;
  m[16] += w11 * q11; m[17] += w11 * q12; m[18] += w11 * q13; m[19] += w11 * q22;
  m[20] += w11 * q23; m[21] += w11 * q33; m[22] += w12 * q22; m[23] += w12 * q23;
  m[24] += w12 * q33; m[25] += w13 * q33; m[26] += w22 * q22; m[27] += w22 * q23;
  m[28] += w22 * q33; m[29] += w23 * q33; m[30] += w33 * q33;
}

// R (upper, packed as RI) = the Gram matrix of the basis
// [p1, p2, p3, p1^2/2, p2^2/2, p3^2/2, p1p2, p1p3, p2p3] from its moments
__device__ inline void gram_from_moments(const double *m, double *R) {
  const double h = 0.5, q = 0.25;
  R[RI(0, 0)] = m[0];      R[RI(0, 1)] = m[1];      R[RI(0, 2)] = m[2];
  R[RI(0, 3)] = h * m[6];  R[RI(0, 4)] = h * m[9];  R[RI(0, 5)] = h * m[11];
  R[RI(0, 6)] = m[7];      R[RI(0, 7)] = m[8];      R[RI(0, 8)] = m[10];
  R[RI(1, 1)] = m[3];      R[RI(1, 2)] = m[4];      R[RI(1, 3)] = h * m[7];
  R[RI(1, 4)] = h * m[12]; R[RI(1, 5)] = h * m[14]; R[RI(1, 6)] = m[9];
  R[RI(1, 7)] = m[10];     R[RI(1, 8)] = m[13];
  R[RI(2, 2)] = m[5];      R[RI(2, 3)] = h * m[8];  R[RI(2, 4)] = h * m[13];
  R[RI(2, 5)] = h * m[15]; R[RI(2, 6)] = m[10];     R[RI(2, 7)] = m[11];
  R[RI(2, 8)] = m[14];
  R[RI(3, 3)] = q * m[16]; R[RI(3, 4)] = q * m[19]; R[RI(3, 5)] = q * m[21];
  R[RI(3, 6)] = h * m[17]; R[RI(3, 7)] = h * m[18]; R[RI(3, 8)] = h * m[20];
  R[RI(4, 4)] = q * m[26]; R[RI(4, 5)] = q * m[28]; R[RI(4, 6)] = h * m[22];
  R[RI(4, 7)] = h * m[23]; R[RI(4, 8)] = h * m[27];
  R[RI(5, 5)] = q * m[30]; R[RI(5, 6)] = h * m[24]; R[RI(5, 7)] = h * m[25];
  R[RI(5, 8)] = h * m[29];
  R[RI(6, 6)] = m[19];     R[RI(6, 7)] = m[20];     R[RI(6, 8)] = m[23];
  R[RI(7, 7)] = m[21];     R[RI(7, 8)] = m[24];
  R[RI(8, 8)] = m[28];
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

// ||R||_F ||R^-1||_F for the packed upper triangular R (R^T R = B^T B): an
// upper bound on cond_2(R) = cond_2(B) that exceeds it by at most a factor
// of 9.  Column c of R^-1 is R^-1 e_c by back substitution (zero below c).
__device__ inline double cond_frob(const double *R) {
  double nr = 0.0;
#pragma unroll
  for (int q = 0; q < 45; ++q) nr += R[q] * R[q];
  double id[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) id[k] = 1.0 / R[RI(k, k)];
  double ni = 0.0;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    double v[9];
#pragma unroll
    for (int k = 8; k >= 0; --k) {
      if (k > c) {
        v[k] = 0.0;
        continue;
      }
      double t = (k == c) ? 1.0 : 0.0;
#pragma unroll
      for (int q = k + 1; q <= c; ++q) t -= R[RI(k, q)] * v[q];
      v[k] = t * id[k];
      ni += v[k] * v[k];
    }
  }
  return sqrt(nr * ni);
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

// Per-row constants of a fit: the row's position, 1/h and 1/maxdis^2.  The
// reference divides by h and twice by maxdis for every neighbour; here each is
// one product with a reciprocal taken once per row (a rounding-level
// difference; a double division costs about ten instructions).
struct FitRow {
  double xi, yi, zi, ih, inv_md2;
};

__device__ inline FitRow fit_row(const double *__restrict__ cx, const double *__restrict__ cy,
                                 const double *__restrict__ cz, const double *__restrict__ ch,
                                 const double *__restrict__ dist, int M, int i, int lm) {
  FitRow f;
  f.xi = cx[i];
  f.yi = cy[i];
  f.zi = cz[i];
  f.ih = 1.0 / ch[i];
  const double md = row_maxdis(dist, M, i, lm);
  f.inv_md2 = 1.0 / (md * md);
  return f;
}

// w = (exp(-d^2/maxdis^2) - e^-4) / (1 - e^-4)
__device__ inline double weight_of(double d, double inv_md2) {
  constexpr double E4 = 0x1.2c155b8213cf4p-6;  // exp(-4) rounded to double
  constexpr double INV_NORM = 1.0 / (1.0 - E4);
  return (exp(-(d * d) * inv_md2) - E4) * INV_NORM;
}

// Neighbour slot of a fit: the weight and the wrapped offset (x_m - x_i) / h.
__device__ inline double fit_offset(const Geo &G, const FitRow &f, double d, double x, double y,
                                    double z, double *p) {
  p[0] = pwrap_s(x - f.xi, G.len[0], G.bc) * f.ih;
  p[1] = pwrap_s(y - f.yi, G.len[1], G.bc) * f.ih;
  p[2] = pwrap_s(z - f.zi, G.len[2], G.bc) * f.ih;
  return weight_of(d, f.inv_md2);
}

// Neighbour slot of a fit: the weight and the unweighted basis row of the
// wrapped offset (x_m - x_i) / h.
__device__ inline double fit_slot(const Geo &G, const FitRow &f, double d, double x, double y,
                                  double z, double *r) {
  basis_row(pwrap_s(x - f.xi, G.len[0], G.bc) * f.ih, pwrap_s(y - f.yi, G.len[1], G.bc) * f.ih,
            pwrap_s(z - f.zi, G.len[2], G.bc) * f.ih, r);
  return weight_of(d, f.inv_md2);
}

// Visit the lm neighbour slots of row i in order, body(j, m, d, x, y, z, e)
// with m = nbr_ij, d = dist_ij, (x, y, z) the position of m and e =
// extra(m).  The loads run two slots ahead: slot j+2's index and distance and
// slot j+1's gathers are in flight while slot j is computed, so a thread does
// not wait for its chain of dependent loads (index, then gather) on every
// slot.  Loads past the last slot repeat the last slot.
template <class Extra, class Body>
__device__ inline void for_slots(int M, int lm, int i, const int *__restrict__ nbr,
                                 const double *__restrict__ dist, const double *__restrict__ cx,
                                 const double *__restrict__ cy, const double *__restrict__ cz,
                                 Extra &&extra, Body &&body) {
  const int last = lm - 1;
  auto at = [&](int j) { return (size_t)(j < last ? j : last) * M + i; };
  int mC = nbr[at(0)];
  double dC = dist[at(0)];
  int mB = nbr[at(1)];
  double dB = dist[at(1)];
  double xC = cx[mC], yC = cy[mC], zC = cz[mC], eC = extra(mC);
  for (int j = 0; j < lm; ++j) {
    const int mA = nbr[at(j + 2)];
    const double dA = dist[at(j + 2)];
    const double xB = cx[mB], yB = cy[mB], zB = cz[mB], eB = extra(mB);
    body(j, mC, dC, xC, yC, zC, eC);
    mC = mB; dC = dB; xC = xB; yC = yB; zC = zB; eC = eB;
    mB = mA; dB = dA;
  }
}

struct NoExtra {
  __device__ double operator()(int) const { return 0.0; }
};

// Packed boxes: the box of every row, the first row of every box and the
// boxes' mass factors; all null for a single-box context.  The periodic
// potential system pins the first row of every box (row 0 alone without
// boxes), and the density right-hand side carries the box's mass factor.
struct Boxes {
  const int *rowbox, *rstart;
  const double *mass;
};

__device__ inline bool pin_row(const Boxes &B, int pin, int i) {
  return pin && (B.rowbox ? B.rstart[B.rowbox[i]] == i : i == 0);
}

__device__ inline double box_mass(const Boxes &B, int i) {
  return B.rowbox ? B.mass[B.rowbox[i]] : 1.0;
}

inline Boxes boxes_of(apc_ctx *c) {
  if (c->nbox > 1) return Boxes{c->rowbox.p, c->rstart.p, c->boxmass.p};
  return Boxes{nullptr, nullptr, nullptr};
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

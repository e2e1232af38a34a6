// Thin-plate-spline subdomain smoothing in FP64 (gfx950).
//
// Per subdomain with n centers c_j and values v (reference:
// src/fluidimage/calcul/interpolate/thin_plate_spline.py:168-216 and
// thin_plate_spline_subdom.py:128-159,224-280):
//   [[K + sI, P], [P^T, 0]] [w; a] = [v; 0],  K_ij = r^2 ln(r^2) / 2,
//   P = [1, y, x].
// Solved here by the null-space method:
//   P = Q [R; 0] (3 Householder reflectors, compact WY form Q = I - V T V^T),
//   B' = Q^T (K + sI) Q = B - V Z^T - Z V^T  (one pass over B: Y = B V),
//   C = B'[3:, 3:] is symmetric positive definite (K is conditionally
//   positive definite of order 2) -> blocked Cholesky C = L L^T,
//   z = C^-1 (Q^T v)[3:],  a = R^-1 ((Q^T v)[:3] - B'[:3, 3:] z),
//   w = Q [0; z],  smoothed = v - s w  (= [K P] [w; a]).
// One factorisation serves both components and every outlier iteration.
// One workgroup per subdomain; matrices live in a caller workspace.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <string>

#include <algorithm>
#include <atomic>
#include <mutex>

#include "../../include/piv_hip.h"
#include "last_error.hpp"
#include "tuning.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kNB = 32;       // Cholesky / triangular-solve block size
constexpr int kMaxN = 2048;   // centers per subdomain
constexpr int kMaxLatticeA = 4095;            // largest lattice table: 4096 x 4096

// 1 / x and 1 / sqrt(x) for positive normal x: the hardware estimates
// refined by two Newton-Raphson steps (y <- y (2 - x y), y <- y (3 - x y^2) / 2)
__device__ __forceinline__ double rcp_d(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  y = fma(y, fma(-x, y, 1.0), y);
  return y;
}

__device__ __forceinline__ double rsqrt_d(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}

// ln x for positive normal x, to about 2 ulp: x = 2^e m with m in
// [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m - 1) / (m + 1), |s| < 0.1716,
// by the odd series s + s^3 / 3 + ... + s^19 / 19 (the first omitted term is
// below 2.3e-17 relative); e ln 2 added in two parts (the high part has 42
// significant bits, so e ln2_hi is exact for every exponent)
__device__ __forceinline__ double ln_pos(double x) {
  constexpr double kLn2Hi = 0x1.62e42fefa3800p-1, kLn2Lo = 0x1.ef35793c76730p-45;
  int e = __builtin_amdgcn_frexp_exp(x);  // x = mant 2^e, mant in [1/2, 1)
  double m = __builtin_amdgcn_frexp_mant(x);
  if (m < 0.70710678118654752440) {
    m = 2.0 * m;
    e -= 1;
  }
  const double s = (m - 1.0) * rcp_d(m + 1.0);
  const double z = s * s;
  double p = 1.0 / 19.0;
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  const double t = 2.0 * s;
  const double de = double(e);
  return fma(de, kLn2Hi, fma(de, kLn2Lo, fma(t * z, p, t)));
}

// r^2 ln(r^2) / 2, zero at r = 0
__device__ __forceinline__ double tps_phi(double r2) {
  return r2 != 0.0 ? r2 * ln_pos(r2) * 0.5 : 0.0;
}

// r^2 of a difference, one expression for every caller (the lattice table
// and the direct evaluation round alike)
__device__ __forceinline__ double tps_r2(double dx, double dy) { return fma(dy, dy, dx * dx); }

// Matrix region of a subdomain of n centers (m = n - 3): E = B'[3:, :3]
// (m x 3, row-major), then C = B'[3:, 3:] with rows of tps_mat_ld(m)
// doubles (m rounded up to 16) from a 128-byte boundary, the region itself
// starting on one: a 32-column block of a row of C is two cache lines.
__host__ __device__ __forceinline__ int64_t tps_mat_ld(int m) { return (int64_t(m) + 15) & ~int64_t(15); }
__host__ __device__ __forceinline__ int64_t tps_mat_e_size(int m) {
  return (3 * int64_t(m) + 15) & ~int64_t(15);
}
__host__ __device__ __forceinline__ int64_t tps_mat_size(int n) {
  const int m = n > 3 ? n - 3 : 0;
  return tps_mat_e_size(m) + int64_t(m) * tps_mat_ld(m);
}

// Explicit saddle-point inverse of a shared pattern of n nodes: X (n x n) in
// rows of tps_x_ld(n) doubles, Y in columns n .. n + 2 of the rows (one
// product with X's rows extended gives w and a), padded so that a pair of
// columns from n + 2 stays in the row.
__host__ __device__ __forceinline__ int64_t tps_x_ld(int n) { return (int64_t(n) + 4 + 15) & ~int64_t(15); }
__host__ __device__ __forceinline__ int64_t tps_x_size(int n) {
  return n > 0 ? int64_t(n) * tps_x_ld(n) : 0;
}

// The kernel function at offsets that are multiples of a lattice step:
// table[(a_max + 1) b + a] = tps_phi(tps_r2(a step, b step)) for
// 0 <= a, b <= a_max (a from |dx|, b from |dy|: neighbouring points along a
// row read neighbouring entries). Where both |dx| and |dy| are exactly a step, b step
// with a, b <= a_max, the table entry is the value the direct evaluation
// gives (same operands, same expression; the signs drop out of the
// squares); any other difference is evaluated directly.
struct PhiLattice {
  const double* table;
  double step, inv_step;
  int a_max;
  int lds_words;  // nonzero: the kernels copy the table (this many doubles) into LDS
};

constexpr int kLatticeLdsBytes = 16 * 1024;  // largest table staged in LDS

// `table`: the lattice table or its LDS copy
__device__ __forceinline__ double tps_phi_at(double dx, double dy, const PhiLattice& lat,
                                            const double* table) {
  if (table != nullptr) {
    const double ax = fabs(dx), ay = fabs(dy);
    const double fa = rint(ax * lat.inv_step), fb = rint(ay * lat.inv_step);
    if (fa <= double(lat.a_max) && fb <= double(lat.a_max) && fa * lat.step == ax &&
        fb * lat.step == ay)
      return table[int(fb) * (lat.a_max + 1) + int(fa)];
  }
  return tps_phi(tps_r2(dx, dy));
}

// a coordinate on the lattice (the test tps_phi_at applies to differences)
__device__ __forceinline__ bool on_lattice(double v, const PhiLattice& lat) {
  return rint(fabs(v) * lat.inv_step) * lat.step == fabs(v);
}

// the table entry of a difference known to lie on the lattice (every
// center a multiple of the step, differences within a_max): the index
// tps_phi_at computes, without its check (clamped for safety)
__device__ __forceinline__ double tps_phi_on_lattice(double dx, double dy, const PhiLattice& lat,
                                                    const double* table) {
  const int a = min(int(rint(fabs(dx) * lat.inv_step)), lat.a_max);
  const int b = min(int(rint(fabs(dy) * lat.inv_step)), lat.a_max);
  return table[b * (lat.a_max + 1) + a];
}

// the lattice table in LDS at `lds` when it was sized for it (else the
// global table); call before a barrier of the whole workgroup
__device__ __forceinline__ const double* stage_lattice(const PhiLattice& lat, double* lds) {
  if (lat.table == nullptr || lat.lds_words == 0) return lat.table;
  for (int e = int(threadIdx.x); e < lat.lds_words; e += int(blockDim.x)) lds[e] = lat.table[e];
  return lds;
}

__global__ void k_tps_phi_table(double step, int a_max, double* __restrict__ table) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = a_max + 1;
  if (e >= w * w) return;
  const int b = e / w, a = e % w;
  table[e] = tps_phi(tps_r2(double(a) * step, double(b) * step));
}

PhiLattice make_lattice(const piv_tps_lattice* l) {
  PhiLattice out{nullptr, 0.0, 0.0, 0, 0};
  if (l != nullptr && l->table != nullptr && l->step > 0.0 && l->a_max >= 0) {
    out.table = l->table;
    out.step = l->step;
    out.inv_step = 1.0 / l->step;
    out.a_max = l->a_max;
    const int words = (l->a_max + 1) * (l->a_max + 1);
    if (size_t(words) * sizeof(double) <= size_t(kLatticeLdsBytes)) out.lds_words = words;
  }
  return out;
}

template <int NT = kThreads>
__device__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < NT / 64; ++k) s += red[k];
  __syncthreads();
  return s;
}

__device__ double block_max(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = red[0];
  for (int k = 1; k < kThreads / 64; ++k) s = fmax(s, red[k]);
  __syncthreads();
  return s;
}

// refl layout per subdomain: tau[0..2], R upper (R00 R01 R02 R11 R12 R22)
// [3..8], T upper (T00 T01 T02 T11 T12 T22) [9..14]
// vwork layout per slot: V[0..2], Z[0..2]

// ---------------------------------------------------------------------------
// 1. Householder QR of P, matrix B = K + sI, Y = B V, Z
// ---------------------------------------------------------------------------
// GEN: every center on the lattice of the table, which covers the
// differences within a subdomain (the table entry without the lattice
// check); only B[:, :3] is stored, the factorisation generates the rest
template <int NT, bool GEN>
__device__ __forceinline__ void tps_prep_one(
    int b, const double* __restrict__ xs, const double* __restrict__ ys,
    const int64_t* __restrict__ dom_off, const int32_t* __restrict__ slot_pt,
    const int64_t* __restrict__ mat_off, const double* __restrict__ smoothing_d,
    double* __restrict__ mats, double* __restrict__ refl, double* __restrict__ vwork,
    double* __restrict__ slot_xy, int64_t nslots, int32_t* __restrict__ status, int nm,
    PhiLattice lat, const int32_t* __restrict__ order) {
  static_assert(NT >= kThreads && NT % 64 == 0, "block sums over the first kThreads threads");
  constexpr int kWaves = NT / 64;
  const int d = order != nullptr ? order[b] : b;
  if (d < 0) return;  // an order entry for another kernel (solved by patterns)
  const int64_t s0 = dom_off[d];
  const int n = int(dom_off[d + 1] - s0);
  if (n == 0) return;  // a subdomain slot without centers (device-sized batches)
  double* E = mats + mat_off[d];
  const int64_t ld = tps_mat_ld(n - 3);
  double* C = E + tps_mat_e_size(n - 3);
  double* RF = refl + 16 * d;
  double* VW = vwork + 6 * s0;
  // dynamic LDS, nm >= n doubles per array: cy, cx, P (3 columns, then
  // their reflected copy, then Y = B V), V (3 Householder vectors); then
  // the lattice table when it is staged
  extern __shared__ double dyn[];
  double* cy = dyn;
  double* cx = dyn + nm;
#define PM(k, j) dyn[(2 + (k)) * nm + (j)]
#define VM(k, j) dyn[(5 + (k)) * nm + (j)]
  __shared__ double red[NT / 64];
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  // every block-wide sum takes its terms over the first kThreads threads
  // (the others add zeros after them): the sums, hence the results, are the
  // same for every NT, i.e. for every launch size
  const bool rt = tid < kThreads;
  const double smoothing = smoothing_d[d];
  const double* ltab = stage_lattice(lat, dyn + 8 * nm);
  for (int j = tid; j < n; j += NT) {
    const int p = slot_pt[s0 + j];
    cy[j] = ys[p];
    cx[j] = xs[p];
    slot_xy[s0 + j] = xs[p];  // center coordinates by slot, for the evaluation
    slot_xy[nslots + s0 + j] = ys[p];
    PM(0, j) = 1.0;
    PM(1, j) = ys[p];
    PM(2, j) = xs[p];
    VM(0, j) = VM(1, j) = VM(2, j) = 0.0;
  }
  __syncthreads();
  // ---- Householder reflectors (LAPACK dgeqr2 convention) ----
  double tau[3];
  for (int k = 0; k < 3; ++k) {
    double part = 0.0;
    if (rt)
      for (int j = k + 1 + tid; j < n; j += kThreads) part += PM(k, j) * PM(k, j);
    const double sig = block_sum<NT>(part, red);
    const double alpha = PM(k, k);
    double beta, t, scale;
    if (sig == 0.0) {
      beta = alpha;
      t = 0.0;
      scale = 0.0;
    } else {
      beta = -copysign(sqrt(alpha * alpha + sig), alpha);
      t = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    tau[k] = t;
    for (int j = tid; j < n; j += NT) {
      double v = 0.0;
      if (j == k) v = 1.0;
      else if (j > k) v = PM(k, j) * scale;
      VM(k, j) = v;
    }
    __syncthreads();
    if (tid == 0) PM(k, k) = beta;
    // apply H_k to the remaining columns of P
    for (int c = k + 1; c < 3; ++c) {
      double pp = 0.0;
      if (rt)
        for (int j = k + tid; j < n; j += kThreads) pp += VM(k, j) * PM(c, j);
      const double dot = block_sum<NT>(pp, red);
      for (int j = k + tid; j < n; j += NT) PM(c, j) -= t * VM(k, j) * dot;
      __syncthreads();
    }
  }
  // T of the compact WY form (dlarft, forward, columnwise)
  double p01 = 0.0, p02 = 0.0, p12 = 0.0;
  for (int j = tid; rt && j < n; j += kThreads) {
    p01 += VM(0, j) * VM(1, j);
    p02 += VM(0, j) * VM(2, j);
    p12 += VM(1, j) * VM(2, j);
  }
  const double vtv01 = block_sum<NT>(p01, red);
  const double vtv02 = block_sum<NT>(p02, red);
  const double vtv12 = block_sum<NT>(p12, red);
  const double T00 = tau[0], T11 = tau[1], T22 = tau[2];
  const double T01 = -tau[1] * T00 * vtv01;
  // T[0:2, 2] = -tau2 * T[0:2, 0:2] * V[:, 0:2]^T v2
  const double T02 = -tau[2] * (T00 * vtv02 + T01 * vtv12);
  const double T12 = -tau[2] * (T11 * vtv12);
  if (tid == 0) {
    RF[0] = tau[0];
    RF[1] = tau[1];
    RF[2] = tau[2];
    RF[3] = PM(0, 0);
    RF[4] = PM(1, 0);
    RF[5] = PM(2, 0);
    RF[6] = PM(1, 1);
    RF[7] = PM(2, 1);
    RF[8] = PM(2, 2);
    RF[9] = T00;
    RF[10] = T01;
    RF[11] = T02;
    RF[12] = T11;
    RF[13] = T12;
    RF[14] = T22;
    const bool singular = n < 3 || PM(0, 0) == 0.0 || PM(1, 1) == 0.0 || PM(2, 2) == 0.0;
    if (singular) status[d] = 1;
  }
  // ---- B = K + s I and Y = B V: row i by one wavefront, its lanes over
  //      all columns j (B_ij = B_ji bit for bit: the kernel function sees
  //      the squares of the differences), so that Y_i is a wavefront
  //      reduction and nothing is scattered. The lower triangle is stored,
  //      or (GEN) only B[:, :3], the projection terms of the solves, when
  //      the factorisation generates the other elements with the same
  //      values where it first reads them.
  //      Y goes to the P slots (P is no longer used: R and the reflectors
  //      are saved above) ----
  double m9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int lane = tid & 63, wave = tid >> 6;
  __syncthreads();  // thread 0 has read R out of P
  for (int i = wave; i < n; i += kWaves) {
    const double yi = cy[i], xi = cx[i];
    double* Ei = E + 3 * int64_t(i - 3);
    double* Ci = C + int64_t(i - 3) * ld - 3;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    for (int j = lane; j < n; j += 64) {
      double b = GEN ? tps_phi_on_lattice(cx[j] - xi, cy[j] - yi, lat, ltab)
                     : tps_phi_at(cx[j] - xi, cy[j] - yi, lat, ltab);
      if (i == j) b += smoothing;
      if (i >= 3) {  // rows 0..2 of B are not used by the solves
        if (j < 3) Ei[j] = b;
        else if (!GEN && j <= i) Ci[j] = b;
      }
      r0 += b * VM(0, j);
      r1 += b * VM(1, j);
      r2 += b * VM(2, j);
    }
    for (int off = 32; off > 0; off >>= 1) {
      r0 += __shfl_xor(r0, off, 64);
      r1 += __shfl_xor(r1, off, 64);
      r2 += __shfl_xor(r2, off, 64);
    }
    if (lane == 0) {
      PM(0, i) = r0;
      PM(1, i) = r1;
      PM(2, i) = r2;
    }
  }
  __syncthreads();
  // ---- M = V^T Y, Z = Y T - 1/2 V T^T M T ----
  for (int j = tid; rt && j < n; j += kThreads) {
    const double y0 = PM(0, j), y1 = PM(1, j), y2 = PM(2, j);
    const double v0 = VM(0, j), v1 = VM(1, j), v2 = VM(2, j);
    m9[0] += v0 * y0; m9[1] += v0 * y1; m9[2] += v0 * y2;
    m9[3] += v1 * y0; m9[4] += v1 * y1; m9[5] += v1 * y2;
    m9[6] += v2 * y0; m9[7] += v2 * y1; m9[8] += v2 * y2;
    // keep Y in the Z slots until M is known
    VW[6 * j + 0] = v0;
    VW[6 * j + 1] = v1;
    VW[6 * j + 2] = v2;
    VW[6 * j + 3] = y0;
    VW[6 * j + 4] = y1;
    VW[6 * j + 5] = y2;
  }
  double m[9];
  for (int e = 0; e < 9; ++e) m[e] = block_sum<NT>(m9[e], red);
  const double T[3][3] = {{T00, T01, T02}, {0.0, T11, T12}, {0.0, 0.0, T22}};
  double MT[3][3], TtMT[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double sacc = 0;
      for (int k = 0; k < 3; ++k) sacc += m[3 * a + k] * T[k][b];
      MT[a][b] = sacc;
    }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double sacc = 0;
      for (int k = 0; k < 3; ++k) sacc += T[k][a] * MT[k][b];
      TtMT[a][b] = sacc;
    }
  for (int j = tid; j < n; j += NT) {
    const double v[3] = {VW[6 * j + 0], VW[6 * j + 1], VW[6 * j + 2]};
    const double y[3] = {VW[6 * j + 3], VW[6 * j + 4], VW[6 * j + 5]};
    for (int b = 0; b < 3; ++b) {
      double z = 0.0;
      for (int k = 0; k < 3; ++k) z += y[k] * T[k][b];
      double corr = 0.0;
      for (int k = 0; k < 3; ++k) corr += v[k] * TtMT[k][b];
      VW[6 * j + 3 + b] = z - 0.5 * corr;
    }
  }
#undef PM
#undef VM
}

// the subdomain order[b] (b if order is NULL) for b = blockIdx.x, or with
// STRIDE for b = blockIdx.x, + gridDim.x, ... < nitems (an order lists its
// subdomains first, then -1: for launches over slots most of which are
// unused; the loop costs registers, so the one-slot form stays for the
// others)
template <int NT, bool GEN, bool STRIDE>
__global__ __launch_bounds__(NT) void k_tps_prep(
    const double* __restrict__ xs, const double* __restrict__ ys,
    const int64_t* __restrict__ dom_off, const int32_t* __restrict__ slot_pt,
    const int64_t* __restrict__ mat_off, const double* __restrict__ smoothing_d,
    double* __restrict__ mats, double* __restrict__ refl, double* __restrict__ vwork,
    double* __restrict__ slot_xy, int64_t nslots, int32_t* __restrict__ status, int nm,
    PhiLattice lat, const int32_t* __restrict__ order, int nitems) {
  if constexpr (!STRIDE) {
    tps_prep_one<NT, GEN>(blockIdx.x, xs, ys, dom_off, slot_pt, mat_off, smoothing_d, mats, refl,
                          vwork, slot_xy, nslots, status, nm, lat, order);
  } else {
  for (int b = blockIdx.x; b < nitems; b += gridDim.x) {
    if (order != nullptr && order[b] < 0) break;
    tps_prep_one<NT, GEN>(b, xs, ys, dom_off, slot_pt, mat_off, smoothing_d, mats, refl, vwork,
                          slot_xy, nslots, status, nm, lat, order);
    __syncthreads();  // LDS reused by the next subdomain
  }
  }
}

// ---------------------------------------------------------------------------
// 2. B' = B - V Z^T - Z V^T applied on the fly, C = B'[3:, 3:] = L L^T by a
//    left-looking blocked Cholesky (block columns of kNB = 32). For block
//    column J (rows r0 = 32 J .. m):
//    a) update: C(r0:m, J) -= L(r0:m, 0:r0) L(J, 0:r0)^T on the FP64 matrix
//       cores (v_mfma_f64_16x16x4), one 16-row strip per wavefront at a time,
//       operands read from L2, the result written back in place;
//    b) the 32 x 32 diagonal block is factorised and inverted by wavefront 0
//       in registers;
//    c) panel: L(r0+32:m, J) = C(r0+32:m, J) inv(L_JJ)^T, again on the
//       matrix cores with inv(L_JJ) in LDS.
//    The inverse of every diagonal block is kept, transposed, in the unused
//    upper triangle of that block (the diagonal keeps L_ii) for the
//    triangular solves.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wz_corr(const double* VW, int i, int j) {
  const double* a = VW + 6 * i;
  const double* b = VW + 6 * j;
  return a[0] * b[3] + a[1] * b[4] + a[2] * b[5] + a[3] * b[0] + a[4] * b[1] + a[5] * b[2];
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef double dvec4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[4 v + l / 16][l % 16], v = 0..3
__device__ __forceinline__ dvec4 mfma_f64(double a, double b, dvec4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <int NT>
__device__ __forceinline__ void tps_chol_one(int b, const int64_t* __restrict__ dom_off,
                                             const int64_t* __restrict__ mat_off,
                                             double* __restrict__ mats,
                                             const double* __restrict__ vwork,
                                             int32_t* __restrict__ status,
                                             const double* __restrict__ slot_xy, int64_t nslots,
                                             const double* __restrict__ smoothing_d,
                                             PhiLattice lat, bool gen,
                                             const int32_t* __restrict__ order) {
  const int d = order != nullptr ? order[b] : b;
  if (d < 0) return;  // an order entry for another kernel (solved by patterns)
  const int64_t s0 = dom_off[d];
  const int n = int(dom_off[d + 1] - s0);
  if (n == 0) return;  // a subdomain slot without centers (device-sized batches)
  const int m = n - 3;
  double* E = mats + mat_off[d];
  const int64_t ld = tps_mat_ld(m);
  double* A = E + tps_mat_e_size(m);  // C(i, j) = A[i ld + j]
  // C before any update is B'[3:, 3:] without the rank-6 term, i.e. the
  // kernel function of two centers plus s on the diagonal. With `gen` (every
  // center on the lattice of the table) it is generated where it is first
  // read and the matrix build stores only B[:, :3]; else it is read from
  // the matrix the build stored
  extern __shared__ double chol_dyn[];  // the lattice table, when staged
  const double* ltab = stage_lattice(lat, chol_dyn);
  const double* SX = slot_xy + s0 + 3;
  const double* SY = slot_xy + nslots + s0 + 3;
  const double smoothing = smoothing_d[d];
  auto fresh = [&](int row, int col) {
    double b = tps_phi_on_lattice(SX[col] - SX[row], SY[col] - SY[row], lat, ltab);
    if (row == col) b += smoothing;
    return b;
  };
  const double* VW = vwork + 6 * s0;
  constexpr int kCholThreads = NT, kCholWaves = NT / 64;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;  // MFMA operand row / k group
  __shared__ double inv[kNB][kNB + 1];       // inv(L_JJ)
  __shared__ double lbuf[kNB][kNB + 1];      // L_JJ
  __shared__ double rdiag[kNB];              // 1 / L_ii
  __shared__ double colj[64];                // pivot column of the diagonal factorisation
  __shared__ double tbuf[16][17];            // L21 X11
  __shared__ double ptile[kCholWaves][16][kNB + 1];  // per-wavefront strip transposition
  __shared__ int bad;
  __shared__ int next_pair;                  // look-ahead work counter
  if (tid == 0) bad = 0;
  // B'[3:, :3] (kept for the solves)
  for (int e = tid; e < m * 3; e += kCholThreads) E[e] -= wz_corr(VW, 3 + e / 3, e % 3);
  const double* VC = VW + 18;          // V, Z rows of C's index space
  const int nblk = (m + kNB - 1) / kNB;
  // C(r0t:m, Jt) -= [rank-6 term] + L(r0t:m, kb:ke) L(Jt, kb:ke)^T for the
  // strip pair sp (rows r0t + 32 sp .. + 31) on the FP64 matrix cores; with
  // keep_diag, pair 0 also leaves the updated diagonal block in lbuf
  auto update_pair = [&](int Jt, int sp, int kb, int ke, bool with_wz, bool keep_diag,
                         bool first_touch) {
    const int r0t = Jt * kNB;
    const int bst = min(kNB, m - r0t);
    {
      const int row0 = r0t + 32 * sp;
      const int ra0 = min(row0 + li, m - 1);       // A operand rows (clamped)
      const int ra1 = min(row0 + 16 + li, m - 1);
      const int rb0 = min(r0t + li, m - 1);        // B operand rows: block Jt
      const int rb1 = min(r0t + 16 + li, m - 1);
      const double* pa0 = A + int64_t(ra0) * ld + 4 * lk;
      const double* pa1 = A + int64_t(ra1) * ld + 4 * lk;
      const double* pb0 = A + int64_t(rb0) * ld + 4 * lk;
      const double* pb1 = A + int64_t(rb1) * ld + 4 * lk;
      // the current values of the strip pair's block (rows 4 v + lk (+16),
      // columns li (+16)), requested with the first operands: this
      // wavefront alone writes them during the call
      double cur[2][4][2];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = min(row0 + 16 * h + 4 * v + lk, m - 1);
          if (first_touch && gen) {
            cur[h][v][0] = fresh(row, r0t + min(li, bst - 1));
            cur[h][v][1] = fresh(row, r0t + min(16 + li, bst - 1));
          } else {
            const double* src = A + int64_t(row) * ld + r0t;
            cur[h][v][0] = src[min(li, bst - 1)];
            cur[h][v][1] = src[min(16 + li, bst - 1)];
          }
        }
      dvec4 c00 = {0.0, 0.0, 0.0, 0.0}, c01 = c00, c10 = c00, c11 = c00;
      // the rank-6 term V Z^T + Z V^T of B' as two more k-steps:
      // a_i = [V_i, Z_i], b_j = [Z_j, V_j] (k < 6, zero beyond)
#pragma unroll
      for (int s2 = 0; s2 < 2; ++s2) {
        if (!with_wz) break;
        const int k = 4 * s2 + lk;
        const bool kin = k < 6;
        const int ka = kin ? k : 0, kb = kin ? (k + 3) % 6 : 0;
        const double a0 = kin ? VC[6 * ra0 + ka] : 0.0;
        const double a1 = kin ? VC[6 * ra1 + ka] : 0.0;
        const double b0 = kin ? VC[6 * rb0 + kb] : 0.0;
        const double b1 = kin ? VC[6 * rb1 + kb] : 0.0;
        c00 = mfma_f64(a0, b0, c00);
        c01 = mfma_f64(a0, b1, c01);
        c10 = mfma_f64(a1, b0, c10);
        c11 = mfma_f64(a1, b1, c11);
      }
      // k = k0 + 4 lk + q: the lane reads 4 consecutive columns per operand;
      // the operands of step k0 + 16 are requested while step k0 runs
      double op[2][16];
      auto fetch = [&](double* o, int k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          o[q] = pa0[k0 + q];
          o[4 + q] = pa1[k0 + q];
          o[8 + q] = pb0[k0 + q];
          o[12 + q] = pb1[k0 + q];
        }
      };
      if (kb < ke) fetch(op[0], kb);
      for (int k0 = kb; k0 < ke; k0 += 32) {
        // two steps per iteration, ping-pong buffers
        if (k0 + 16 < ke) fetch(op[1], k0 + 16);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          c00 = mfma_f64(op[0][q], op[0][8 + q], c00);
          c01 = mfma_f64(op[0][q], op[0][12 + q], c01);
          c10 = mfma_f64(op[0][4 + q], op[0][8 + q], c10);
          c11 = mfma_f64(op[0][4 + q], op[0][12 + q], c11);
        }
        if (k0 + 16 >= ke) break;
        if (k0 + 32 < ke) fetch(op[0], k0 + 32);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          c00 = mfma_f64(op[1][q], op[1][8 + q], c00);
          c01 = mfma_f64(op[1][q], op[1][12 + q], c01);
          c10 = mfma_f64(op[1][4 + q], op[1][8 + q], c10);
          c11 = mfma_f64(op[1][4 + q], op[1][12 + q], c11);
        }
      }
      // C(row, col) -= (wz + update), rows 4 v + lk (+16), columns li (+16)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = row0 + 16 * h + 4 * v + lk;
          if (row < m) {
            double* dst = A + int64_t(row) * ld + r0t;
            const int col0 = li, col1 = 16 + li;
            const double u0 = h ? c10[v] : c00[v];
            const double u1 = h ? c11[v] : c01[v];
            if (col0 < bst) {
              const double r = cur[h][v][0] - u0;
              dst[col0] = r;
              if (keep_diag && sp == 0) lbuf[16 * h + 4 * v + lk][col0] = r;  // diagonal block
            }
            if (col1 < bst) {
              const double r = cur[h][v][1] - u1;
              dst[col1] = r;
              if (keep_diag && sp == 0) lbuf[16 * h + 4 * v + lk][col1] = r;
            }
          }
        }
      }
    }
  };
  auto npairs_of = [&](int Jt) { return ((m - Jt * kNB + 15) / 16 + 1) / 2; };
  if (tid == 0) next_pair = 0;
  __syncthreads();
  for (int J = 0; J < nblk; ++J) {
    const int r0 = J * kNB;
    const int bs = min(kNB, m - r0);
    const int nstrip = (m - r0 + 15) / 16;
    // a) the diagonal block of column J finished by wavefront 0: its rank-32
    //    update with block J - 1 (the earlier blocks were applied while the
    //    previous diagonal block was factorised), or for J = 0 the rank-6
    //    term alone; the rows below are finished in the fused panel (c)
    // b) diagonal block by wavefront 0 (C_JJ is in lbuf, written there by
    //    this wavefront in a): right-looking factorisation in registers
    //    (lane i < 32 holds row i; column j of the trailing matrix is
    //    published in LDS at step j and read back by every lane: a_ic -=
    //    (a_ij / a_jj) a_cj, l_ij = a_ij / sqrt(a_jj)), then inv(L_JJ) by
    //    16 x 16 blocks: the two diagonal blocks by forward substitution
    //    in registers (lane c: column c of the upper one, lane 16 + c:
    //    column c of the lower one), the off-diagonal block -X22 (L21 X11)
    //    on the matrix cores
    if (wave == 0) {
      update_pair(J, 0, J > 0 ? r0 - kNB : 0, r0, J == 0, true, J == 0);
      wave_sync();
      // the lane index through an opaque move: the lane-dependent selects
      // below stay in this block instead of being hoisted out of the J loop
      int dl;
      __asm__ volatile("v_mov_b32 %0, %1" : "=v"(dl) : "v"(lane));
      const bool drow = dl < kNB;
      bool singular = false;
      {
        double acc[kNB];
#pragma unroll
        for (int c = 0; c < kNB; ++c)
          acc[c] = (dl < bs && c <= dl && c < bs) ? lbuf[dl < kNB ? dl : 0][c] : 0.0;
#pragma unroll
        for (int j = 0; j < kNB; ++j) {
          // (lanes 32..63 hold zero rows and run along without branches)
          colj[dl] = acc[j];
          wave_sync();
          double col[kNB];  // the pivot column, read at once
#pragma unroll
          for (int c = j; c < kNB; ++c) col[c] = colj[c];
          double dj = col[j];
          if (j >= bs) dj = 1.0;
          if (!(dj > 0.0)) {
            singular = true;
            dj = 1.0;
          }
          const double inv_d = rcp_d(dj);
          const double rj = rsqrt_d(dj);
          // column j of L_JJ (lane j: L_jj = dj rj), trailing update
          const double t = acc[j] * inv_d;
#pragma unroll
          for (int c = j + 1; c < kNB; ++c) acc[c] -= t * col[c];
          acc[j] = j < bs ? acc[j] * rj : (dl == j ? 1.0 : 0.0);
          wave_sync();
        }
        if (drow) {
#pragma unroll
          for (int c = 0; c < kNB; ++c) lbuf[dl][c] = c <= dl ? acc[c] : 0.0;
          rdiag[dl] = rcp_d(lbuf[dl][dl]);  // written by this lane just above
        }
      }
      if (singular && dl == 0) bad = 1;
      wave_sync();
      if (drow) {
        // X11 (o = 0) and X22 (o = 16): column c, x_i = (delta_ic -
        // sum_{k < i} L_ik x_k) / L_ii (x_k = 0 for k < c)
        const int o = dl < 16 ? 0 : 16, c = dl - o;
        double xc[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          double sx = i == c ? 1.0 : 0.0;
#pragma unroll
          for (int k = 0; k < i; ++k) sx -= lbuf[o + i][o + k] * xc[k];
          xc[i] = i < c ? 0.0 : sx * rdiag[o + i];
          inv[o + i][o + c] = xc[i];
          inv[o + c][16 - o + i] = 0.0;  // the strictly upper 16 x 16 block stays zero
          __asm__ volatile("" ::: "memory");  // keep the row loads of lbuf in their step
        }
      }
      wave_sync();
      {
        // T = L21 X11, then X21 = -X22 T (16 x 16 x 16 each, k = 4 s + lk)
        dvec4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
          t = mfma_f64(lbuf[16 + li][4 * s4 + lk], inv[4 * s4 + lk][li], t);
#pragma unroll
        for (int v = 0; v < 4; ++v) tbuf[4 * v + lk][li] = t[v];
        wave_sync();
        dvec4 x21 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
          x21 = mfma_f64(inv[16 + li][16 + 4 * s4 + lk], tbuf[4 * s4 + lk][li], x21);
#pragma unroll
        for (int v = 0; v < 4; ++v) inv[16 + 4 * v + lk][li] = -x21[v];
      }
      wave_sync();
      if (dl < bs) {
        // L_JJ (lower with the diagonal) and inv(L_JJ)^T (strictly upper)
        // into the matrix, row `dl`
        double* blk = A + int64_t(r0 + dl) * ld + r0;
        for (int c = 0; c < bs; ++c) blk[c] = c <= dl ? lbuf[dl][c] : inv[c][dl];
      }
    }
    // look-ahead: blocks 0..J-1 (and the rank-6 term) applied to block column
    // J + 1 by the other wavefronts, pairs of strips taken from a counter;
    // wavefront 0 joins after the diagonal block
    if (J + 1 < nblk) {
      const int np = npairs_of(J + 1);
      for (;;) {
        int sp = 0;
        if (lane == 0) sp = atomicAdd(&next_pair, 1);
        sp = __builtin_amdgcn_readfirstlane(sp);
        if (sp >= np) break;
        update_pair(J + 1, sp, 0, r0, true, false, true);
      }
    }
    __syncthreads();
    if (tid == 0) next_pair = 0;  // the next look-ahead's counter (read after the barrier below)
    // c) rows below the diagonal block, by strips of 16: the finishing
    //    update of a) for these rows (rank 32 with block J - 1, or the rank-6
    //    term for J = 0) in registers, transposed through the wavefront's LDS
    //    tile, then L(rows, J) = C(rows, J) inv^T, D[i][j] = sum_k C[i][k]
    //    inv[j][k] (k <= j): block column J is read and written once here
    if (bs == kNB) {
      const int kb = J > 0 ? r0 - kNB : 0;
      double(*tile)[kNB + 1] = ptile[wave];
      // the update's right operand (block row J) is the same for every strip
      double ob0[8], ob1[8];
      if (J > 0) {
        const double* pb0 = A + int64_t(r0 + li) * ld + kb + 4 * lk;
        const double* pb1 = A + int64_t(r0 + 16 + li) * ld + kb + 4 * lk;
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            ob0[4 * h + q] = pb0[16 * h + q];
            ob1[4 * h + q] = pb1[16 * h + q];
          }
      } else {
        // rank-6 term V Z^T + Z V^T as two k-steps: b_j = [Z_j, V_j]
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
          const int k = 4 * s2 + lk;
          const bool kin = k < 6;
          const int kbb = kin ? (k + 3) % 6 : 0;
          ob0[s2] = kin ? VC[6 * (r0 + li) + kbb] : 0.0;
          ob1[s2] = kin ? VC[6 * (r0 + 16 + li) + kbb] : 0.0;
        }
      }
      // a strip's current values and update operand (rows clamped below m)
      auto load_strip = [&](int st, double* cur, double* oa) {
        const int row0 = r0 + 16 * st;
        const int ra = min(row0 + li, m - 1);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = min(row0 + 4 * v + lk, m - 1);
          if (J == 0 && gen) {  // column 0 has had no look-ahead: first read
            cur[v] = fresh(row, r0 + li);
            cur[4 + v] = fresh(row, r0 + 16 + li);
          } else {
            const double* src = A + int64_t(row) * ld + r0;
            cur[v] = src[li];
            cur[4 + v] = src[16 + li];
          }
        }
        if (J > 0) {
          const double* pa = A + int64_t(ra) * ld + kb + 4 * lk;
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int q = 0; q < 4; ++q) oa[4 * h + q] = pa[16 * h + q];
        } else {
#pragma unroll
          for (int s2 = 0; s2 < 2; ++s2) {
            const int k = 4 * s2 + lk;
            oa[s2] = k < 6 ? VC[6 * ra + k] : 0.0;
          }
        }
      };
      for (int st = 2 + wave; st < nstrip; st += kCholWaves) {
        const int row0 = r0 + 16 * st;
        double cur[8], oa[8];
        load_strip(st, cur, oa);
        dvec4 u0 = {0.0, 0.0, 0.0, 0.0}, u1 = {0.0, 0.0, 0.0, 0.0};
        if (J > 0) {
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            u0 = mfma_f64(oa[e], ob0[e], u0);
            u1 = mfma_f64(oa[e], ob1[e], u1);
          }
        } else {
#pragma unroll
          for (int s2 = 0; s2 < 2; ++s2) {
            u0 = mfma_f64(oa[s2], ob0[s2], u0);
            u1 = mfma_f64(oa[s2], ob1[s2], u1);
          }
        }
        // the finished rows (layout D[4 v + lk][li]) into the tile, read
        // back as the panel's left operand (row li, column 16 h + 4 lk + q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          tile[4 * v + lk][li] = cur[v] - u0[v];
          tile[4 * v + lk][16 + li] = cur[4 + v] - u1[v];
        }
        wave_sync();
        double a8[8];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int q = 0; q < 4; ++q) a8[4 * h + q] = tile[li][16 * h + 4 * lk + q];
        wave_sync();  // the tile is read before the next strip writes it
        dvec4 c0 = {0.0, 0.0, 0.0, 0.0}, c1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int k = 16 * h + 4 * lk + q;
            c0 = mfma_f64(a8[4 * h + q], inv[li][k], c0);
            c1 = mfma_f64(a8[4 * h + q], inv[16 + li][k], c1);
          }
        // the strip's reads of C(rows, J) are complete before the writes
        // (each strip is read and written by one wavefront only)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = row0 + 4 * v + lk;
          if (row < m) {
            double* dst = A + int64_t(row) * ld + r0;
            dst[li] = c0[v];
            dst[16 + li] = c1[v];
          }
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0 && bad) status[d] = 1;
}

template <int NT, bool STRIDE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void k_tps_chol(
    const int64_t* __restrict__ dom_off, const int64_t* __restrict__ mat_off,
    double* __restrict__ mats, const double* __restrict__ vwork, int32_t* __restrict__ status,
    const double* __restrict__ slot_xy, int64_t nslots, const double* __restrict__ smoothing_d,
    PhiLattice lat, bool gen, const int32_t* __restrict__ order, int nitems) {
  if constexpr (!STRIDE) {
    tps_chol_one<NT>(blockIdx.x, dom_off, mat_off, mats, vwork, status, slot_xy, nslots,
                     smoothing_d, lat, gen, order);
  } else {
  for (int b = blockIdx.x; b < nitems; b += gridDim.x) {
    if (order != nullptr && order[b] < 0) break;
    tps_chol_one<NT>(b, dom_off, mat_off, mats, vwork, status, slot_xy, nslots, smoothing_d, lat,
                     gen, order);
    __syncthreads();  // LDS reused by the next subdomain
  }
  }
}

// ---------------------------------------------------------------------------
// 3. solves with the outlier iterations, both components
// ---------------------------------------------------------------------------
// x <- Q^T x (apply H1, H2, H3) or Q x (apply H3, H2, H1), two vectors
__device__ void apply_q(double* x0, double* x1, int n, const double* VW, const double* tau,
                        bool transpose, double* red) {
  for (int kk = 0; kk < 3; ++kk) {
    const int k = transpose ? kk : 2 - kk;
    double p0 = 0.0, p1 = 0.0;
    for (int j = threadIdx.x; j < n; j += kThreads) {
      p0 += VW[6 * j + k] * x0[j];
      p1 += VW[6 * j + k] * x1[j];
    }
    const double d0 = block_sum(p0, red), d1 = block_sum(p1, red);
    for (int j = threadIdx.x; j < n; j += kThreads) {
      x0[j] -= tau[k] * VW[6 * j + k] * d0;
      x1[j] -= tau[k] * VW[6 * j + k] * d1;
    }
    __syncthreads();
  }
}

// solve C z = b in place (x[c][3 + i], c = 0, 1) with C = L L^T, using the
// inverted diagonal blocks (strict upper triangle of each diagonal block
// holds inv(L_JJ)^T, the diagonal holds L_ii). Both sweeps are left-looking
// so that every block step is one coalesced pass over a block row (forward)
// or a block column (backward) of L, reduced across threads, followed by a
// 32 x 32 triangular matrix-vector product out of LDS. The sweeps are
// latency-bound: every thread issues its loads of a step in batches
// (kBatch doubles in flight) before it uses them, and the sums keep the
// order of one element after the other.
constexpr int kBatch = 16;

// the diagonal block J of L into LDS (zero outside the bs x bs block)
__device__ __forceinline__ void load_diag_block(double (*dblk)[kNB + 1], const double* A,
                                                int64_t ld, int r0, int bs) {
  double v[kNB * kNB / kThreads];
#pragma unroll
  for (int t = 0; t < kNB * kNB / kThreads; ++t) {
    const int e = threadIdx.x + kThreads * t;
    const int i = e / kNB, c = e % kNB;
    const int ii = min(i, bs - 1), cc = min(c, bs - 1);
    v[t] = A[int64_t(r0 + ii) * ld + r0 + cc];
  }
#pragma unroll
  for (int t = 0; t < kNB * kNB / kThreads; ++t) {
    const int e = threadIdx.x + kThreads * t;
    const int i = e / kNB, c = e % kNB;
    dblk[i][c] = (i < bs && c < bs) ? v[t] : 0.0;
  }
}

__device__ void chol_solve(double* xa, double* xb, const double* A, int64_t ld, int m) {
  static_assert(kThreads == 256 && kNB == 32, "thread layout of chol_solve");
  __shared__ double dblk[kNB][kNB + 1];
  __shared__ double tmp[2][kNB];
  __shared__ double part[2][8][kNB];
  const int tid = threadIdx.x;
  const int nblk = (m + kNB - 1) / kNB;
  double* x0 = xa + 3;
  double* x1 = xb + 3;
  // forward: y_J = inv(L_JJ) (b_J - L(J, 0:r0) y(0:r0))
  for (int J = 0; J < nblk; ++J) {
    const int r0 = J * kNB;
    const int bs = min(kNB, m - r0);
    load_diag_block(dblk, A, ld, r0, bs);
    {
      // thread = (row i, part p): columns 32 q + 4 p .. + 3 of row r0 + i,
      // kBatch / 4 blocks q per batch of loads
      const int i = tid >> 3, p = tid & 7;
      double s0 = 0.0, s1 = 0.0;
      if (i < bs) {
        const double* Lr = A + int64_t(r0 + i) * ld + 4 * p;
        constexpr int QB = kBatch / 4;
        for (int q0 = 0; q0 < J; q0 += QB) {
          double l[kBatch];
#pragma unroll
          for (int qq = 0; qq < QB; ++qq) {
            const int q = min(q0 + qq, J - 1);
#pragma unroll
            for (int u = 0; u < 4; ++u) l[4 * qq + u] = Lr[32 * q + u];
          }
#pragma unroll
          for (int qq = 0; qq < QB; ++qq) {
            if (q0 + qq < J) {
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const int k = 32 * (q0 + qq) + 4 * p + u;
                s0 += l[4 * qq + u] * x0[k];
                s1 += l[4 * qq + u] * x1[k];
              }
            }
          }
        }
      }
#pragma unroll
      for (int o = 4; o >= 1; o >>= 1) {
        s0 += __shfl_xor(s0, o, 64);
        s1 += __shfl_xor(s1, o, 64);
      }
      if (p == 0 && i < bs) {
        tmp[0][i] = x0[r0 + i] - s0;
        tmp[1][i] = x1[r0 + i] - s1;
      }
    }
    __syncthreads();
    if (tid < 64) {
      const int c = tid & 31, comp = tid >> 5;
      if (c < bs) {
        // y_c = tmp_c / L_cc + sum_{k < c} inv(L)[c][k] tmp_k
        double sacc = tmp[comp][c] / dblk[c][c];
#pragma unroll 4
        for (int k = 0; k < kNB; ++k)
          if (k < c) sacc += dblk[k][c] * tmp[comp][k];
        (comp ? x1 : x0)[r0 + c] = sacc;
      }
    }
    __syncthreads();
  }
  // backward: z_J = inv(L_JJ)^T (y_J - L(r0+bs:m, J)^T z(r0+bs:m))
  for (int J = nblk - 1; J >= 0; --J) {
    const int r0 = J * kNB;
    const int bs = min(kNB, m - r0);
    load_diag_block(dblk, A, ld, r0, bs);
    {
      // thread = (column c, row group g): rows r0 + bs + g, + 8, ...,
      // kBatch rows per batch of loads
      const int c = tid & 31, g = tid >> 5;
      double s0 = 0.0, s1 = 0.0;
      if (c < bs) {
        const double* Lc = A + r0 + c;
        for (int i0 = r0 + bs + g; i0 < m; i0 += 8 * kBatch) {
          double l[kBatch];
#pragma unroll
          for (int u = 0; u < kBatch; ++u) l[u] = Lc[int64_t(min(i0 + 8 * u, m - 1)) * ld];
#pragma unroll
          for (int u = 0; u < kBatch; ++u) {
            const int i = i0 + 8 * u;
            if (i < m) {
              s0 += l[u] * x0[i];
              s1 += l[u] * x1[i];
            }
          }
        }
      }
      part[0][g][c] = s0;
      part[1][g][c] = s1;
    }
    __syncthreads();
    if (tid < 64) {
      const int c = tid & 31, comp = tid >> 5;
      double sacc = 0.0;
#pragma unroll
      for (int g = 0; g < 8; ++g) sacc += part[comp][g][c];
      tmp[comp][c] = c < bs ? (comp ? x1 : x0)[r0 + c] - sacc : 0.0;
    }
    __syncthreads();
    if (tid < 64) {
      const int c = tid & 31, comp = tid >> 5;
      if (c < bs) {
        // z_c = tmp_c / L_cc + sum_{k > c} inv(L)[k][c] tmp_k
        double sacc = tmp[comp][c] / dblk[c][c];
#pragma unroll 4
        for (int k = 0; k < kNB; ++k)
          if (k > c && k < bs) sacc += dblk[c][k] * tmp[comp][k];
        (comp ? x1 : x0)[r0 + c] = sacc;
      }
    }
    __syncthreads();
  }
}

// ---- the same solves for NV right-hand sides at once (x[v nm + j] in LDS):
// one pass over the factor serves them all ----
// block sums of NV values (NT = kThreads)
template <int NV>
__device__ void block_sum_n(double v[NV], double* red) {
#pragma unroll
  for (int q = 0; q < NV; ++q)
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) red[q * (kThreads / 64) + (threadIdx.x >> 6)] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double t = 0.0;
    for (int k = 0; k < kThreads / 64; ++k) t += red[q * (kThreads / 64) + k];
    v[q] = t;
  }
  __syncthreads();
}

template <int NV>
__device__ void apply_q_n(double* x, int nm, int n, const double* VW, const double* tau,
                          bool transpose, double* red) {
  for (int kk = 0; kk < 3; ++kk) {
    const int k = transpose ? kk : 2 - kk;
    double p[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) p[q] = 0.0;
    for (int j = threadIdx.x; j < n; j += kThreads) {
      const double vk = VW[6 * j + k];
#pragma unroll
      for (int q = 0; q < NV; ++q) p[q] += vk * x[q * nm + j];
    }
    block_sum_n<NV>(p, red);
    for (int j = threadIdx.x; j < n; j += kThreads) {
      const double vk = tau[k] * VW[6 * j + k];
#pragma unroll
      for (int q = 0; q < NV; ++q) x[q * nm + j] -= vk * p[q];
    }
    __syncthreads();
  }
}

// LDS scratch of chol_solve_n (doubles): the diagonal block, the block
// right-hand sides and the partial sums of up to 8 vectors
constexpr int kSolveScratch = kNB * (kNB + 1) + 8 * kNB + 8 * 8 * kNB;

// C z = b in place for NV vectors (x[q nm + 3 + i]); as chol_solve. scr:
// kSolveScratch doubles of LDS.
template <int NV>
__device__ void chol_solve_n(double* x, int nm, const double* A, int64_t ld, int m, double* scr) {
  static_assert(kThreads == 256 && kNB == 32 && NV <= 8, "thread layout of chol_solve_n");
  double(*dblk)[kNB + 1] = reinterpret_cast<double(*)[kNB + 1]>(scr);
  double(*tmp)[kNB] = reinterpret_cast<double(*)[kNB]>(scr + kNB * (kNB + 1));
  double(*part)[8][kNB] = reinterpret_cast<double(*)[8][kNB]>(scr + kNB * (kNB + 1) + 8 * kNB);
  const int tid = threadIdx.x;
  const int nblk = (m + kNB - 1) / kNB;
  double* xs = x + 3;
  for (int J = 0; J < nblk; ++J) {
    const int r0 = J * kNB;
    const int bs = min(kNB, m - r0);
    load_diag_block(dblk, A, ld, r0, bs);
    {
      const int i = tid >> 3, p = tid & 7;
      double sv[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) sv[q] = 0.0;
      if (i < bs) {
        const double* Lr = A + int64_t(r0 + i) * ld + 4 * p;
        constexpr int QB = kBatch / 4;
        for (int q0 = 0; q0 < J; q0 += QB) {
          double l[kBatch];
#pragma unroll
          for (int qq = 0; qq < QB; ++qq) {
            const int qb = min(q0 + qq, J - 1);
#pragma unroll
            for (int u = 0; u < 4; ++u) l[4 * qq + u] = Lr[32 * qb + u];
          }
#pragma unroll
          for (int qq = 0; qq < QB; ++qq) {
            if (q0 + qq < J) {
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const int k = 32 * (q0 + qq) + 4 * p + u;
#pragma unroll
                for (int q = 0; q < NV; ++q) sv[q] += l[4 * qq + u] * xs[q * nm + k];
              }
            }
          }
        }
      }
#pragma unroll
      for (int q = 0; q < NV; ++q)
#pragma unroll
        for (int o = 4; o >= 1; o >>= 1) sv[q] += __shfl_xor(sv[q], o, 64);
      if (p == 0 && i < bs)
#pragma unroll
        for (int q = 0; q < NV; ++q) tmp[q][i] = xs[q * nm + r0 + i] - sv[q];
    }
    __syncthreads();
    if (tid < 32 * NV) {
      const int c = tid & 31, comp = tid >> 5;
      if (c < bs) {
        double sacc = tmp[comp][c] / dblk[c][c];
#pragma unroll 4
        for (int k = 0; k < kNB; ++k)
          if (k < c) sacc += dblk[k][c] * tmp[comp][k];
        xs[comp * nm + r0 + c] = sacc;
      }
    }
    __syncthreads();
  }
  for (int J = nblk - 1; J >= 0; --J) {
    const int r0 = J * kNB;
    const int bs = min(kNB, m - r0);
    load_diag_block(dblk, A, ld, r0, bs);
    {
      const int c = tid & 31, g = tid >> 5;
      double sv[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) sv[q] = 0.0;
      if (c < bs) {
        const double* Lc = A + r0 + c;
        for (int i0 = r0 + bs + g; i0 < m; i0 += 8 * kBatch) {
          double l[kBatch];
#pragma unroll
          for (int u = 0; u < kBatch; ++u) l[u] = Lc[int64_t(min(i0 + 8 * u, m - 1)) * ld];
#pragma unroll
          for (int u = 0; u < kBatch; ++u) {
            const int i = i0 + 8 * u;
            if (i < m)
#pragma unroll
              for (int q = 0; q < NV; ++q) sv[q] += l[u] * xs[q * nm + i];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < NV; ++q) part[q][g][c] = sv[q];
    }
    __syncthreads();
    if (tid < 32 * NV) {
      const int c = tid & 31, comp = tid >> 5;
      double sacc = 0.0;
#pragma unroll
      for (int g = 0; g < 8; ++g) sacc += part[comp][g][c];
      tmp[comp][c] = c < bs ? xs[comp * nm + r0 + c] - sacc : 0.0;
    }
    __syncthreads();
    if (tid < 32 * NV) {
      const int c = tid & 31, comp = tid >> 5;
      if (c < bs) {
        double sacc = tmp[comp][c] / dblk[c][c];
#pragma unroll 4
        for (int k = 0; k < kNB; ++k)
          if (k > c && k < bs) sacc += dblk[c][k] * tmp[comp][k];
        xs[comp * nm + r0 + c] = sacc;
      }
    }
    __syncthreads();
  }
}

// saddle_solve for NV right-hand sides x[q nm + j]; the polynomial terms go
// to aout[q][0..2] (LDS, written by thread 0, read after the call)
// (w, a) = M^-1 [x; 0] for NV right-hand sides (see saddle_solve); scr:
// solve_scratch(NV) doubles of LDS.
template <int NV>
__device__ void saddle_solve_n(double* x, int nm, int n, const double* VW, const double* tau,
                               const double* A, int64_t ld, const double* E, const double* RF,
                               double* red, double* scr, double (*aout)[3]) {
  const int tid = threadIdx.x;
  const int m = n - 3;
  // Q^T x; the first three entries (the a-part of the rhs) are left alone by the
  // Cholesky solve of the others
  apply_q_n<NV>(x, nm, n, VW, tau, true, red);
  if (m > 0) {
    chol_solve_n<NV>(x, nm, A, ld, m, scr);
  }
  // a = R^-1 (vt - B'[:3, 3:] z), B'[:3, 3 + j] = B'[3 + j][k] = E[j][k]
  for (int k = 0; k < 3; ++k) {
    double bz[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) bz[q] = 0.0;
    for (int j = tid; j < m; j += kThreads) {
      const double b = E[3 * int64_t(j) + k];
#pragma unroll
      for (int q = 0; q < NV; ++q) bz[q] += b * x[q * nm + 3 + j];
    }
    block_sum_n<NV>(bz, red);
    if (tid == 0)
#pragma unroll
      for (int q = 0; q < NV; ++q) aout[q][k] = x[q * nm + k] - bz[q];
  }
  if (tid == 0)
    for (int q = 0; q < NV; ++q) {
      const double r0 = aout[q][0], r1 = aout[q][1], r2 = aout[q][2];
      const double a2 = r2 / RF[8];
      const double a1 = (r1 - RF[7] * a2) / RF[6];
      aout[q][0] = (r0 - RF[4] * a1 - RF[5] * a2) / RF[3];
      aout[q][1] = a1;
      aout[q][2] = a2;
      x[q * nm] = x[q * nm + 1] = x[q * nm + 2] = 0.0;
    }
  __syncthreads();
  apply_q_n<NV>(x, nm, n, VW, tau, false, red);
}

// (w, a) = M^-1 [x; 0] for the two right-hand sides x0, x1 (length n, in
// LDS) of the saddle-point system M = [[K + sI, P], [P^T, 0]] of n centers,
// by the null-space factorisation of M (VW, tau, RF: the reflectors, R and
// T; A, ld: the Cholesky factor of C; E = B'[3:, :3]): on return x0, x1 hold
// w and a[c] the polynomial terms [1, y, x]. Every thread of the block
// calls it.
__device__ void saddle_solve(double* x0, double* x1, int n, const double* VW, const double* tau,
                             const double* A, int64_t ld, const double* E, const double* RF,
                             double* red, double a[2][3]) {
  const int tid = threadIdx.x;
  const int m = n - 3;
  double* x[2] = {x0, x1};
  // rhs in the rotated basis; its a-part kept in registers: vt[0..2]
  apply_q(x0, x1, n, VW, tau, true, red);
  double vt[2][3];
  for (int c = 0; c < 2; ++c)
    for (int k = 0; k < 3; ++k) vt[c][k] = x[c][k];
  __syncthreads();
  if (m > 0) chol_solve(x0, x1, A, ld, m);
  // a = R^-1 (vt - B'[:3, 3:] z), B'[:3, 3 + j] = B'[3 + j][k] = E[j][k]
  double bz[2][3];
  for (int k = 0; k < 3; ++k) {
    double p0 = 0.0, p1 = 0.0;
    for (int j = tid; j < m; j += kThreads) {
      const double b = E[3 * int64_t(j) + k];
      p0 += b * x0[3 + j];
      p1 += b * x1[3 + j];
    }
    bz[0][k] = block_sum(p0, red);
    bz[1][k] = block_sum(p1, red);
  }
  for (int c = 0; c < 2; ++c) {
    const double r0 = vt[c][0] - bz[c][0], r1 = vt[c][1] - bz[c][1], r2 = vt[c][2] - bz[c][2];
    a[c][2] = r2 / RF[8];
    a[c][1] = (r1 - RF[7] * a[c][2]) / RF[6];
    a[c][0] = (r0 - RF[4] * a[c][1] - RF[5] * a[c][2]) / RF[3];
  }
  // w = Q [0; z]
  if (tid < 2) {
    x[tid][0] = x[tid][1] = x[tid][2] = 0.0;
  }
  __syncthreads();
  apply_q(x0, x1, n, VW, tau, false, red);
}

__device__ __forceinline__ void tps_solve_one(
    int b, const float* __restrict__ vx, const float* __restrict__ vy,
    const int64_t* __restrict__ dom_off, const int32_t* __restrict__ slot_pt,
    const int64_t* __restrict__ mat_off, const double* __restrict__ mats,
    const double* __restrict__ refl, const double* __restrict__ vwork,
    const double* __restrict__ smoothing_d,
    double threshold, double* __restrict__ smooth_slots, int64_t nslots,
    double* __restrict__ weights, int64_t wstride, double* __restrict__ stats,
    const int32_t* __restrict__ status, int nm, const int32_t* __restrict__ order) {
  const int d = order != nullptr ? order[b] : b;
  if (d < 0) return;  // an order entry for another kernel (solved by patterns)
  if (status[d]) return;
  const int64_t s0 = dom_off[d];
  const int n = int(dom_off[d + 1] - s0);
  if (n == 0) return;  // a subdomain slot without centers (device-sized batches)
  const int m = n - 3;
  const double* E = mats + mat_off[d];
  const int64_t ld = tps_mat_ld(m);
  const double* A = E + tps_mat_e_size(m);
  const double* RF = refl + 16 * d;
  const double* VW = vwork + 6 * s0;
  const double smoothing = smoothing_d[d];
  const int tid = threadIdx.x;
  // dynamic LDS (nm >= n, even): work vectors x[2][nm], then the current
  // values val[2][nm] (float32, as the reference)
  extern __shared__ double dyn[];
  double* x[2] = {dyn, dyn + nm};
  float* val[2] = {reinterpret_cast<float*>(dyn + 2 * nm),
                   reinterpret_cast<float*>(dyn + 2 * nm) + nm};
  __shared__ double red[kThreads / 64];
  __shared__ int go[2];
  __shared__ int nerr[2];
  const double tau[3] = {RF[0], RF[1], RF[2]};
  for (int j = tid; j < n; j += kThreads) {
    const int p = slot_pt[s0 + j];
    val[0][j] = vx[p];
    val[1][j] = vy[p];
  }
  if (tid < 2) {
    go[tid] = 1;
  }
  __syncthreads();
  double nfixed[2] = {0.0, 0.0}, maxdiff[2] = {0.0, 0.0};
  int count[2] = {0, 0};
  for (int it = 0; it < 11; ++it) {
    if (!go[0] && !go[1]) break;
    // rhs, then (w, a) = M^-1 [v; 0]
    for (int j = tid; j < n; j += kThreads) {
      x[0][j] = val[0][j];
      x[1][j] = val[1][j];
    }
    __syncthreads();
    double a[2][3];
    saddle_solve(x[0], x[1], n, VW, tau, A, ld, E, RF, red, a);
    // smoothed = v - s w, differences, erratic vectors
    for (int c = 0; c < 2; ++c) {
      if (!go[c]) continue;  // uniform across the block
      count[c] += 1;
      double dmax = 0.0;
      int ne = 0;
      for (int j = tid; j < n; j += kThreads) {
        const double v = double(val[c][j]);
        const double sm = v - smoothing * x[c][j];
        smooth_slots[c * nslots + s0 + j] = sm;
        const double diff = fabs(sm - v);
        dmax = fmax(dmax, diff);
        if (threshold >= 0.0 && diff > threshold) ++ne;
      }
      double* W = weights + c * wstride + s0 + 3 * int64_t(d);
      for (int j = tid; j < n; j += kThreads) W[j] = x[c][j];
      if (tid == 0) {
        W[n] = a[c][0];
        W[n + 1] = a[c][1];
        W[n + 2] = a[c][2];
      }
      const double dm = block_max(dmax, red);
      const double nesum = block_sum(double(ne), red);
      if (count[c] == 1) maxdiff[c] = dm;
      const bool more = threshold >= 0.0 && nesum > 0.0 && count[c] < 11;
      if (more) {
        nfixed[c] += nesum;
        // values[erratic] = smoothed[erratic] (stored as float32)
        for (int j = tid; j < n; j += kThreads) {
          const double v = double(val[c][j]);
          const double sm = v - smoothing * x[c][j];
          if (fabs(sm - v) > threshold) val[c][j] = float(sm);
        }
      }
      __syncthreads();
      if (tid == 0) go[c] = more ? 1 : 0;
      __syncthreads();
    }
  }
  if (tid == 0) {
    for (int c = 0; c < 2; ++c) {
      stats[6 * d + 3 * c + 0] = nfixed[c];
      stats[6 * d + 3 * c + 1] = maxdiff[c];
      stats[6 * d + 3 * c + 2] = count[c];
    }
  }
}

template <bool STRIDE>
__global__ __launch_bounds__(kThreads) void k_tps_solve(
    const float* __restrict__ vx, const float* __restrict__ vy,
    const int64_t* __restrict__ dom_off, const int32_t* __restrict__ slot_pt,
    const int64_t* __restrict__ mat_off, const double* __restrict__ mats,
    const double* __restrict__ refl, const double* __restrict__ vwork,
    const double* __restrict__ smoothing_d, double threshold,
    double* __restrict__ smooth_slots, int64_t nslots, double* __restrict__ weights,
    int64_t wstride, double* __restrict__ stats, const int32_t* __restrict__ status, int nm,
    const int32_t* __restrict__ order, int nitems) {
  if constexpr (!STRIDE) {
    tps_solve_one(blockIdx.x, vx, vy, dom_off, slot_pt, mat_off, mats, refl, vwork, smoothing_d,
                  threshold, smooth_slots, nslots, weights, wstride, stats, status, nm, order);
  } else {
  for (int b = blockIdx.x; b < nitems; b += gridDim.x) {
    if (order != nullptr && order[b] < 0) break;
    tps_solve_one(b, vx, vy, dom_off, slot_pt, mat_off, mats, refl, vwork, smoothing_d,
                  threshold, smooth_slots, nslots, weights, wstride, stats, status, nm, order);
    __syncthreads();  // LDS reused by the next subdomain
  }
  }
}

// ---------------------------------------------------------------------------
// 4. geometry: selection of the finite vectors, subdomain membership
//    (thin_plate_spline_subdom.py:53-126,161-222), with the subdomain grid
//    limits computed on the host from five scalars
// ---------------------------------------------------------------------------
constexpr int kMaxSlots = 4;  // subdomains per point (20 % buffers: <= 2 x 2)

// inclusive prefix sum over a 1024-thread block
__device__ int block_scan_1024(int v, int* tmp, int& total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  if (lane == 63) tmp[wv] = v;
  __syncthreads();
  if (threadIdx.x < 16) {
    int x = tmp[threadIdx.x];
    for (int off = 1; off < 16; off <<= 1) {
      const int u = __shfl_up(x, off, 16);
      if (threadIdx.x >= off) x += u;
    }
    tmp[16 + threadIdx.x] = x;
  }
  __syncthreads();
  total = tmp[16 + 15];
  const int add = wv > 0 ? tmp[16 + wv - 1] : 0;
  __syncthreads();
  return v + add;
}

// one block of 1024 threads: ordered compaction of the vectors whose dx and
// dy are finite (singlepass.py:450-463) and the bounds of their positions;
// consecutive threads read consecutive vectors, four tiles of 1024 per
// barrier round (wavefront ballots, per-wavefront counts in LDS)
__global__ __launch_bounds__(1024) void k_tps_select(int n, const float* __restrict__ dx,
                                                     const float* __restrict__ dy,
                                                     const double* __restrict__ xs,
                                                     const double* __restrict__ ys,
                                                     double* cx, double* cy, float* vx,
                                                     float* vy, int32_t* idx, double* info,
                                                     double lat_step) {
  // field blockIdx.x: elements [f n, (f + 1) n) in and out
  {
    const int64_t o = int64_t(blockIdx.x) * n;
    dx += o;
    dy += o;
    xs += o;
    ys += o;
    cx += o;
    cy += o;
    vx += o;
    vy += o;
    idx += o;
    info += 7 * blockIdx.x;
  }
  const double inv_step = lat_step > 0.0 ? 1.0 / lat_step : 0.0;
  constexpr int kTiles = 4;  // tiles of 1024 vectors per barrier round
  __shared__ int wcount[2][kTiles][16];  // valid vectors per tile and wavefront (double-buffered)
  __shared__ double red[4][16];
  double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
  int base = 0;
  int unsorted = 0;   // some ys[i] > ys[i + 1]
  int off_lattice = lat_step > 0.0 ? 0 : 1;  // some selected position not a multiple of the step
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t below = (uint64_t(1) << lane) - 1;
  for (int t0 = 0, round = 0; t0 < n; t0 += 1024 * kTiles, round ^= 1) {
    float a[kTiles], b[kTiles];
    uint64_t mask[kTiles];
#pragma unroll
    for (int s = 0; s < kTiles; ++s) {
      const int i = t0 + 1024 * s + int(threadIdx.x);
      a[s] = 0.f;
      b[s] = 0.f;
      bool f = false;
      if (i < n) {
        a[s] = dx[i];
        b[s] = dy[i];
        f = a[s] == a[s] && b[s] == b[s];
        if (i + 1 < n && !(ys[i] <= ys[i + 1])) unsorted = 1;
      }
      mask[s] = __ballot(f);
      if (lane == 0) wcount[round][s][wv] = int(__popcll(mask[s]));
    }
    __syncthreads();
    // vector (tile s, lane of wavefront wv) goes after the valid vectors of
    // the tiles before s and of the wavefronts before wv in tile s
    int start = base;
#pragma unroll
    for (int s = 0; s < kTiles; ++s) {
      int in_tile = 0, tile_total = 0;
      for (int w = 0; w < 16; ++w) {
        const int c = wcount[round][s][w];
        in_tile += w < wv ? c : 0;
        tile_total += c;
      }
      const int i = t0 + 1024 * s + int(threadIdx.x);
      if ((mask[s] >> lane) & 1) {
        const int pos = start + in_tile + int(__popcll(mask[s] & below));
        const double x = xs[i], y = ys[i];
        cx[pos] = x;
        cy[pos] = y;
        vx[pos] = a[s];
        vy[pos] = b[s];
        idx[pos] = i;
        if (rint(x * inv_step) * lat_step != x || rint(y * inv_step) * lat_step != y)
          off_lattice = 1;
        xmin = fmin(xmin, x);
        xmax = fmax(xmax, x);
        ymin = fmin(ymin, y);
        ymax = fmax(ymax, y);
      }
      start += tile_total;
    }
    base = start;
  }
  unsorted = __syncthreads_or(unsorted);
  off_lattice = __syncthreads_or(off_lattice);
  for (int off = 32; off > 0; off >>= 1) {
    xmin = fmin(xmin, __shfl_xor(xmin, off, 64));
    xmax = fmax(xmax, __shfl_xor(xmax, off, 64));
    ymin = fmin(ymin, __shfl_xor(ymin, off, 64));
    ymax = fmax(ymax, __shfl_xor(ymax, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = xmin;
    red[1][threadIdx.x >> 6] = xmax;
    red[2][threadIdx.x >> 6] = ymin;
    red[3][threadIdx.x >> 6] = ymax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 16; ++k) {
      red[0][0] = fmin(red[0][0], red[0][k]);
      red[1][0] = fmax(red[1][0], red[1][k]);
      red[2][0] = fmin(red[2][0], red[2][k]);
      red[3][0] = fmax(red[3][0], red[3][k]);
    }
    info[0] = base;
    info[1] = red[0][0];
    info[2] = red[1][0];
    info[3] = red[2][0];
    info[4] = red[3][0];
    info[5] = unsorted ? 0.0 : 1.0;
    info[6] = off_lattice ? 0.0 : 1.0;
  }
}

struct SubGrid {
  int nsx, nsy;
  const double* lim;  // lmin_x[nsx], lmax_x[nsx], lmin_y[nsy], lmax_y[nsy]
  __device__ double xlo(int i) const { return lim[i]; }
  __device__ double xhi(int i) const { return lim[nsx + i]; }
  __device__ double ylo(int i) const { return lim[2 * nsx + i]; }
  __device__ double yhi(int i) const { return lim[2 * nsx + nsy + i]; }
  __device__ bool inside(int d, double x, double y) const {
    const int ix = d % nsx, iy = d / nsx;
    return x >= xlo(ix) && x < xhi(ix) && y >= ylo(iy) && y < yhi(iy);
  }
  // partition-of-unity weight of subdomain d (thin_plate_spline_subdom.py:208-222)
  __device__ double coef(int d, double x, double y) const {
    const int ix = d % nsx, iy = d / nsx;
    const double xc = 0.5 * (xlo(ix) + xhi(ix)), yc = 0.5 * (ylo(iy) + yhi(iy));
    const double dxm = xhi(ix) - xc, dym = yhi(iy) - yc;
    const double dx2 = (x - xc) * (x - xc) / (dxm * dxm);
    const double dy2 = (y - yc) * (y - yc) / (dym * dym);
    return (1 - dx2) * (1 - dy2) + 0.001;
  }
};

__device__ __forceinline__ SubGrid field_grid(const piv_tps_field& f, const double* lim) {
  return SubGrid{f.nsx, f.nsy, lim + f.lim_off};
}

// first index in [0, n) with a[i] >= v (a ascending), n if none, by one
// wavefront: 64 samples a level narrow the range 64-fold, then one ballot
__device__ int wave_lower_bound(const double* a, int n, double v) {
  const int lane = threadIdx.x & 63;
  int lo = 0, hi = n;  // the answer is in [lo, hi]
  while (hi - lo > 64) {
    const int s = (hi - lo + 63) / 64;
    const int j = lo + lane * s;
    const int c = __popcll(__ballot(j < hi && a[j] < v));
    if (c == 0) return lo;
    const int nlo = lo + (c - 1) * s + 1;
    hi = min(lo + c * s, hi);
    lo = nlo;
  }
  const int j = lo + lane;
  return lo + __popcll(__ballot(j < hi && a[j] < v));
}

// the range of centers [first, last) of a field that may lie in subdomain d
// (those with y in [ylo, yhi) when the centers are sorted by y, else all),
// by one wavefront
__device__ void wave_member_range(const piv_tps_field& f, const SubGrid& g, int d,
                                  const double* cy, int& first, int& last) {
  first = 0;
  last = f.count;
  if (f.sorted_y) {
    const int iy = d / g.nsx;
    first = wave_lower_bound(cy, f.count, g.ylo(iy));
    last = wave_lower_bound(cy, f.count, g.yhi(iy));
  }
}

// One workgroup of four wavefronts per (global) subdomain slot: the
// candidates [first, last) split into four contiguous segments, one per
// wavefront, tested 64 at a time with four chunks' loads in flight and
// counted by ballots (no barrier inside a segment)
constexpr int kSlotChunks = 4;

// the members of subdomain d among the candidates [s0, s1) (one wavefront):
// their number, and with WRITE their center indices from slot_pt[base] on,
// in increasing order
template <bool WRITE>
__device__ int segment_members(const SubGrid& g, int d, const double* cx, const double* cy,
                               int s0, int s1, int center_base, int32_t* slot_pt,
                               int64_t base) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  int c = 0;
  for (int c0 = s0; c0 < s1; c0 += 64 * kSlotChunks) {
    bool in[kSlotChunks];
#pragma unroll
    for (int u = 0; u < kSlotChunks; ++u) {
      const int i = c0 + 64 * u + lane;
      in[u] = i < s1 && g.inside(d, cx[i], cy[i]);
    }
#pragma unroll
    for (int u = 0; u < kSlotChunks; ++u) {
      const unsigned long long m = __ballot(in[u]);
      if (WRITE && in[u]) slot_pt[base + c + __popcll(m & below)] = center_base + c0 + 64 * u + lane;
      c += __popcll(m);
    }
  }
  return c;
}

// segment w (of 4) of [first, last), 64-aligned
__device__ __forceinline__ void segment_of(int first, int last, int w, int& s0, int& s1) {
  const int len = ((last - first + 255) / 256) * 64;
  s0 = min(last, first + w * len);
  s1 = min(last, s0 + len);
}

// members of (global) subdomain blockIdx.x: count
__global__ __launch_bounds__(256) void k_tps_count(const piv_tps_field* __restrict__ fields,
                                                   const int32_t* __restrict__ dom_field,
                                                   const double* __restrict__ lim,
                                                   const double* __restrict__ cx_all,
                                                   const double* __restrict__ cy_all,
                                                   int32_t* counts) {
  const int fi = dom_field[blockIdx.x];
  if (fi < 0) {  // a subdomain slot beyond its field's grid (device-sized batches)
    if (threadIdx.x == 0) counts[blockIdx.x] = 0;
    return;
  }
  const piv_tps_field f = fields[fi];
  const SubGrid g = field_grid(f, lim);
  const int d = blockIdx.x - f.dom_base;
  const double* cx = cx_all + f.center_base;
  const double* cy = cy_all + f.center_base;
  int first, last, s0, s1;
  wave_member_range(f, g, d, cy, first, last);
  segment_of(first, last, threadIdx.x >> 6, s0, s1);
  const int c = segment_members<false>(g, d, cx, cy, s0, s1, 0, nullptr, 0);
  __shared__ int part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// members of (global) subdomain blockIdx.x in increasing center order (np.where)
__global__ __launch_bounds__(256) void k_tps_fill(const piv_tps_field* __restrict__ fields,
                                                  const int32_t* __restrict__ dom_field,
                                                  const double* __restrict__ lim,
                                                  const double* __restrict__ cx_all,
                                                  const double* __restrict__ cy_all,
                                                  const int64_t* __restrict__ dom_off,
                                                  int32_t* slot_pt) {
  const int fi = dom_field[blockIdx.x];
  // inactive subdomain, or one of a field the offsets left out (no slots)
  if (fi < 0 || dom_off[blockIdx.x + 1] == dom_off[blockIdx.x]) return;
  const piv_tps_field f = fields[fi];
  const SubGrid g = field_grid(f, lim);
  const int d = blockIdx.x - f.dom_base;
  const int center_base = f.center_base;
  const double* cx = cx_all + center_base;
  const double* cy = cy_all + center_base;
  const int w = threadIdx.x >> 6;
  int first, last, s0, s1;
  wave_member_range(f, g, d, cy, first, last);
  segment_of(first, last, w, s0, s1);
  // each segment's count, then its members from the segments before it on
  __shared__ int part[4];
  const int c = segment_members<false>(g, d, cx, cy, s0, s1, 0, nullptr, 0);
  if ((threadIdx.x & 63) == 0) part[w] = c;
  __syncthreads();
  int64_t base = dom_off[blockIdx.x];
  for (int k = 0; k < w; ++k) base += part[k];
  segment_members<true>(g, d, cx, cy, s0, s1, center_base, slot_pt, base);
}

// per point: its subdomains in increasing order (at most kMaxSlots), their
// weights and the sum of the weights. Centers mode (dom_off != NULL): the
// slot is the position of the point in the member list of the subdomain.
// field blockIdx.y; centers mode (dom_off != NULL): its centers
// cx[center_base + p], p < count, outputs at center_base + p, and for
// count <= p < m an empty entry (no slot, weight 0, norm 1); new-position
// mode: the m positions px[p], outputs at field * m + p
__global__ void k_tps_point_slots(const piv_tps_field* __restrict__ fields,
                                  const double* __restrict__ lim, int m,
                                  const double* __restrict__ px_all,
                                  const double* __restrict__ py_all,
                                  const int64_t* __restrict__ dom_off,
                                  const int32_t* __restrict__ slot_pt, int32_t* out_slot,
                                  double* out_coef, double* out_norm, int32_t* status) {
  const piv_tps_field f = fields[blockIdx.y];
  const SubGrid g = field_grid(f, lim);
  const bool centers = dom_off != nullptr;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m) return;
  const int64_t in = centers ? int64_t(f.center_base) + p : p;
  const int64_t o = centers ? int64_t(f.center_base) + p : int64_t(blockIdx.y) * m + p;
  const int center_base = f.center_base, dom_base = f.dom_base;
  out_slot += kMaxSlots * o;
  out_coef += kMaxSlots * o;
  out_norm += o;
  // beyond the field's valid vectors, or a field left out (no grid): empty
  if ((centers && p >= f.count) || g.nsx == 0) {
    for (int k = 0; k < kMaxSlots; ++k) {
      out_slot[k] = -1;
      out_coef[k] = 0.0;
    }
    out_norm[0] = 1.0;
    return;
  }
  const double x = px_all[in], y = py_all[in];
  int xs_[kMaxSlots], ys_[kMaxSlots];
  int nx = 0, ny = 0;
  for (int i = 0; i < g.nsx; ++i)
    if (x >= g.xlo(i) && x < g.xhi(i)) {
      if (nx < kMaxSlots) xs_[nx] = i;
      ++nx;
    }
  for (int i = 0; i < g.nsy; ++i)
    if (y >= g.ylo(i) && y < g.yhi(i)) {
      if (ny < kMaxSlots) ys_[ny] = i;
      ++ny;
    }
  int k = 0;
  double norm = 0.0;
  if (nx * ny > kMaxSlots) status[0] = 1;
  for (int a = 0; a < min(ny, kMaxSlots); ++a)
    for (int b = 0; b < min(nx, kMaxSlots); ++b) {
      if (k >= kMaxSlots) continue;
      const int d = ys_[a] * g.nsx + xs_[b];
      const double cf = g.coef(d, x, y);
      int slot = dom_base + d;
      if (centers) {
        const int key = center_base + p;
        int64_t lo = dom_off[dom_base + d], hi = dom_off[dom_base + d + 1];
        while (lo < hi) {
          const int64_t mid = (lo + hi) / 2;
          if (slot_pt[mid] < key) lo = mid + 1;
          else hi = mid;
        }
        slot = int(lo);
      }
      out_slot[k] = slot;
      out_coef[k] = cf;
      norm += cf;
      ++k;
    }
  for (; k < kMaxSlots; ++k) {
    out_slot[k] = -1;
    out_coef[k] = 0.0;
  }
  out_norm[0] = norm;
}

// ---------------------------------------------------------------------------
// 4b. subdomain grids and offsets on the device: a batch sized by
//     capacities instead of host reads (thin_plate_spline_subdom.py:53-126)
// ---------------------------------------------------------------------------
// One workgroup per field: the subdomain counts of the reference from the
// count and bounds of its selected centers (info[f][7] of k_tps_select), the
// np.linspace limits (y[i] = i step + start, the last one = stop, rounded
// as numpy rounds them), the 20 % buffers and the rescaled smoothing; field
// records, limits, per-subdomain smoothing and subdomain-to-field map of
// dom_cap subdomain slots per field (slots beyond the field's grid: -1).
__global__ __launch_bounds__(256) void k_tps_grid(
    const double* __restrict__ info, int n, double subdom_size, double smoothing_coef,
    double coef_buffer, int dom_cap, double gen_width, int require_lattice,
    piv_tps_field* __restrict__ fields, double* __restrict__ lim,
    int32_t* __restrict__ dom_field, double* __restrict__ smoothing_dom,
    int32_t* __restrict__ fstat) {
#pragma clang fp contract(off)
  const int f = blockIdx.x;
  const double* in = info + 7 * f;
  const int count = int(in[0]);
  const double x_min = in[1], x_max = in[2], y_min = in[3], y_max = in[4];
  __shared__ int sh_nsx, sh_nsy, sh_st;
  __shared__ double sh_smooth;
  if (threadIdx.x == 0) {
    int st = PIV_TPS_FIELD_OK, nsx = 0, nsy = 0;
    double smooth = 0.0;
    const double range_x = x_max - x_min, range_y = y_max - y_min;
    if (count == 0) {
      st = PIV_TPS_FIELD_NO_VECTOR;
    } else if (range_x == 0.0 || range_y == 0.0) {
      st = PIV_TPS_FIELD_DEGENERATE;
    } else {
      const double aspect = range_y / range_x;
      const double nb = double(count) / subdom_size;
      const double fx = floor(sqrt(nb / aspect));
      const double fy = ceil(sqrt(nb * aspect));
      // int(...) or 1; a grid beyond the capacity is left to the host
      if (!(fx <= double(dom_cap)) || !(fy <= double(dom_cap))) {
        st = PIV_TPS_FIELD_CAPACITY;
      } else {
        nsx = fx >= 1.0 ? int(fx) : 1;
        nsy = fy >= 1.0 ? int(fy) : 1;
        if (int64_t(nsx) * nsy > dom_cap) {
          st = PIV_TPS_FIELD_CAPACITY;
        } else {
          const double step_x = range_x / double(nsx), step_y = range_y / double(nsy);
          // x_dom[1] - x_dom[0]: x_dom[0] = 0 step + x_min = x_min
          const double x1 = nsx > 1 ? 1.0 * step_x + x_min : x_max;
          const double y1 = nsy > 1 ? 1.0 * step_y + y_min : y_max;
          const double per = double(count) / double(nsx * nsy);
          smooth = smoothing_coef * sqrt((x1 - x_min) * (y1 - y_min) / per);
        }
      }
    }
    if (st == PIV_TPS_FIELD_OK && require_lattice && in[6] == 0.0) st = PIV_TPS_FIELD_OFF_LATTICE;
    if (st != PIV_TPS_FIELD_OK) nsx = nsy = 0;
    sh_nsx = nsx;
    sh_nsy = nsy;
    sh_st = st;
    sh_smooth = smooth;
  }
  __syncthreads();
  const int nsx = sh_nsx, nsy = sh_nsy;
  int st = sh_st;
  const int64_t lim_off = int64_t(f) * 4 * dom_cap;
  if (st == PIV_TPS_FIELD_OK) {
    const double range_x = x_max - x_min, range_y = y_max - y_min;
    const double step_x = range_x / double(nsx), step_y = range_y / double(nsy);
    const double bx = coef_buffer * range_x / double(nsx);
    const double by = coef_buffer * range_y / double(nsy);
    double* L = lim + lim_off;
    int wide = 0;
    for (int i = threadIdx.x; i < nsx; i += blockDim.x) {
      const double a = double(i) * step_x + x_min;
      const double b = i + 1 < nsx ? double(i + 1) * step_x + x_min : x_max;
      L[i] = a - bx;
      L[nsx + i] = b + bx;
      if (gen_width > 0.0 && !(L[nsx + i] - L[i] <= gen_width)) wide = 1;
    }
    for (int i = threadIdx.x; i < nsy; i += blockDim.x) {
      const double a = double(i) * step_y + y_min;
      const double b = i + 1 < nsy ? double(i + 1) * step_y + y_min : y_max;
      L[2 * nsx + i] = a - by;
      L[2 * nsx + nsy + i] = b + by;
      if (gen_width > 0.0 && !(L[2 * nsx + nsy + i] - L[2 * nsx + i] <= gen_width)) wide = 1;
    }
    // a subdomain wider than the kernel-function table (generated matrices)
    if (__syncthreads_or(wide)) st = PIV_TPS_FIELD_OFF_LATTICE;
  }
  const bool ok = st == PIV_TPS_FIELD_OK;
  const int nd = ok ? nsx * nsy : 0;
  for (int d = threadIdx.x; d < dom_cap; d += blockDim.x) {
    dom_field[int64_t(f) * dom_cap + d] = d < nd ? f : -1;
    smoothing_dom[int64_t(f) * dom_cap + d] = sh_smooth;
  }
  if (threadIdx.x == 0) {
    piv_tps_field r;
    r.nsx = ok ? nsx : 0;
    r.nsy = ok ? nsy : 0;
    r.count = ok ? count : 0;
    r.center_base = f * n;
    r.dom_base = f * dom_cap;
    r.lim_off = int(lim_off);
    r.sorted_y = in[5] != 0.0 ? 1 : 0;
    fields[f] = r;
    fstat[f] = st;
  }
}

// inclusive prefix sum of int64 over a 1024-thread block
__device__ int64_t block_scan_1024_64(int64_t v, int64_t* tmp, int64_t& total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  if (lane == 63) tmp[wv] = v;
  __syncthreads();
  if (threadIdx.x < 16) {
    int64_t x = tmp[threadIdx.x];
    for (int off = 1; off < 16; off <<= 1) {
      const int64_t u = __shfl_up(x, off, 16);
      if (threadIdx.x >= off) x += u;
    }
    tmp[16 + threadIdx.x] = x;
  }
  __syncthreads();
  total = tmp[16 + 15];
  const int64_t add = wv > 0 ? tmp[16 + wv - 1] : 0;
  __syncthreads();
  return v + add;
}

constexpr int kOffsetsMaxFields = 2048;

// One workgroup of 1024 threads for the whole batch: checks the member
// counts of every subdomain (an empty subdomain or too many centers fails
// its field, as the host sizing does), admits the fields in order while
// their slots and matrices fit the capacities, then writes the prefix sums
// dom_off / mat_off over the nd = nfields dom_cap subdomain slots (the
// slots of failed fields get no room) and the order that takes the largest
// subdomains first. totals: slots, matrix doubles, largest subdomain,
// subdomains with centers.
__global__ __launch_bounds__(1024) void k_tps_offsets(
    int nfields, int dom_cap, piv_tps_field* __restrict__ fields,
    const int32_t* __restrict__ counts, const int32_t* __restrict__ dom_field,
    int64_t slots_cap, int64_t mats_cap, int n_max_cap, int32_t* __restrict__ fstat,
    int64_t* __restrict__ dom_off, int64_t* __restrict__ mat_off, int32_t* __restrict__ order,
    int64_t* __restrict__ totals) {
  __shared__ int flags[kOffsetsMaxFields];
  __shared__ int hist[kMaxN + 2];
  __shared__ int64_t tmp64[32];
  __shared__ int tmpi[32];
  __shared__ int64_t fslots[kOffsetsMaxFields];
  __shared__ int64_t fmats[kOffsetsMaxFields];
  const int nd = nfields * dom_cap;
  const int tid = threadIdx.x;
  for (int f = tid; f < nfields; f += 1024) {
    flags[f] = 0;
    fslots[f] = 0;
    fmats[f] = 0;
  }
  for (int b = tid; b < kMaxN + 2; b += 1024) hist[b] = 0;
  __syncthreads();
  // 1. per-subdomain checks: bit 0 empty, bit 1 more than kMaxN, bit 2 above n_max_cap
  //    and each field's slots and matrix doubles (integer sums: any order)
  for (int d = tid; d < nd; d += 1024) {
    const int f = dom_field[d];
    if (f < 0 || fstat[f] != PIV_TPS_FIELD_OK) continue;
    const int c = counts[d];
    int bits = 0;
    if (c == 0) bits = 1;
    else if (c > kMaxN) bits = 2;
    else if (c > n_max_cap) bits = 4;
    if (bits) atomicOr(&flags[f], bits);
    atomicAdd(reinterpret_cast<unsigned long long*>(&fslots[f]), (unsigned long long)c);
    atomicAdd(reinterpret_cast<unsigned long long*>(&fmats[f]),
              (unsigned long long)tps_mat_size(c));
  }
  __syncthreads();
  // 2. statuses (an empty subdomain first, as the host sizing checks it
  //    first)
  for (int f = tid; f < nfields; f += 1024) {
    int st = fstat[f];
    const int b = flags[f];
    if (st == PIV_TPS_FIELD_OK && b)
      st = (b & 1) ? PIV_TPS_FIELD_EMPTY_SUBDOMAIN
                   : ((b & 2) ? PIV_TPS_FIELD_TOO_MANY : PIV_TPS_FIELD_CAPACITY);
    flags[f] = st;
  }
  __syncthreads();
  // 3. admission in field order while the capacities hold
  if (tid == 0) {
    int64_t cs = 0, cm = 0;
    for (int f = 0; f < nfields; ++f) {
      if (flags[f] != PIV_TPS_FIELD_OK) continue;
      const int64_t ma = fmats[f];
      if (cs + fslots[f] > slots_cap || cm + ma > mats_cap) {
        flags[f] = PIV_TPS_FIELD_CAPACITY;
        continue;
      }
      cs += fslots[f];
      cm += ma;
    }
  }
  __syncthreads();
  for (int f = tid; f < nfields; f += 1024) {
    fstat[f] = flags[f];
    if (flags[f] != PIV_TPS_FIELD_OK) {  // no grid from here on: points get no slots
      fields[f].nsx = 0;
      fields[f].nsy = 0;
      fields[f].count = 0;
    }
  }
  // 4. prefix sums of the admitted subdomains' slots and matrix sizes; the
  //    histogram of their sizes
  //    (kPer consecutive subdomains a thread, so one round covers 8192)
  int64_t base_s = 0, base_m = 0;
  int n_max = 0, active = 0;
  constexpr int kPer = 8;
  for (int c0 = 0; c0 < nd; c0 += 1024 * kPer) {
    int cs[kPer];
    int64_t ls = 0, lm = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int d = c0 + kPer * tid + k;
      int c = 0;
      if (d < nd) {
        const int f = dom_field[d];
        if (f >= 0 && flags[f] == PIV_TPS_FIELD_OK) c = counts[d];
      }
      cs[k] = c;
      ls += c;
      lm += c > 0 ? tps_mat_size(c) : 0;
    }
    int64_t ts, tm;
    int64_t os = base_s + block_scan_1024_64(ls, tmp64, ts) - ls;
    int64_t om = base_m + block_scan_1024_64(lm, tmp64, tm) - lm;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int d = c0 + kPer * tid + k;
      const int c = cs[k];
      if (d < nd) {
        dom_off[d] = os;
        mat_off[d] = om;
        atomicAdd(&hist[c], 1);
      }
      os += c;
      om += c > 0 ? tps_mat_size(c) : 0;
      n_max = max(n_max, c);
      active += c > 0 ? 1 : 0;
    }
    base_s += ts;
    base_m += tm;
  }
  if (tid == 0) {
    dom_off[nd] = base_s;
    mat_off[nd] = base_m;
  }
  // reductions of n_max and active
  for (int off = 32; off > 0; off >>= 1) {
    n_max = max(n_max, __shfl_xor(n_max, off, 64));
    active += __shfl_xor(active, off, 64);
  }
  __syncthreads();
  if ((tid & 63) == 0) {
    tmp64[tid >> 6] = n_max;
    tmp64[16 + (tid >> 6)] = active;
  }
  __syncthreads();
  // 5. order: sizes descending (a counting sort; the order within a size
  //    is free, results do not depend on it)
  if (tid == 0) {
    int64_t nm = 0, na = 0;
    for (int w = 0; w < 16; ++w) {
      nm = nm > tmp64[w] ? nm : tmp64[w];
      na += tmp64[16 + w];
    }
    totals[0] = base_s;
    totals[1] = base_m;
    totals[2] = nm;
    totals[3] = na;
  }
  {
    // hist[b] <- the subdomains of sizes above b (an exclusive scan of the
    // bins from the largest size down, kPer bins a thread)
    constexpr int kBins = kMaxN + 2, kPer = (kBins + 1023) / 1024;
    int h[kPer];
    int local = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int r = kPer * tid + k;
      h[k] = r < kBins ? hist[kBins - 1 - r] : 0;
      local += h[k];
    }
    int total;
    int run = block_scan_1024(local, tmpi, total) - local;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int r = kPer * tid + k;
      if (r < kBins) hist[kBins - 1 - r] = run;
      run += h[k];
    }
  }
  __syncthreads();
  for (int d = tid; d < nd; d += 1024) {
    const int f = dom_field[d];
    const int c = (f >= 0 && flags[f] == PIV_TPS_FIELD_OK) ? counts[d] : 0;
    order[atomicAdd(&hist[c], 1)] = d;
  }
}

// The evaluation at new positions that are the field's own centres: the
// blended smoothed value there ([K P][w; a] = v - s w at the centres). For
// field f (blockIdx.y) and the rank o < count of a selected vector (vector
// q = idx[f n + o]): if new position q is that vector's position, out gets
// the smoothed value and skip[f n + q] = 1 (out and skip start at zero).
__global__ void k_tps_at_centers(int n, const double* __restrict__ info,
                                 const int32_t* __restrict__ idx, const double* __restrict__ xs,
                                 const double* __restrict__ ys, const double* __restrict__ new_x,
                                 const double* __restrict__ new_y,
                                 const double* __restrict__ smoothed, int64_t ntot,
                                 double* __restrict__ out, uint8_t* __restrict__ skip) {
  const int f = blockIdx.y;
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  const int count = int(info[7 * f]);
  if (o >= count || o >= n) return;
  const int64_t base = int64_t(f) * n;
  const int q = idx[base + o];
  if (q < 0 || q >= n) return;
  if (new_x[q] == xs[base + q] && new_y[q] == ys[base + q]) {
    out[base + q] = smoothed[base + o];
    out[ntot + base + q] = smoothed[ntot + base + o];
    skip[base + q] = 1;
  }
}

// ---------------------------------------------------------------------------
// 4c. subdomains sharing one factorisation (centers on a grid)
// ---------------------------------------------------------------------------
// When every center of a field is a node of the pass grid gx[ngx] x gy[ngy]
// (the first pass; passes on deformed images), the members S of a
// subdomain are the valid nodes of its box F (the grid nodes inside its
// half-open limits). Boxes with the same numbers of columns and rows (and
// the same spacings: the grid is regular but for a last column or row
// appended at a smaller step) hold the same node set up to a translation;
// with the polynomial part P = [1, y - y0, x - x0] relative to the box's
// first node their saddle-point matrices M_F are identical, so one
// factorisation per field and pattern serves every such box. A subdomain
// missing few nodes R = F \ S (the invalid vectors) is solved with it and a
// capacitance correction: with v' = v on S, 0 on R,
//   M_F [w; a] = [v'; 0] + E_R lambda,  w_R = 0,
// whose rows of S are M_S [w_S; a] = [v_S; 0]. For G = M_F^-1 E_R and
// H = G_w[R, :] (|R| x |R|, positive definite when M_S is regular),
// lambda = -H^-1 y_w[R] with y = M_F^-1 [v'; 0], and (w, a) = y + G lambda.
// The outlier iterations of the reference change v only, so G and H are
// computed once per subdomain.
// Rank of workgroup b among nb when the 8 XCDs of the chip (each with its
// own L2) take workgroups round-robin (b to XCD b % 8): XCD x gets the
// contiguous ranks [x q + min(x, r), ...), q = nb / 8, r = nb % 8. A
// bijection of [0, nb).
__device__ __forceinline__ int xcd_rank(int b, int nb) {
  constexpr int kXcd = 8;
  const int x = b % kXcd, j = b / kXcd;
  const int q = nb / kXcd, r = nb % kXcd;
  return x * q + min(x, r) + j;
}

constexpr int kCapMaxR = 12;  // missing nodes of a subdomain solved with a shared factor
constexpr int kCapX = 32;     // ... with a pattern's explicit inverse
static_assert(kCapX <= 32, "k_tps_shared_x inverts H with two lanes per row");
constexpr int kPatMaxPerField = 64;

// first index in [0, n) with a[i] >= v (a ascending), n if none
__device__ __forceinline__ int lower_bound_d(const double* a, int n, double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One workgroup per field: the box of every subdomain on the grid, and the
// field's patterns. dmode[d]: 1 solved with pattern dpat[d] (global index
// f npat_cap + id), 0 direct factorisation, -1 no centers. dbox[4 d]:
// first column, columns, first row, rows. pat_n[p]: nodes of pattern p (0:
// unused), pat_box[4 p] its reference box, pat_smooth[p] its smoothing.
__global__ __launch_bounds__(256) void k_tps_patterns(
    int dom_cap, const piv_tps_field* __restrict__ fields, const double* __restrict__ lim,
    const int32_t* __restrict__ fstat, const int32_t* __restrict__ dom_field,
    const int32_t* __restrict__ counts, const double* __restrict__ smoothing,
    const double* __restrict__ gx, int ngx, int irr_x, const double* __restrict__ gy, int ngy,
    int irr_y, int cap_r, int n_max_cap, int npat_cap, int32_t* __restrict__ dmode,
    int32_t* __restrict__ dpat, int32_t* __restrict__ dbox, int32_t* __restrict__ pat_n,
    int32_t* __restrict__ pat_box, double* __restrict__ pat_smooth) {
  const int f = blockIdx.x;
  const int64_t d0 = int64_t(f) * dom_cap;
  const bool ok = fstat[f] == PIV_TPS_FIELD_OK;
  const piv_tps_field fr = fields[f];
  const SubGrid g = field_grid(fr, lim);
  for (int d = threadIdx.x; d < dom_cap; d += blockDim.x) {
    const int64_t dd = d0 + d;
    dmode[dd] = -1;
    dpat[dd] = -1;
    if (!ok || dom_field[dd] != f || counts[dd] == 0) continue;
    const int ix = d % g.nsx, iy = d / g.nsx;
    const int j0 = lower_bound_d(gx, ngx, g.xlo(ix)), j1 = lower_bound_d(gx, ngx, g.xhi(ix));
    const int i0 = lower_bound_d(gy, ngy, g.ylo(iy)), i1 = lower_bound_d(gy, ngy, g.yhi(iy));
    dbox[4 * dd + 0] = j0;
    dbox[4 * dd + 1] = j1 - j0;
    dbox[4 * dd + 2] = i0;
    dbox[4 * dd + 3] = i1 - i0;
    dmode[dd] = 0;
  }
  for (int q = threadIdx.x; q < npat_cap; q += blockDim.x) {
    pat_n[int64_t(f) * npat_cap + q] = 0;
    pat_smooth[int64_t(f) * npat_cap + q] = smoothing[d0];
  }
  __syncthreads();
  if (!ok) return;
  // pattern keys (columns, rows, and whether the box ends on an irregular
  // last column / row of the grid) of the subdomains that may share, kept as
  // dpat = -(key + 2) until they are assigned
  for (int d = threadIdx.x; d < dom_cap; d += blockDim.x) {
    const int64_t dd = d0 + d;
    if (dmode[dd] != 0) continue;
    const int j0 = dbox[4 * dd], nc = dbox[4 * dd + 1], i0 = dbox[4 * dd + 2],
              nr = dbox[4 * dd + 3];
    const int nf = nc * nr;
    if (nf < 4 || nf > n_max_cap || nf - counts[dd] > cap_r) continue;  // direct
    const int key = ((nc * 4096 + nr) * 2 + ((j0 + nc == ngx && irr_x) ? 1 : 0)) * 2 +
                    ((i0 + nr == ngy && irr_y) ? 1 : 0);
    dpat[dd] = -(key + 2);
  }
  // pattern r: the key of the first unassigned subdomain in index order,
  // assigned to every subdomain of that key (patterns numbered in order of
  // first occurrence); beyond the cap the rest are factorised alone
  __shared__ int first_sh[4];
  const int cap = min(npat_cap, kPatMaxPerField);
  for (int r = 0; r < cap; ++r) {
    __syncthreads();
    int dmin = INT_MAX;
    for (int d = threadIdx.x; d < dom_cap; d += blockDim.x)
      if (dpat[d0 + d] <= -2) {
        dmin = d;
        break;
      }
    for (int off = 32; off > 0; off >>= 1) dmin = min(dmin, __shfl_xor(dmin, off, 64));
    if ((threadIdx.x & 63) == 0) first_sh[threadIdx.x >> 6] = dmin;
    __syncthreads();
    dmin = min(min(first_sh[0], first_sh[1]), min(first_sh[2], first_sh[3]));
    if (dmin == INT_MAX) break;
    const int64_t dm = d0 + dmin;
    const int code = dpat[dm];
    const int64_t p = int64_t(f) * npat_cap + r;
    if (threadIdx.x == 0) {
      pat_n[p] = dbox[4 * dm + 1] * dbox[4 * dm + 3];
#pragma unroll
      for (int k = 0; k < 4; ++k) pat_box[4 * p + k] = dbox[4 * dm + k];
    }
    __syncthreads();  // every thread read the key before it is replaced
    for (int d = dmin + threadIdx.x; d < dom_cap; d += blockDim.x) {
      const int64_t dd = d0 + d;
      if (dpat[dd] == code) {
        dmode[dd] = 1;
        dpat[dd] = int(p);
      }
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < dom_cap; d += blockDim.x)
    if (dpat[d0 + d] <= -2) dpat[d0 + d] = -1;
}

// inclusive max-scan over a block of 1024 threads (tmp: 32 ints of LDS)
__device__ int block_scan_max_1024(int v, int* tmp) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v = max(v, u);
  }
  if (lane == 63) tmp[wv] = v;
  __syncthreads();
  if (threadIdx.x < 16) {
    int x = tmp[threadIdx.x];
    for (int off = 1; off < 16; off <<= 1) {
      const int u = __shfl_up(x, off, 16);
      if (threadIdx.x >= off) x = max(x, u);
    }
    tmp[16 + threadIdx.x] = x;
  }
  __syncthreads();
  const int before = wv > 0 ? tmp[16 + wv - 1] : INT_MIN;
  __syncthreads();
  return max(v, before);
}

// One workgroup of 1024 threads: offsets of the patterns' centers, matrices
// and explicit inverses (prefix sums of their sizes), the orders of the
// direct subdomains (from `order`, largest first) and of the shared ones (by
// index; -1 beyond), and the patterns in use (plist[0 .. plist[npat]), -1
// beyond).
__global__ __launch_bounds__(1024) void k_tps_pattern_layout(
    int npat, const int32_t* __restrict__ pat_n, int nd,
    const int32_t* __restrict__ order, const int32_t* __restrict__ dmode,
    int64_t* __restrict__ pdom_off, int64_t* __restrict__ pmat_off, int64_t* __restrict__ pxo,
    int32_t* __restrict__ order_direct, int32_t* __restrict__ order_shared,
    int32_t* __restrict__ plist) {
  __shared__ int64_t tmp64[32];
  __shared__ int tmp[32];
  const int tid = threadIdx.x;
  int64_t base_s = 0, base_m = 0, base_x = 0;
  int base_u = 0;
  for (int c0 = 0; c0 < npat; c0 += 1024) {
    const int p = c0 + tid;
    const int c = p < npat ? pat_n[p] : 0;
    int tu;
    const int iu = block_scan_1024(c > 0 ? 1 : 0, tmp, tu);
    if (c > 0) plist[base_u + iu - 1] = p;
    base_u += tu;
    int64_t ts, tm, tx;
    const int64_t is = block_scan_1024_64(int64_t(c), tmp64, ts);
    const int64_t im = block_scan_1024_64(c > 0 ? tps_mat_size(c) : 0, tmp64, tm);
    const int64_t ix = block_scan_1024_64(tps_x_size(c), tmp64, tx);
    if (p < npat) {
      pdom_off[p] = base_s + is - c;
      pmat_off[p] = base_m + im - (c > 0 ? tps_mat_size(c) : 0);
      pxo[p] = base_x + ix - tps_x_size(c);
    }
    base_s += ts;
    base_m += tm;
    base_x += tx;
  }
  if (tid == 0) {
    pdom_off[npat] = base_s;
    pmat_off[npat] = base_m;
    pxo[npat] = base_x;
    plist[npat] = base_u;
  }
  for (int p = base_u + tid; p < npat; p += 1024) plist[p] = -1;
  __syncthreads();
  // the direct subdomains in the order of `order`; the shared ones in index
  // order: a field's subdomains are consecutive and most of them share the
  // field's interior pattern, so the workgroups running at one time read
  // few factors (which stay in the L2 cache)
  int bd = 0, bs = 0;
  for (int c0 = 0; c0 < nd; c0 += 1024) {
    const int b = c0 + tid;
    const int d = b < nd ? order[b] : -1;
    const int md = d >= 0 ? dmode[d] : -1;
    const int ms = b < nd ? dmode[b] : -1;
    int td, tsh;
    const int idd = block_scan_1024(md == 0 ? 1 : 0, tmp, td);
    const int ish = block_scan_1024(ms == 1 ? 1 : 0, tmp, tsh);
    if (md == 0) order_direct[bd + idd - 1] = d;
    if (ms == 1) order_shared[bs + ish - 1] = b;
    bd += td;
    bs += tsh;
  }
  for (int b = bd + tid; b < nd; b += 1024) order_direct[b] = -1;
  for (int b = bs + tid; b < nd; b += 1024) order_shared[b] = -1;
  if (tid == 0) order_shared[nd] = bs;
}

// The centers of the patterns in use (workgroups striding over plist): grid
// nodes relative to the box's first node, row-major (the order of the
// members of a subdomain), and an identity slot list.
__global__ __launch_bounds__(256) void k_tps_pattern_nodes(
    int npat, const int32_t* __restrict__ plist, const int32_t* __restrict__ pat_n,
    const int32_t* __restrict__ pat_box, const int64_t* __restrict__ pdom_off,
    const double* __restrict__ gx, const double* __restrict__ gy, double* __restrict__ pxy,
    int32_t* __restrict__ pslot, int64_t pxy_stride) {
  const int npu = plist[npat];
  for (int u = blockIdx.x; u < npu; u += gridDim.x) {
    const int p = plist[u];
    const int n = pat_n[p];
    const int j0 = pat_box[4 * p], nc = pat_box[4 * p + 1], i0 = pat_box[4 * p + 2];
    const int64_t o = pdom_off[p];
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      const int r = k / nc, c = k % nc;
      pxy[o + k] = gx[j0 + c] - gx[j0];
      pxy[pxy_stride + o + k] = gy[i0 + r] - gy[i0];
      pslot[o + k] = int(o + k);
    }
  }
}

// inclusive scan over a block of 256 threads (tmp: 8 ints of LDS)
__device__ int block_scan_256(int v, int* tmp, int& total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  if (lane == 63) tmp[wv] = v;
  __syncthreads();
  int before = 0;
  for (int k = 0; k < wv; ++k) before += tmp[k];
  total = tmp[0] + tmp[1] + tmp[2] + tmp[3];
  __syncthreads();
  return v + before;
}

// Field blockIdx.x (256 threads): its shared subdomains (a segment of the
// shared order, which is by index) reordered by pattern, stably, into
// order_sorted: the subdomains of one pattern follow one another, so that
// the groups of k_tps_shared_x are full and the workgroups running at one
// time on an XCD read few patterns' inverses (from its L2 cache).
__global__ __launch_bounds__(256) void k_tps_shared_sort(
    int dom_cap, int npat_cap, const int32_t* __restrict__ pat_n,
    const int32_t* __restrict__ dpat, const int32_t* __restrict__ order_shared, int nd,
    int32_t* __restrict__ order_sorted) {
  __shared__ int tmp[8];
  const int f = blockIdx.x, tid = threadIdx.x;
  const int bs = order_shared[nd];
  // the field's segment [lo, hi) of the shared order (ascending subdomains)
  auto first_at_least = [&](int d) {
    int lo = 0, hi = bs;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (order_shared[mid] < d) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const int lo = first_at_least(f * dom_cap), hi = first_at_least((f + 1) * dom_cap);
  int pos = lo;
  for (int q = 0; q < npat_cap; ++q) {
    const int p = f * npat_cap + q;
    if (pat_n[p] == 0) continue;
    for (int c0 = lo; c0 < hi; c0 += 256) {
      const int b = c0 + tid;
      const int d = b < hi ? order_shared[b] : -1;
      const bool in = d >= 0 && dpat[d] == p;
      int t;
      const int r = block_scan_256(in ? 1 : 0, tmp, t);
      if (in) order_sorted[pos + r - 1] = d;
      pos += t;
    }
  }
  if (f == 0)
    for (int b = bs + tid; b <= nd; b += 256) order_sorted[b] = b == nd ? bs : -1;
}

// One workgroup of 1024 threads: the groups of k_tps_shared_x, runs of one
// pattern in the sorted shared order cut every `group` subdomains
// (groups[g] the first entry of group g, groups[nd + 1 + g] its size,
// groups[nd] their number).
__global__ __launch_bounds__(1024) void k_tps_shared_groups(
    int nd, const int32_t* __restrict__ dpat, const int32_t* __restrict__ order, int group,
    int32_t* __restrict__ groups) {
  __shared__ int tmp[32];
  __shared__ int last_run;
  const int tid = threadIdx.x;
  const int bs = order[nd];
  int32_t* gcnt = groups + nd + 1;
  for (int b = tid; b < nd; b += 1024) gcnt[b] = 0;
  __syncthreads();  // the zeroed sizes, for the whole block
  int run = -1, ng = 0;
  for (int c0 = 0; c0 < bs; c0 += 1024) {
    const int b = c0 + tid;
    const bool valid = b < bs;
    const int pb = valid ? dpat[order[b]] : -1;
    const int pp = valid && b > 0 ? dpat[order[b - 1]] : -1;
    const bool first = valid && (b == 0 || pb != pp);
    const int rs = max(run, block_scan_max_1024(first ? b : -1, tmp));
    const bool gs = valid && (b - rs) % group == 0;
    int tg;
    const int ig = block_scan_1024(gs ? 1 : 0, tmp, tg);
    const int gi = ng + ig - 1;
    if (gs) groups[gi] = b;
    if (valid) atomicAdd(&gcnt[gi], 1);
    if (tid == 1023) last_run = rs;
    __syncthreads();
    run = last_run;
    ng += tg;
    __syncthreads();
  }
  if (tid == 0) groups[nd] = ng;
}

// Solves of the subdomains that share their pattern's factorisation, with
// the capacitance correction of their missing nodes and the reference's
// outlier iterations (thin_plate_spline_subdom.py:236-280, as k_tps_solve),
// in two kernels:
//
//   k_tps_shared_columns: checks every shared subdomain (pattern factor
//     regular, centers on the grid, at most cap_r missing nodes) and, when
//     nodes are missing, computes the columns G = M_F^-1 E_R (up to
//     kSharedNV right-hand sides per pass over the factor), their
//     polynomial parts and the Cholesky factor of H = G_w[R, :];
//   k_tps_shared_iter: the outlier iterations, two right-hand sides per pass
//     over the factor, in a fifth of the LDS of the first kernel so that
//     five workgroups share a CU (the passes are bound by the latency of
//     the triangular solves' chain of block rows, so more of them in flight
//     is what speeds them up); it writes what k_tps_solve and k_tps_prep
//     write for a subdomain (smoothed values and weights by slot, the
//     polynomial part in image coordinates, stats, slot_xy).
//
// Both map workgroup b to the shared order by xcd_rank (order[gridDim.x]:
// the number of shared subdomains).
constexpr int kSharedNV = 8;

// LDS scratch of chol_solve_n<NV> (doubles): the diagonal block, the block
// right-hand sides and the partial sums of NV vectors
constexpr int solve_scratch(int nv) { return kNB * (kNB + 1) + 8 * kNB + nv * 8 * kNB; }

// dynamic LDS bytes of k_tps_shared_columns<nv> and k_tps_shared_iter
constexpr size_t shared_columns_lds(int nv, int nm) {
  return (size_t(nv + 1) * nm + solve_scratch(nv)) * sizeof(double);
}
constexpr size_t shared_iter_lds(int nv, int nm) {
  return (size_t(nv + 2) * nm + solve_scratch(nv)) * sizeof(double);
}

// members of subdomain d on its box: fidx[j] (node of member j), fmem[k]
// (member of node k, -1 for a missing node); false when a center is off the
// grid. Every thread of the block calls it (one barrier inside).
__device__ bool shared_members(int n, int nf, int64_t s0, const int32_t* slot_pt,
                               const int32_t* idx, const double* xs, const double* ys,
                               const double* gx, const double* gy, int ngx, int j0, int nc,
                               int i0, int* fidx, int* fmem, int* bad) {
  for (int k = threadIdx.x; k < nf; k += kThreads) fmem[k] = -1;
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += kThreads) {
    const int c = slot_pt[s0 + j];
    const int q = idx[c];
    const int row = q / ngx, col = q % ngx;
    const int k = (row - i0) * nc + (col - j0);
    const bool on = col - j0 >= 0 && col - j0 < nc && k >= 0 && k < nf && xs[c] == gx[col] &&
                    ys[c] == gy[row];
    if (!on) {
      *bad = 1;
      continue;
    }
    fidx[j] = k;
    fmem[k] = j;
  }
  __syncthreads();
  return *bad == 0;
}

// the missing nodes in node order (at most `list` of them listed); their
// number
__device__ int shared_missing(int nf, const int* fmem, int cap_r, int* rlist, int* nr_sh,
                              int list = kCapMaxR) {
  if (threadIdx.x == 0) {
    int r = 0;
    for (int k = 0; k < nf && r <= cap_r; ++k)
      if (fmem[k] < 0) {
        if (r < list) rlist[r] = k;
        ++r;
      }
    *nr_sh = r;
  }
  __syncthreads();
  return *nr_sh;
}

// Launched twice: NV = 4 checks every shared subdomain and computes the
// columns of those missing at most 4 nodes; NV = kSharedNV those of the
// others (checked already: status 0, more than 4 missing nodes).
template <int NV>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NV <= 4 ? 3 : 2))) void
k_tps_shared_columns(const int32_t* __restrict__ order, const int32_t* __restrict__ dpat,
                     const int32_t* __restrict__ dbox, const int32_t* __restrict__ pat_n,
                     const int64_t* __restrict__ pdom_off, const int64_t* __restrict__ pmat_off,
                     const double* __restrict__ pmats, const double* __restrict__ prefl,
                     const double* __restrict__ pvwork, const int32_t* __restrict__ pstatus,
                     const double* __restrict__ gx, const double* __restrict__ gy, int ngx,
                     const double* __restrict__ xs, const double* __restrict__ ys,
                     const int32_t* __restrict__ idx, const int64_t* __restrict__ dom_off,
                     const int32_t* __restrict__ slot_pt, int32_t* __restrict__ status,
                     double* __restrict__ gmat, double* __restrict__ hfac,
                     double* __restrict__ gpoly, int cap_r, int nm) {
  const int nvalid = order[gridDim.x];
  if (int(blockIdx.x) >= nvalid) return;
  const int d = order[xcd_rank(blockIdx.x, nvalid)];
  if (d < 0) return;
  constexpr bool kFirst = NV <= 4;  // the launch that checks every subdomain
  const int p = dpat[d];
  const int tid = threadIdx.x;
  if (!kFirst && status[d]) return;
  if (pstatus[p]) {
    if (tid == 0) status[d] = 1;
    return;
  }
  const int64_t s0 = dom_off[d];
  const int n = int(dom_off[d + 1] - s0);
  const int nf = pat_n[p];
  const int j0 = dbox[4 * d], nc = dbox[4 * d + 1], i0 = dbox[4 * d + 2];
  const int64_t ps0 = pdom_off[p];
  const double* E = pmats + pmat_off[p];
  const int m = nf - 3;
  const int64_t ld = tps_mat_ld(m);
  const double* A = E + tps_mat_e_size(m);
  const double* RF = prefl + 16 * p;
  const double* VW = pvwork + 6 * ps0;
  const double tau[3] = {RF[0], RF[1], RF[2]};
  // dynamic LDS: x[NV][nm] (doubles), fidx[nm], fmem[nm], then the solve's
  // scratch
  extern __shared__ double dyn[];
  double* x = dyn;
  int* fidx = reinterpret_cast<int*>(dyn + NV * nm);
  int* fmem = fidx + nm;
  double* scr = dyn + (NV + 1) * nm;
  __shared__ double red[NV * (kThreads / 64)];
  __shared__ int rlist[kCapMaxR];
  __shared__ int nr_sh;
  __shared__ double H[kCapMaxR][kCapMaxR + 1];
  __shared__ double a8[NV][3];
  __shared__ int bad;
  if (tid == 0) bad = 0;
  if (!shared_members(n, nf, s0, slot_pt, idx, xs, ys, gx, gy, ngx, j0, nc, i0, fidx, fmem,
                      &bad)) {
    if (tid == 0) status[d] = 2;  // a center off the grid: solved again by host sizing
    return;
  }
  const int nr = shared_missing(nf, fmem, cap_r, rlist, &nr_sh);
  if (nr > cap_r || nr > kCapMaxR || n + nr != nf) {
    if (tid == 0) status[d] = 2;
    return;
  }
  if (nr == 0 || (kFirst ? nr > NV : nr <= 4)) return;  // the other launch's
  // G columns g0 .. g0 + NV - 1 per pass over the factor
  double* G = gmat + int64_t(d) * nm * cap_r;  // columns of G_w, nm apart
  double* GA = gpoly + int64_t(d) * kCapMaxR * 3;
  for (int g0 = 0; g0 < nr; g0 += NV) {
    for (int k = tid; k < nf; k += kThreads)
#pragma unroll
      for (int q = 0; q < NV; ++q)
        x[q * nm + k] = (g0 + q < nr && k == rlist[g0 + q]) ? 1.0 : 0.0;
    __syncthreads();
    saddle_solve_n<NV>(x, nm, nf, VW, tau, A, ld, E, RF, red, scr, a8);
    for (int q = 0; q < NV; ++q) {
      const int r = g0 + q;
      if (r >= nr) break;
      for (int k = tid; k < nf; k += kThreads) G[int64_t(r) * nm + k] = x[q * nm + k];
      if (tid < nr) H[tid][r] = x[q * nm + rlist[tid]];
      if (tid < 3) GA[3 * r + tid] = a8[q][tid];
    }
    __syncthreads();
  }
  // H = L_H L_H^T in place (lower triangle), one thread
  if (tid == 0) {
    int ok = 1;
    for (int j = 0; j < nr && ok; ++j) {
      double dj = H[j][j];
      for (int k = 0; k < j; ++k) dj -= H[j][k] * H[j][k];
      if (!(dj > 0.0)) {
        ok = 0;
        break;
      }
      dj = sqrt(dj);
      H[j][j] = dj;
      for (int i = j + 1; i < nr; ++i) {
        double v = H[i][j];
        for (int k = 0; k < j; ++k) v -= H[i][k] * H[j][k];
        H[i][j] = v / dj;
      }
    }
    if (!ok) status[d] = 1;  // M_S singular, as the direct factorisation reports it
  }
  __syncthreads();
  double* HF = hfac + int64_t(d) * kCapMaxR * kCapMaxR;
  for (int e = tid; e < nr * nr; e += kThreads) HF[e] = H[e / nr][e % nr];
}

constexpr int kIterNV = 2;

__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void
k_tps_shared_iter(const int32_t* __restrict__ order, const int32_t* __restrict__ dpat,
                  const int32_t* __restrict__ dbox, const int32_t* __restrict__ pat_n,
                  const int64_t* __restrict__ pdom_off, const int64_t* __restrict__ pmat_off,
                  const double* __restrict__ pmats, const double* __restrict__ prefl,
                  const double* __restrict__ pvwork, const double* __restrict__ gx,
                  const double* __restrict__ gy, int ngx, const double* __restrict__ xs,
                  const double* __restrict__ ys, const int32_t* __restrict__ idx,
                  const float* __restrict__ vx, const float* __restrict__ vy,
                  const int64_t* __restrict__ dom_off, const int32_t* __restrict__ slot_pt,
                  const double* __restrict__ smoothing_d, double threshold,
                  double* __restrict__ smooth_slots, int64_t nslots, double* __restrict__ weights,
                  int64_t wstride, double* __restrict__ stats, const int32_t* __restrict__ status,
                  double* __restrict__ slot_xy, const double* __restrict__ gmat,
                  const double* __restrict__ hfac, const double* __restrict__ gpoly, int cap_r,
                  int nm) {
  const int nvalid = order[gridDim.x];
  if (int(blockIdx.x) >= nvalid) return;
  const int d = order[xcd_rank(blockIdx.x, nvalid)];
  if (d < 0 || status[d]) return;
  const int p = dpat[d];
  const int64_t s0 = dom_off[d];
  const int n = int(dom_off[d + 1] - s0);
  const int nf = pat_n[p];
  const int tid = threadIdx.x;
  const int j0 = dbox[4 * d], nc = dbox[4 * d + 1], i0 = dbox[4 * d + 2];
  const int64_t ps0 = pdom_off[p];
  const int m = nf - 3;
  const double* E = pmats + pmat_off[p];
  const int64_t ld = tps_mat_ld(m);
  const double* A = E + tps_mat_e_size(m);
  const double* RF = prefl + 16 * p;
  const double* VW = pvwork + 6 * ps0;
  const double tau[3] = {RF[0], RF[1], RF[2]};
  const double smoothing = smoothing_d[d];
  const double x0 = gx[j0], y0 = gy[i0];
  const double* G = gmat + int64_t(d) * nm * cap_r;
  // dynamic LDS: x[2][nm] (doubles), val[2][nm] (floats), fidx[nm],
  // fmem[nm], then the solve's scratch
  extern __shared__ double dyn[];
  double* x = dyn;
  float* val[2] = {reinterpret_cast<float*>(dyn + kIterNV * nm),
                   reinterpret_cast<float*>(dyn + kIterNV * nm) + nm};
  int* fidx = reinterpret_cast<int*>(dyn + (kIterNV + 1) * nm);
  int* fmem = fidx + nm;
  double* scr = dyn + (kIterNV + 2) * nm;
  __shared__ double red[kIterNV * (kThreads / 64)];
  __shared__ int rlist[kCapMaxR];
  __shared__ int nr_sh;
  __shared__ double H[kCapMaxR][kCapMaxR + 1];
  __shared__ double ga[kCapMaxR][3];
  __shared__ double lam[2][kCapMaxR];
  __shared__ double zl[2][kCapMaxR];
  __shared__ double a2[kIterNV][3];  // polynomial terms of the solved vectors
  __shared__ int go[2];
  __shared__ int bad;
  if (tid == 0) bad = 0;
  // the members (the first kernel checked them); values; coordinates by slot
  shared_members(n, nf, s0, slot_pt, idx, xs, ys, gx, gy, ngx, j0, nc, i0, fidx, fmem, &bad);
  for (int j = tid; j < n; j += kThreads) {
    const int c = slot_pt[s0 + j];
    val[0][j] = vx[c];
    val[1][j] = vy[c];
    slot_xy[s0 + j] = xs[c];
    slot_xy[nslots + s0 + j] = ys[c];
  }
  const int nr = shared_missing(nf, fmem, cap_r, rlist, &nr_sh);
  const double* HF = hfac + int64_t(d) * kCapMaxR * kCapMaxR;
  const double* GA = gpoly + int64_t(d) * kCapMaxR * 3;
  for (int e = tid; e < nr * nr; e += kThreads) H[e / nr][e % nr] = HF[e];
  for (int e = tid; e < 3 * nr; e += kThreads) ga[e / 3][e % 3] = GA[e];
  if (tid < 2) go[tid] = 1;
  __syncthreads();
  double nfixed[2] = {0.0, 0.0}, maxdiff[2] = {0.0, 0.0};
  int count[2] = {0, 0};
  for (int it = 0; it < 11; ++it) {
    if (!go[0] && !go[1]) break;
    // y = M_F^-1 [v'; 0] for the current values (v' = 0 at the missing nodes)
    for (int k = tid; k < nf; k += kThreads) {
      const int j = fmem[k];
      x[k] = j >= 0 ? double(val[0][j]) : 0.0;
      x[nm + k] = j >= 0 ? double(val[1][j]) : 0.0;
    }
    __syncthreads();
    saddle_solve_n<kIterNV>(x, nm, nf, VW, tau, A, ld, E, RF, red, scr, a2);
    // lambda = -H^-1 y_w[R] (two triangular solves with L_H, the forward
    // one's result in zl)
    if (tid < 2) {
      const int c = tid;
      for (int i = 0; i < nr; ++i) {
        double v = -x[c * nm + rlist[i]];
        for (int k = 0; k < i; ++k) v -= H[i][k] * zl[c][k];
        zl[c][i] = v / H[i][i];
      }
      for (int i = nr - 1; i >= 0; --i) {
        double v = zl[c][i];
        for (int k = i + 1; k < nr; ++k) v -= H[k][i] * lam[c][k];
        lam[c][i] = v / H[i][i];
      }
    }
    __syncthreads();
    // w = y + G lambda at the members (a = y_a + G_a lambda: below)
    for (int j = tid; j < n; j += kThreads) {
      const int k = fidx[j];
      double w0 = x[k], w1 = x[nm + k];
      for (int r = 0; r < nr; ++r) {
        const double gk = G[int64_t(r) * nm + k];
        w0 += gk * lam[0][r];
        w1 += gk * lam[1][r];
      }
      x[k] = w0;
      x[nm + k] = w1;
    }
    __syncthreads();
    for (int c = 0; c < 2; ++c) {
      if (!go[c]) continue;  // uniform across the block
      count[c] += 1;
      double dmax = 0.0;
      int ne = 0;
      for (int j = tid; j < n; j += kThreads) {
        const double v = double(val[c][j]);
        const double sm = v - smoothing * x[c * nm + fidx[j]];
        smooth_slots[c * nslots + s0 + j] = sm;
        const double diff = fabs(sm - v);
        dmax = fmax(dmax, diff);
        if (threshold >= 0.0 && diff > threshold) ++ne;
      }
      double* W = weights + c * wstride + s0 + 3 * int64_t(d);
      for (int j = tid; j < n; j += kThreads) W[j] = x[c * nm + fidx[j]];
      if (tid == 0) {
        double wa[3];
        for (int t = 0; t < 3; ++t) {
          double v = a2[c][t];
          for (int k = 0; k < nr; ++k) v += ga[k][t] * lam[c][k];
          wa[t] = v;
        }
        // P = [1, y - y0, x - x0] -> [1, y, x]
        W[n] = wa[0] - wa[1] * y0 - wa[2] * x0;
        W[n + 1] = wa[1];
        W[n + 2] = wa[2];
      }
      const double dm = block_max(dmax, red);
      const double nesum = block_sum(double(ne), red);
      if (count[c] == 1) maxdiff[c] = dm;
      const bool more = threshold >= 0.0 && nesum > 0.0 && count[c] < 11;
      if (more) {
        nfixed[c] += nesum;
        for (int j = tid; j < n; j += kThreads) {
          const double v = double(val[c][j]);
          const double sm = v - smoothing * x[c * nm + fidx[j]];
          if (fabs(sm - v) > threshold) val[c][j] = float(sm);
        }
      }
      __syncthreads();
      if (tid == 0) go[c] = more ? 1 : 0;
      __syncthreads();
    }
  }
  if (tid == 0) {
    for (int c = 0; c < 2; ++c) {
      stats[6 * d + 3 * c + 0] = nfixed[c];
      stats[6 * d + 3 * c + 1] = maxdiff[c];
      stats[6 * d + 3 * c + 2] = count[c];
    }
  }
}

// ---------------------------------------------------------------------------
// 4c. the shared solves by the explicit inverse of each pattern's saddle-
//     point matrix
// ---------------------------------------------------------------------------
// For a pattern of n nodes, M^-1 [v; 0] = [X v; Y^T v] with
//   X = Q [[0, 0], [0, C^-1]] Q^T (n x n, symmetric),
//   Y = Q [I_3; -C^-1 E] R^-T (n x 3),
// from the null-space factorisation (k_tps_prep, k_tps_chol): the
// reflectors Q = H_0 H_1 H_2 = I - V T V^T of P = Q [R; 0], C = L L^T and
// E = B'[3:, :3]. They are built once per pattern:
//   k_tps_tri_inverse  Linv = L^-1, block columns in parallel;
//   k_tps_inv_gram     C^-1 = Linv^T Linv (64 x 64 tiles), over L in the
//                      pattern's matrix region (E is kept);
//   k_tps_xprep        U = C^-1 [V[3:, :] | E] (one pass over C^-1), then the
//                      rank-6 form X = Z - A Cc^T - Cc A^T of the two-sided
//                      reflection (Z = [[0, 0], [0, C^-1]], A = V T,
//                      Cc = U' - A S / 2, U' = Z V = [0; C^-1 V[3:, :]],
//                      S = V^T U') and Y = F - A (V^T F),
//                      F = [R^-T; -C^-1 E R^-T];
//   k_tps_xfill        X, rows of tps_x_ld(n) doubles (Y in columns n .. n + 2,
//                      written by k_tps_xprep).
// A subdomain missing the nodes R of its pattern then needs no solve: its
// capacitance columns are rows of X (G_w = X[:, R], G_a = Y[R, :],
// H = X[R, R]), and each outlier iteration is one product with X
// (k_tps_shared_x: up to four subdomains of one pattern per workgroup, one
// read of X serving all of them).

// Launch shape of these kernels: a fixed grid whose workgroups stride over
// the items (pattern, column or tile or rows) of the patterns in use
// (plist[0 .. plist[npat]), written by k_tps_pattern_layout), most work
// first, instead of a grid over every pattern slot, most of them empty.

// Linv of pattern p, block column J: Linv_JJ = inv(L_JJ) (the factorisation keeps
// inv(L_JJ)^T in the strict upper triangle of the diagonal block, L_JJ's
// diagonal on it), then for I > J, Linv_IJ = -inv(L_II) sum_{K=J}^{I-1}
// L_IK Linv_KJ, the blocks of each term staged in LDS with the next term's
// loads in flight, 4 outputs per thread. Linv_IJ goes to the rows of I and
// the columns of J of Ainv (the layout of the factor's matrix region), its
// transpose to the rows of J and the columns of I; the workgroup reads its
// own earlier blocks back after a barrier.
__device__ void tri_inverse_column(int p, int J, const int32_t* __restrict__ pat_n,
                                   const int64_t* __restrict__ pmat_off,
                                   const double* __restrict__ pmats,
                                   const int32_t* __restrict__ pstatus,
                                   double* __restrict__ pinv) {
  static_assert(kThreads == 256 && kNB == 32, "thread layout of tri_inverse_column");
  const int n = pat_n[p];
  if (n < 4 || pstatus[p]) return;
  const int m = n - 3;
  const int nblk = (m + kNB - 1) / kNB;
  if (J >= nblk) return;
  const int64_t ld = tps_mat_ld(m);
  const double* A = pmats + pmat_off[p] + tps_mat_e_size(m);
  double* Inv = pinv + pmat_off[p] + tps_mat_e_size(m);
  __shared__ double Lb[kNB][kNB + 1];  // L_IK, then inv(L_II)
  __shared__ double Xb[kNB][kNB + 1];  // Linv_KJ, then the sum
  const int tid = threadIdx.x;
  const int r0J = J * kNB, bsJ = min(kNB, m - r0J);
  const int ti = tid >> 3, tc = tid & 7;   // outputs (ti, tc + 8 v), v < 4
  const int er = tid >> 5, ec = tid & 31;  // staged elements (er + 8 t, ec), t < 4
  // inv(L_II) of diagonal block I into Lb (zero beyond the matrix)
  auto stage_inv_diag = [&](int r0, int bs) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int i = er + 8 * t;
      double v = 0.0;
      if (i < bs && ec < bs) {
        if (ec == i) v = 1.0 / A[int64_t(r0 + i) * ld + r0 + i];
        else if (ec < i) v = A[int64_t(r0 + ec) * ld + r0 + i];
      }
      Lb[i][ec] = v;
    }
  };
  // L_IK and Linv_KJ (its lower triangle only for K = J) into registers
  auto load_term = [&](int r0I, int bsI, int K, double* la, double* lx) {
    const int r0K = K * kNB, bsK = min(kNB, m - r0K);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int i = er + 8 * t;
      la[t] = (i < bsI && ec < bsK) ? A[int64_t(r0I + i) * ld + r0K + ec] : 0.0;
      lx[t] = (i < bsK && ec < bsJ && (K > J || ec <= i))
                  ? Inv[int64_t(r0K + i) * ld + r0J + ec]
                  : 0.0;
    }
  };
  stage_inv_diag(r0J, bsJ);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int i = er + 8 * t;
    if (i < bsJ && ec <= i) {
      const double v = Lb[i][ec];
      Inv[int64_t(r0J + i) * ld + r0J + ec] = v;
      if (ec < i) Inv[int64_t(r0J + ec) * ld + r0J + i] = v;
    }
  }
  __syncthreads();
  for (int I = J + 1; I < nblk; ++I) {
    const int r0I = I * kNB, bsI = min(kNB, m - r0I);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    double la[4], lx[4];
    load_term(r0I, bsI, J, la, lx);
    for (int K = J; K < I; ++K) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        Lb[er + 8 * t][ec] = la[t];
        Xb[er + 8 * t][ec] = lx[t];
      }
      __syncthreads();
      if (K + 1 < I) load_term(r0I, bsI, K + 1, la, lx);
#pragma unroll 4
      for (int k = 0; k < kNB; ++k) {
        const double l = Lb[ti][k];
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] += l * Xb[k][tc + 8 * v];
      }
      __syncthreads();
    }
    // Linv_IJ = -inv(L_II) sum
    stage_inv_diag(r0I, bsI);
#pragma unroll
    for (int v = 0; v < 4; ++v) Xb[ti][tc + 8 * v] = acc[v];
    __syncthreads();
    double x[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int k = 0; k < kNB; ++k) {
      const double l = Lb[ti][k];
#pragma unroll
      for (int v = 0; v < 4; ++v) x[v] -= l * Xb[k][tc + 8 * v];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int c = tc + 8 * v;
      if (ti < bsI && c < bsJ) {
        Inv[int64_t(r0I + ti) * ld + r0J + c] = x[v];
        Inv[int64_t(r0J + c) * ld + r0I + ti] = x[v];
      }
    }
    __syncthreads();
  }
}

// items: block column J = w / npu (the long columns first) of pattern
// plist[w % npu]
__global__ __launch_bounds__(kThreads) void k_tps_tri_inverse(
    const int32_t* __restrict__ plist, int npat, int nblk_max, const int32_t* __restrict__ pat_n,
    const int64_t* __restrict__ pmat_off, const double* __restrict__ pmats,
    const int32_t* __restrict__ pstatus, double* __restrict__ pinv) {
  const int npu = plist[npat];
  for (int w = blockIdx.x; w < npu * nblk_max; w += gridDim.x)
    tri_inverse_column(plist[w % npu], w / npu, pat_n, pmat_off, pmats, pstatus, pinv);
}

// C^-1 = Linv^T Linv of pattern p: 64 x 64 tile (I, J), I >= J, number
// t = I (I + 1) / 2 + J: the sum over the 32-row blocks K of Linv_KI^T
// Linv_KJ (rows 64 I on), written at (I, J) and mirrored at (J, I) over L (E
// is kept). Each thread has 4 x 4 outputs; the blocks of each term are
// staged in LDS with the next term's loads in flight. The two threads of a
// mirrored pair inside a diagonal tile add the same products in the same
// order: C^-1 is exactly symmetric.
__device__ void inv_gram_tile(int p, int t, const int32_t* __restrict__ pat_n,
                              const int64_t* __restrict__ pmat_off, double* __restrict__ pmats,
                              const int32_t* __restrict__ pstatus,
                              const double* __restrict__ pinv) {
  static_assert(kThreads == 256 && kNB == 32, "thread layout of inv_gram_tile");
  constexpr int kT = 64;
  const int n = pat_n[p];
  if (n < 4 || pstatus[p]) return;
  const int m = n - 3;
  const int nblk = (m + kNB - 1) / kNB, nt = (m + kT - 1) / kT;
  int I = int((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  while (I * (I + 1) / 2 > t) --I;
  const int J = t - I * (I + 1) / 2;
  if (I >= nt) return;
  const int64_t ld = tps_mat_ld(m);
  const double* Li = pinv + pmat_off[p] + tps_mat_e_size(m);
  double* Xc = pmats + pmat_off[p] + tps_mat_e_size(m);
  __shared__ double Ab[kNB][kT + 1];  // Ab[k][a] = Linv[32 K + k][64 I + a]
  __shared__ double Bb[kNB][kT + 1];  // Bb[k][b] = Linv[32 K + k][64 J + b]
  const int tid = threadIdx.x;
  const int cI = kT * I, cJ = kT * J;
  const int ta = tid >> 4, tb = tid & 15;  // outputs (ta + 16 u, tb + 16 v)
  // staged element e = tid + 256 u: row e / 64, column e % 64 (lower
  // triangle of Linv only)
  auto load = [&](int K, double* la, double* lb) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + kThreads * u;
      const int R = kNB * K + (e >> 6), c = e & 63;
      la[u] = (R < m && cI + c <= R) ? Li[int64_t(R) * ld + cI + c] : 0.0;
      lb[u] = (R < m && cJ + c <= R) ? Li[int64_t(R) * ld + cJ + c] : 0.0;
    }
  };
  double acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
  double la[8], lb[8];
  const int K0 = cI / kNB;
  load(K0, la, lb);
  for (int K = K0; K < nblk; ++K) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + kThreads * u;
      Ab[e >> 6][e & 63] = la[u];
      Bb[e >> 6][e & 63] = lb[u];
    }
    __syncthreads();
    if (K + 1 < nblk) load(K + 1, la, lb);
#pragma unroll 2
    for (int k = 0; k < kNB; ++k) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = Ab[k][ta + 16 * u];
#pragma unroll
      for (int v = 0; v < 4; ++v) b[v] = Bb[k][tb + 16 * v];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = cI + ta + 16 * u, c = cJ + tb + 16 * v;
      if (r < m && c < m) {
        Xc[int64_t(r) * ld + c] = acc[u][v];
        Xc[int64_t(c) * ld + r] = acc[u][v];
      }
    }
}

// items: tile t = w / npu (the tiles of most terms first) of pattern
// plist[w % npu]
__global__ __launch_bounds__(kThreads) void k_tps_inv_gram(
    const int32_t* __restrict__ plist, int npat, int ntile_max, const int32_t* __restrict__ pat_n,
    const int64_t* __restrict__ pmat_off, double* __restrict__ pmats,
    const int32_t* __restrict__ pstatus, const double* __restrict__ pinv) {
  const int npu = plist[npat];
  for (int w = blockIdx.x; w < npu * ntile_max; w += gridDim.x)
    inv_gram_tile(plist[w % npu], w / npu, pat_n, pmat_off, pmats, pstatus, pinv);
}

// per pattern: U' = [0; C^-1 V[3:, :]] and C^-1 E by rows (lanes over rows,
// reading the symmetric C^-1 by columns), into pac (6 doubles per node);
// then A = V T and Cc = U' - A S / 2 over them, and Y^T after X. Dynamic
// LDS: V[3:, :] and E, 6 nm doubles.
__device__ void xprep_pattern(int p, const int32_t* __restrict__ pat_n,
                              const int64_t* __restrict__ pmat_off,
                              const double* __restrict__ pmats,
                              const int64_t* __restrict__ pdom_off,
                              const double* __restrict__ prefl,
                              const double* __restrict__ pvwork,
                              const int32_t* __restrict__ pstatus,
                              const int64_t* __restrict__ pxo, double* __restrict__ pinv,
                              double* __restrict__ pac) {
  const int n = pat_n[p];
  if (n < 4 || pstatus[p]) return;
  const int m = n - 3;
  const int64_t ld = tps_mat_ld(m);
  const double* E = pmats + pmat_off[p];
  const double* Xc = E + tps_mat_e_size(m);
  const double* RF = prefl + 16 * p;
  const double* VW = pvwork + 6 * pdom_off[p];
  double* AC = pac + 6 * pdom_off[p];
  const int64_t ldx = tps_x_ld(n);
  double* Xp = pinv + pxo[p];
  extern __shared__ double dyn[];
  double* VE = dyn;  // [m][6]: V[3 + j][0..2], E[j][0..2]
  __shared__ double red[18 * (kThreads / 64)];
  const int tid = threadIdx.x;
  for (int j = tid; j < m; j += kThreads)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      VE[6 * j + k] = VW[6 * (3 + j) + k];
      VE[6 * j + 3 + k] = E[3 * int64_t(j) + k];
    }
  for (int j = tid; j < 3; j += kThreads)
#pragma unroll
    for (int k = 0; k < 6; ++k) AC[6 * j + k] = 0.0;
  __syncthreads();
  // U rows (slot j = 3 + i)
  for (int i = tid; i < m; i += kThreads) {
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double* col = Xc + i;
    constexpr int U8 = 8;
    int j = 0;
    for (; j + U8 <= m; j += U8) {
      double a[U8];
#pragma unroll
      for (int u = 0; u < U8; ++u) a[u] = col[int64_t(j + u) * ld];
#pragma unroll
      for (int u = 0; u < U8; ++u)
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[k] += a[u] * VE[6 * (j + u) + k];
    }
    for (; j < m; ++j) {
      const double a = col[int64_t(j) * ld];
#pragma unroll
      for (int k = 0; k < 6; ++k) acc[k] += a * VE[6 * j + k];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) AC[6 * (3 + i) + k] = acc[k];
  }
  __syncthreads();
  const double T[3][3] = {{RF[9], RF[10], RF[11]}, {0.0, RF[12], RF[13]}, {0.0, 0.0, RF[14]}};
  // R^-1 (R upper)
  double Ri[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  Ri[0][0] = 1.0 / RF[3];
  Ri[1][1] = 1.0 / RF[6];
  Ri[2][2] = 1.0 / RF[8];
  Ri[0][1] = -RF[4] * Ri[1][1] / RF[3];
  Ri[1][2] = -RF[7] * Ri[2][2] / RF[6];
  Ri[0][2] = -(RF[4] * Ri[1][2] + RF[5] * Ri[2][2]) / RF[3];
  // row j of F = [R^-T; -C^-1 E R^-T] (u: slot j of AC, the C^-1 E part)
  auto f_row = [&](int j, const double* u, double* f) {
    if (j < 3) {
#pragma unroll
      for (int k = 0; k < 3; ++k) f[k] = Ri[k][j];
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) f[k] = -(u[3] * Ri[k][0] + u[4] * Ri[k][1] + u[5] * Ri[k][2]);
    }
  };
  // S = V^T U', G3 = V^T F
  double part[18];
#pragma unroll
  for (int e = 0; e < 18; ++e) part[e] = 0.0;
  for (int j = tid; j < n; j += kThreads) {
    const double v[3] = {VW[6 * j], VW[6 * j + 1], VW[6 * j + 2]};
    const double* u = AC + 6 * j;
    double f[3];
    f_row(j, u, f);
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        part[3 * k + l] += v[k] * u[l];
        part[9 + 3 * k + l] += v[k] * f[l];
      }
  }
  block_sum_n<18>(part, red);
  // Q Z Q^T = Z - A U'^T - U' A^T + A S A^T, S = part[0..8]
  for (int j = tid; j < n; j += kThreads) {
    double* u = AC + 6 * j;
    const double v[3] = {VW[6 * j], VW[6 * j + 1], VW[6 * j + 2]};
    double a[3], f[3], c[3], y[3];
    f_row(j, u, f);
#pragma unroll
    for (int k = 0; k < 3; ++k) a[k] = v[0] * T[0][k] + v[1] * T[1][k] + v[2] * T[2][k];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      c[k] = u[k] - 0.5 * (a[0] * part[k] + a[1] * part[3 + k] + a[2] * part[6 + k]);
      y[k] = f[k] - (a[0] * part[9 + k] + a[1] * part[12 + k] + a[2] * part[15 + k]);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      u[k] = a[k];
      u[3 + k] = c[k];
      Xp[int64_t(j) * ldx + n + k] = y[k];
    }
  }
  __syncthreads();  // LDS reused by the next pattern
}

__global__ __launch_bounds__(kThreads) void k_tps_xprep(
    const int32_t* __restrict__ plist, int npat, const int32_t* __restrict__ pat_n,
    const int64_t* __restrict__ pmat_off, const double* __restrict__ pmats,
    const int64_t* __restrict__ pdom_off, const double* __restrict__ prefl,
    const double* __restrict__ pvwork, const int32_t* __restrict__ pstatus,
    const int64_t* __restrict__ pxo, double* __restrict__ pinv, double* __restrict__ pac) {
  const int npu = plist[npat];
  for (int w = blockIdx.x; w < npu; w += gridDim.x)
    xprep_pattern(plist[w], pat_n, pmat_off, pmats, pdom_off, prefl, pvwork, pstatus, pxo, pinv,
                  pac);
}

// X = Z - A Cc^T - Cc A^T, rows 8 c .. 8 c + 7 (c = w / npu) of pattern
// plist[w % npu]; X[i][j] and X[j][i] add the same two dot products: X is
// exactly symmetric
#pragma clang fp contract(off)
__global__ __launch_bounds__(kThreads) void k_tps_xfill(
    const int32_t* __restrict__ plist, int npat, int nchunk, const int32_t* __restrict__ pat_n,
    const int64_t* __restrict__ pmat_off, const double* __restrict__ pmats,
    const int64_t* __restrict__ pdom_off, const int32_t* __restrict__ pstatus,
    const int64_t* __restrict__ pxo, const double* __restrict__ pac, double* __restrict__ pinv) {
  const int npu = plist[npat];
  for (int w = blockIdx.x; w < npu * nchunk; w += gridDim.x) {
  const int p = plist[w % npu];
  const int n = pat_n[p];
  const int i0 = 8 * (w / npu);
  if (n < 4 || pstatus[p] || i0 >= n) continue;
  const int m = n - 3;
  const int64_t ld = tps_mat_ld(m);
  const double* Xc = pmats + pmat_off[p] + tps_mat_e_size(m);
  const double* AC = pac + 6 * pdom_off[p];
  const int64_t ldx = tps_x_ld(n);
  double* X = pinv + pxo[p];
  for (int i = i0; i < min(i0 + 8, n); ++i) {
    const double* ri = AC + 6 * i;
    const double a0 = ri[0], a1 = ri[1], a2 = ri[2], c0 = ri[3], c1 = ri[4], c2 = ri[5];
    for (int j = threadIdx.x; j < n; j += kThreads) {
      const double* rj = AC + 6 * j;
      const double z = (i >= 3 && j >= 3) ? Xc[int64_t(i - 3) * ld + (j - 3)] : 0.0;
      const double s1 = fma(a2, rj[5], fma(a1, rj[4], a0 * rj[3]));
      const double s2 = fma(c2, rj[2], fma(c1, rj[1], c0 * rj[0]));
      X[int64_t(i) * ldx + j] = z - (s1 + s2);
    }
  }
  }
}
#pragma clang fp contract(on)

// dynamic LDS of k_tps_shared_x<g>: x (floats) [nm][2g] and y [2g][nm + 4]
// (first the matrices H of the subdomains, kCapX (kCapX + 1) doubles each),
// then the values (floats) [g][2][nm] and member nodes [g][nm]
constexpr int shared_x_head(int g, int nm) {
  return g * nm + 2 * g * (nm + 4) > g * kCapX * (kCapX + 1) ? g * nm + 2 * g * (nm + 4)
                                                              : g * kCapX * (kCapX + 1);
}
constexpr size_t shared_x_lds(int g, int nm) {
  return (size_t(shared_x_head(g, nm)) + size_t(g) * nm) * sizeof(double) + size_t(g) * nm * 4;
}
// subdomains per workgroup of k_tps_shared_x: 4 while two workgroups fit a
// CU's LDS, else 2, else 1
constexpr int shared_x_group(int nm) {
  return shared_x_lds(4, nm) <= 80 * 1024 ? 4 : shared_x_lds(2, nm) <= 80 * 1024 ? 2 : 1;
}

// [y; Y^T v'] = [X | Y]^T v' for the K columns of x (v' of component aq[k]
// in column k of xv, rows of NV floats; aq[k] < 0: a zero column, not
// stored): X is symmetric, so lanes over output pairs i, i + 1 read the rows
// of [X | Y] by columns, 16 bytes a load, kU loads in flight; outputs
// i >= nf + 3 read the padding. y of component q goes to yv[q ny + i].
// Each component's sums are the same for every K.
template <int K, int NV>
__device__ __forceinline__ void x_products(const double* X, int64_t ldx, int nf, const float* xv,
                                           double* yv, int ny, const int* aq) {
  for (int i = 2 * int(threadIdx.x); i < nf + 3; i += 2 * kThreads) {
    double acc0[K], acc1[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc0[k] = acc1[k] = 0.0;
    const double* col = X + i;
    constexpr int kU = 8;
    int j = 0;
    for (; j + kU <= nf; j += kU) {
      double2 a[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u)
        a[u] = *reinterpret_cast<const double2*>(col + int64_t(j + u) * ldx);
#pragma unroll
      for (int u = 0; u < kU; ++u)
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const double xq = double(xv[(j + u) * NV + k]);
          acc0[k] += a[u].x * xq;
          acc1[k] += a[u].y * xq;
        }
    }
    for (; j < nf; ++j) {
      const double2 a = *reinterpret_cast<const double2*>(col + int64_t(j) * ldx);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double xq = double(xv[j * NV + k]);
        acc0[k] += a.x * xq;
        acc1[k] += a.y * xq;
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int q = aq[k];
      if (q < 0) continue;
      if (i + 1 < nf + 3)
        *reinterpret_cast<double2*>(yv + q * ny + i) = make_double2(acc0[k], acc1[k]);
      else
        yv[q * ny + i] = acc0[k];
    }
  }
}

// The shared subdomains of group gi (up to G of one pattern, consecutive in
// the shared order) with the explicit inverse: the checks of
// k_tps_shared_columns, then the outlier iterations of k_tps_shared_iter
// (thin_plate_spline_subdom.py:236-280), whose solve is y = X v' (v' the
// values, 0 at the missing nodes), lambda = -H^-1 y[R], w = y + X[:, R]
// lambda, a = Y^T v' + Y[R, :]^T lambda, for the 2 G components at once.
// Each component's arithmetic is independent of the others in the group.
template <int G>
__device__ void shared_x_solve(
    int gi, double* __restrict__ hinv, const int32_t* __restrict__ order,
    const int32_t* __restrict__ groups,
    const int32_t* __restrict__ dpat, const int32_t* __restrict__ dbox,
    const int32_t* __restrict__ pat_n, const int32_t* __restrict__ pstatus,
    const int64_t* __restrict__ pxo, const double* __restrict__ pinv,
    const double* __restrict__ gx, const double* __restrict__ gy, int ngx,
    const double* __restrict__ xs, const double* __restrict__ ys,
    const int32_t* __restrict__ idx, const float* __restrict__ vx,
    const float* __restrict__ vy, const int64_t* __restrict__ dom_off,
    const int32_t* __restrict__ slot_pt, const double* __restrict__ smoothing_d,
    double threshold, double* __restrict__ smooth_slots, int64_t nslots,
    double* __restrict__ weights, int64_t wstride, double* __restrict__ stats,
    int32_t* __restrict__ status, double* __restrict__ slot_xy, int cap_r, int nm, int nd) {
  static_assert(G >= 1 && G <= kThreads / 64, "one wavefront per subdomain for its H");
  constexpr int NV = 2 * G;
  const int b0 = groups[gi], cnt = groups[nd + 1 + gi];
  const int tid = threadIdx.x;
  const int p = dpat[order[b0]];
  if (pstatus[p]) {
    if (tid < cnt) status[order[b0 + tid]] = 1;  // as the direct factorisation reports it
    return;
  }
  const int nf = pat_n[p];
  const int64_t ldx = tps_x_ld(nf);
  const double* X = pinv + pxo[p];
  extern __shared__ double dyn[];
  float* xv = reinterpret_cast<float*>(dyn);  // [nm][NV]
  double* yv = dyn + G * nm;                  // [NV][ny], component-major
  const int ny = nm + 4;
  float* val = reinterpret_cast<float*>(dyn + shared_x_head(G, nm));   // [G][2][nm]
  int* fidx = reinterpret_cast<int*>(dyn + shared_x_head(G, nm) + G * nm);  // [G][nm]
  int* fmem = reinterpret_cast<int*>(yv);  // [G][nm], during the checks
  __shared__ double red[3 * NV * (kThreads / 64)];
  __shared__ int rl[G][kCapX];
  __shared__ int nr_sh[G];
  __shared__ int dom[G];
  __shared__ double ga[G][kCapX][3];
  __shared__ double lam[NV][kCapX];
  __shared__ double wa[NV][3];
  __shared__ int go[NV];
  __shared__ int aq[NV];
  __shared__ int na_sh;
  __shared__ int act[G];
  __shared__ int bad[G];
  __shared__ double nfixed[NV], maxdiff[NV];
  __shared__ int count[NV];
  __shared__ int64_t s0_sh[G];
  __shared__ int n_sh[G], box_sh[G][3];
  // the group's subdomains (those of status 0 are checked: the checks of
  // k_tps_shared_columns, every subdomain at once)
  if (tid < G) {
    const int s = tid;
    const int d = s < cnt ? order[b0 + s] : -1;
    const bool ok = d >= 0 && status[d] == 0;
    dom[s] = d;
    act[s] = ok ? 1 : 0;
    bad[s] = 0;
    nr_sh[s] = 0;
    s0_sh[s] = ok ? dom_off[d] : 0;
    n_sh[s] = ok ? int(dom_off[d + 1] - dom_off[d]) : 0;
    box_sh[s][0] = ok ? dbox[4 * d] : 0;
    box_sh[s][1] = ok ? dbox[4 * d + 1] : 0;
    box_sh[s][2] = ok ? dbox[4 * d + 2] : 0;
  }
  for (int e = tid; e < G * nm; e += kThreads) fmem[e] = -1;
  __syncthreads();
  // members on the pattern's nodes: fidx[s][j] (node of member j), fmem[s][k]
  // (member of node k, -1 for a missing node); a center off the grid marks
  // its subdomain bad. Items e = (s, j) over all subdomains, eight a batch
  // so that their dependent loads overlap
  constexpr int kB = 8;
  for (int e0 = 0; e0 < G * nm; e0 += kB * kThreads) {
    int c[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int e = e0 + u * kThreads + tid;
      const int s = e / nm, j = e % nm;
      c[u] = (s < G && j < n_sh[s]) ? slot_pt[s0_sh[s] + j] : -1;
    }
    int q[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) q[u] = c[u] >= 0 ? idx[c[u]] : 0;
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      if (c[u] < 0) continue;
      const int e = e0 + u * kThreads + tid;
      const int s = e / nm, j = e % nm;
      const int j0 = box_sh[s][0], nc = box_sh[s][1], i0 = box_sh[s][2];
      const int row = q[u] / ngx, col = q[u] % ngx;
      const int k = (row - i0) * nc + (col - j0);
      const bool on = col - j0 >= 0 && col - j0 < nc && k >= 0 && k < nf &&
                      xs[c[u]] == gx[col] && ys[c[u]] == gy[row];
      if (!on) {
        bad[s] = 1;
        continue;
      }
      fidx[s * nm + j] = k;
      fmem[s * nm + k] = j;
    }
  }
  __syncthreads();
  // the missing nodes in node order (at most kCapX listed), wavefront s
  if ((tid >> 6) < G && act[tid >> 6] && !bad[tid >> 6]) {
    const int s = tid >> 6, lane = tid & 63;
    int nr = 0;
    for (int k0 = 0; k0 < nf; k0 += 64) {
      const int k = k0 + lane;
      const bool miss = k < nf && fmem[s * nm + k] < 0;
      const unsigned long long m = __ballot(miss);
      const int rank = nr + __popcll(m & ((1ull << lane) - 1ull));
      if (miss && rank < kCapX) rl[s][rank] = k;
      nr += __popcll(m);
    }
    if (lane == 0) nr_sh[s] = nr;
  }
  __syncthreads();
  if (tid < G && act[tid]) {
    const int s = tid, nr = nr_sh[s];
    if (bad[s] || nr > cap_r || nr > kCapX || n_sh[s] + nr != nf) {
      status[dom[s]] = 2;  // solved again by host sizing
      act[s] = 0;
    }
  }
  __syncthreads();
  // values and coordinates by slot
  for (int e0 = 0; e0 < G * nm; e0 += kB * kThreads) {
    int c[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int e = e0 + u * kThreads + tid;
      const int s = e / nm, j = e % nm;
      c[u] = (s < G && act[s] && j < n_sh[s]) ? slot_pt[s0_sh[s] + j] : -1;
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      if (c[u] < 0) continue;
      const int e = e0 + u * kThreads + tid;
      const int s = e / nm, j = e % nm;
      val[2 * s * nm + j] = vx[c[u]];
      val[(2 * s + 1) * nm + j] = vy[c[u]];
      slot_xy[s0_sh[s] + j] = xs[c[u]];
      slot_xy[nslots + s0_sh[s] + j] = ys[c[u]];
    }
  }
  __syncthreads();
  // H = X[R, R] (into the free LDS of x and y, nr x (kCapX + 1) per
  // subdomain), G_a = Y[R, :]
  double* Hs = dyn;
  for (int s = 0; s < G; ++s) {
    if (!act[s]) continue;
    const int nr = nr_sh[s];
    double* Hm = Hs + s * kCapX * (kCapX + 1);
    for (int e = tid; e < nr * nr; e += kThreads)
      Hm[(e / nr) * (kCapX + 1) + e % nr] = X[int64_t(rl[s][e / nr]) * ldx + rl[s][e % nr]];
    for (int e = tid; e < 3 * nr; e += kThreads)
      ga[s][e / 3][e % 3] = X[int64_t(rl[s][e / 3]) * ldx + nf + e % 3];
  }
  __syncthreads();
  // H^-1 by Gauss-Jordan elimination in place (H is symmetric positive
  // definite exactly when every pivot is positive), wavefront s, lanes over
  // rows; then to hinv (rows of kCapX doubles)
  if ((tid >> 6) < G && act[tid >> 6]) {
    const int s = tid >> 6, lane = tid & 63;
    const int nr = nr_sh[s];
    double* Hm = Hs + s * kCapX * (kCapX + 1);
    bool okh = true;
    for (int k = 0; k < nr; ++k) {
      const double piv = Hm[k * (kCapX + 1) + k];
      if (!(piv > 0.0)) {
        okh = false;
        break;
      }
      const double rp = 1.0 / piv;
      double rk = 0.0;
      if (lane < nr) rk = lane == k ? rp : Hm[k * (kCapX + 1) + lane] * rp;
      wave_sync();
      if (lane < nr) Hm[k * (kCapX + 1) + lane] = rk;
      wave_sync();
      // row i = lane % 32, columns j = lane / 32, + 2, ...
      const int i = lane & 31;
      if (i < nr && i != k) {
        double* hi = Hm + i * (kCapX + 1);
        const double f = hi[k];
#pragma unroll 4
        for (int j = lane >> 5; j < nr; j += 2)
          hi[j] = j == k ? -f * rp : hi[j] - f * Hm[k * (kCapX + 1) + j];
      }
      wave_sync();
    }
    if (okh) {
      double* hv = hinv + int64_t(dom[s]) * kCapX * kCapX;
      for (int e = lane; e < nr * kCapX; e += 64) {
        const int i = e / kCapX, j = e % kCapX;
        if (j < nr) hv[e] = Hm[i * (kCapX + 1) + j];
      }
    } else if (lane == 0) {
      status[dom[s]] = 1;  // M_S singular, as the direct factorisation reports it
      act[s] = 0;
    }
  }
  __syncthreads();
  if (tid < NV) {
    go[tid] = act[tid >> 1];
    nfixed[tid] = 0.0;
    maxdiff[tid] = 0.0;
    count[tid] = 0;
  }
  __syncthreads();
  for (int it = 0; it < 11; ++it) {
    bool any = false;
    for (int q = 0; q < NV; ++q) any = any || go[q] != 0;
    if (!any) break;
    // the running components, compacted: column k of x holds component
    // aq[k] (k < na; the columns up to the product's even width are zero)
    if (tid == 0) {
      int k = 0;
      for (int q = 0; q < NV; ++q)
        if (go[q]) aq[k++] = q;
      na_sh = k;
      for (; k < NV; ++k) aq[k] = -1;
    }
    // v' of every running component (0 at the missing nodes)
    for (int e = tid; e < nf * NV; e += kThreads) xv[e] = 0.0f;
    __syncthreads();
    const int na = na_sh;
    for (int k = 0; k < na; ++k) {
      const int q = aq[k], s = q >> 1;
      const int d = order[b0 + s];
      const int n = int(dom_off[d + 1] - dom_off[d]);
      for (int j = tid; j < n; j += kThreads) xv[fidx[s * nm + j] * NV + k] = val[q * nm + j];
    }
    __syncthreads();
    // [y; Y^T v'] = [X | Y]^T v' for the running components only
    if (na <= 2) x_products<2, NV>(X, ldx, nf, xv, yv, ny, aq);
    else if constexpr (NV >= 4) {
      if (na <= 4) x_products<4, NV>(X, ldx, nf, xv, yv, ny, aq);
      else if constexpr (NV >= 8) {
        if (na <= 6) x_products<6, NV>(X, ldx, nf, xv, yv, ny, aq);
        else x_products<8, NV>(X, ldx, nf, xv, yv, ny, aq);
      }
    }
    __syncthreads();
    if (tid < NV)
#pragma unroll
      for (int k = 0; k < 3; ++k) wa[tid][k] = yv[tid * ny + nf + k];  // Y^T v'
    __syncthreads();
    // lambda = -H^-1 y[R]: thread (q, r)
    for (int e = tid; e < NV * kCapX; e += kThreads) {
      const int q = e / kCapX, r = e % kCapX, s = q >> 1;
      const int nr = nr_sh[s];
      if (!go[q] || r >= nr) continue;
      const double* hr = hinv + int64_t(dom[s]) * kCapX * kCapX + r * kCapX;
      double v = 0.0;
      for (int k0 = 0; k0 < nr; k0 += 8) {
        double h[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) h[u] = k0 + u < nr ? hr[k0 + u] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (k0 + u < nr) v += h[u] * yv[q * ny + rl[s][k0 + u]];
      }
      lam[q][r] = -v;
    }
    __syncthreads();
    // a = Y^T v' + G_a lambda
    if (tid < NV && go[tid]) {
      const int q = tid, s = q >> 1;
      const int nr = nr_sh[s];
      for (int t = 0; t < 3; ++t) {
        double v = wa[q][t];
        for (int r = 0; r < nr; ++r) v += ga[s][r][t] * lam[q][r];
        wa[q][t] = v;
      }
    }
    __syncthreads();
    // w = y + X[:, R] lambda (items (s, k) over every subdomain at once)
    for (int e = tid; e < G * nf; e += kThreads) {
      const int s = e / nf, k = e - s * nf;
      if (!go[2 * s] && !go[2 * s + 1]) continue;
      const int nr = nr_sh[s];
      {
        double w0 = yv[2 * s * ny + k], w1 = yv[(2 * s + 1) * ny + k];
        for (int r0 = 0; r0 < nr; r0 += 8) {
          double g[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) g[u] = r0 + u < nr ? X[int64_t(rl[s][r0 + u]) * ldx + k] : 0.0;
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (r0 + u < nr) {
              w0 += g[u] * lam[2 * s][r0 + u];
              w1 += g[u] * lam[2 * s + 1][r0 + u];
            }
        }
        yv[2 * s * ny + k] = w0;
        yv[(2 * s + 1) * ny + k] = w1;
      }
    }
    __syncthreads();
    // smoothed values, weights and outliers of every running component:
    // per-wavefront maxima and counts, combined by thread 0
    for (int q = 0; q < NV; ++q) {
      if (!go[q]) continue;
      const int s = q >> 1, c = q & 1;
      const int d = order[b0 + s];
      const int64_t s0 = dom_off[d];
      const int n = int(dom_off[d + 1] - s0);
      const double smoothing = smoothing_d[d];
      double* W = weights + c * wstride + s0 + 3 * int64_t(d);
      double dm = 0.0, ne = 0.0;
      for (int j = tid; j < n; j += kThreads) {
        const double v = double(val[q * nm + j]);
        const double wq = yv[q * ny + fidx[s * nm + j]];
        const double sm = v - smoothing * wq;
        smooth_slots[c * nslots + s0 + j] = sm;
        const double diff = fabs(sm - v);
        dm = fmax(dm, diff);
        if (threshold >= 0.0 && diff > threshold) ne += 1.0;
        W[j] = wq;
      }
      for (int off = 32; off > 0; off >>= 1) {
        dm = fmax(dm, __shfl_xor(dm, off, 64));
        ne += __shfl_xor(ne, off, 64);
      }
      if ((tid & 63) == 0) {
        red[q * (kThreads / 64) + (tid >> 6)] = dm;
        red[(NV + q) * (kThreads / 64) + (tid >> 6)] = ne;
      }
      if (tid == 0) {
        // P = [1, y - y0, x - x0] -> [1, y, x]
        const double x0 = gx[dbox[4 * d]], y0 = gy[dbox[4 * d + 2]];
        W[n] = wa[q][0] - wa[q][1] * y0 - wa[q][2] * x0;
        W[n + 1] = wa[q][1];
        W[n + 2] = wa[q][2];
      }
    }
    __syncthreads();
    // go[q] becomes whether component q iterates again
    if (tid == 0)
      for (int q = 0; q < NV; ++q) {
        if (!go[q]) continue;
        double dm = red[q * (kThreads / 64)], ne = red[(NV + q) * (kThreads / 64)];
        for (int k = 1; k < kThreads / 64; ++k) {
          dm = fmax(dm, red[q * (kThreads / 64) + k]);
          ne += red[(NV + q) * (kThreads / 64) + k];
        }
        count[q] += 1;
        if (count[q] == 1) maxdiff[q] = dm;
        const bool more = threshold >= 0.0 && ne > 0.0 && count[q] < 11;
        if (more) nfixed[q] += ne;
        go[q] = more ? 1 : 0;
      }
    __syncthreads();
    // the outliers take their smoothed values
    for (int q = 0; q < NV; ++q) {
      if (!go[q]) continue;
      const int s = q >> 1;
      const int d = order[b0 + s];
      const int n = int(dom_off[d + 1] - dom_off[d]);
      const double smoothing = smoothing_d[d];
      for (int j = tid; j < n; j += kThreads) {
        const double v = double(val[q * nm + j]);
        const double sm = v - smoothing * yv[q * ny + fidx[s * nm + j]];
        if (fabs(sm - v) > threshold) val[q * nm + j] = float(sm);
      }
    }
  }
  if (tid == 0) {
#pragma unroll
    for (int s = 0; s < G; ++s) {
      if (!act[s]) continue;
      const int d = order[b0 + s];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        stats[6 * d + 3 * c + 0] = nfixed[2 * s + c];
        stats[6 * d + 3 * c + 1] = maxdiff[2 * s + c];
        stats[6 * d + 3 * c + 2] = count[2 * s + c];
      }
    }
  }
  __syncthreads();  // LDS reused by the next group
}

// the groups (groups[nd] of them), XCD-contiguous ranges of them per
// workgroup stride: a fixed grid instead of one workgroup per subdomain slot
template <int G>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_tps_shared_x(
    int nd, double* __restrict__ hinv, const int32_t* __restrict__ order,
    const int32_t* __restrict__ groups,
    const int32_t* __restrict__ dpat, const int32_t* __restrict__ dbox,
    const int32_t* __restrict__ pat_n, const int32_t* __restrict__ pstatus,
    const int64_t* __restrict__ pxo, const double* __restrict__ pinv,
    const double* __restrict__ gx, const double* __restrict__ gy, int ngx,
    const double* __restrict__ xs, const double* __restrict__ ys,
    const int32_t* __restrict__ idx, const float* __restrict__ vx,
    const float* __restrict__ vy, const int64_t* __restrict__ dom_off,
    const int32_t* __restrict__ slot_pt, const double* __restrict__ smoothing_d,
    double threshold, double* __restrict__ smooth_slots, int64_t nslots,
    double* __restrict__ weights, int64_t wstride, double* __restrict__ stats,
    int32_t* __restrict__ status, double* __restrict__ slot_xy, int cap_r, int nm) {
  const int ng = groups[nd];
  const int ngrid = int(gridDim.x) & ~7;  // a multiple of the XCD count
  if (int(blockIdx.x) >= ngrid) return;
  for (int w = blockIdx.x; w < ng; w += ngrid)
    shared_x_solve<G>(xcd_rank(w, ng), hinv, order, groups, dpat, dbox, pat_n, pstatus, pxo, pinv, gx,
                      gy, ngx, xs, ys, idx, vx, vy, dom_off, slot_pt, smoothing_d, threshold,
                      smooth_slots, nslots, weights, wstride, stats, status, slot_xy, cap_r, nm,
                      nd);
}

// ---------------------------------------------------------------------------
// 4d. the smoothing as a linear map of the values (differentiable node field)
// ---------------------------------------------------------------------------
// Without the outlier iterations, the smoothed values of a subdomain are
// v - s w with [w; a] = M^-1 [v; 0] (thin_plate_spline.py:168-216), i.e.
// (I - s Sw) v for the influence matrix Sw = (M^-1)[:n, :n]. Sw is built
// once per subdomain (column j: the w part of the solution for [e_j; 0],
// eight columns per pass of the multi-right-hand-side solve); the blend of
// the subdomains' smoothed values at the centers is the reference's
// (thin_plate_spline_subdom.py:145-158). The map and its transpose then cost
// one matrix-vector product per subdomain and field.

// Sw of subdomain blockIdx.x, row-major n x n (ld n) at infl + d stride
__global__ __launch_bounds__(kThreads) void k_tps_influence(
    const int64_t* __restrict__ dom_off, const int64_t* __restrict__ mat_off,
    const double* __restrict__ mats, const double* __restrict__ refl,
    const double* __restrict__ vwork, const int32_t* __restrict__ status, int nm,
    double* __restrict__ infl, int64_t stride) {
  const int d = blockIdx.x;
  if (status[d]) return;
  const int64_t s0 = dom_off[d];
  const int n = int(dom_off[d + 1] - s0);
  if (n == 0) return;
  const int m = n - 3;
  const double* E = mats + mat_off[d];
  const int64_t ld = tps_mat_ld(m);
  const double* A = E + tps_mat_e_size(m);
  const double* RF = refl + 16 * d;
  const double* VW = vwork + 6 * s0;
  const double tau[3] = {RF[0], RF[1], RF[2]};
  const int tid = threadIdx.x;
  // dynamic LDS: x[8][nm], then the solve's scratch
  extern __shared__ double dyn[];
  double* x = dyn;
  double* scr = dyn + 8 * nm;
  __shared__ double red[8 * (kThreads / 64)];
  __shared__ double aout[8][3];
  double* S = infl + int64_t(d) * stride;
  for (int j0 = 0; j0 < n; j0 += 8) {
    for (int k = tid; k < n; k += kThreads)
#pragma unroll
      for (int q = 0; q < 8; ++q) x[q * nm + k] = k == j0 + q ? 1.0 : 0.0;
    __syncthreads();
    saddle_solve_n<8>(x, nm, n, VW, tau, A, ld, E, RF, red, scr, aout);
    for (int k = tid; k < n; k += kThreads)
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (j0 + q < n) S[int64_t(k) * n + j0 + q] = x[q * nm + k];
    __syncthreads();
  }
}

// out[c][f nslots + s0 + j] = in_j - s (Sw in)_j (transpose 0) or
// in_j - s (Sw^T in)_j (transpose 1), in[c][f nslots + s0 + j]; subdomain
// blockIdx.y, field blockIdx.x (consecutive workgroups read one Sw).
// cstride: offset of component 1. Sw rows are read by wavefronts (lanes
// over k), its columns by threads (thread j, loop over k).
__global__ __launch_bounds__(kThreads) void k_tps_node_apply(
    const int64_t* __restrict__ dom_off, const int32_t* __restrict__ status,
    const double* __restrict__ smoothing_d, const double* __restrict__ infl, int64_t stride,
    const double* __restrict__ in, double* __restrict__ out, int64_t nslots, int64_t cstride,
    int transpose) {
  const int d = blockIdx.y;
  const int f = blockIdx.x;
  if (status[d]) return;
  const int64_t s0 = dom_off[d];
  const int n = int(dom_off[d + 1] - s0);
  if (n == 0) return;
  const double sm = smoothing_d[d];
  const double* S = infl + int64_t(d) * stride;
  const int64_t base = int64_t(f) * nslots + s0;
  extern __shared__ double dyn[];
  double* xv[2] = {dyn, dyn + n};
  const int tid = threadIdx.x;
  for (int j = tid; j < n; j += kThreads) {
    xv[0][j] = in[base + j];
    xv[1][j] = in[cstride + base + j];
  }
  __syncthreads();
  if (transpose) {
    for (int j = tid; j < n; j += kThreads) {
      double a0 = 0.0, a1 = 0.0;
      for (int k = 0; k < n; ++k) {
        const double w = S[int64_t(k) * n + j];
        a0 += w * xv[0][k];
        a1 += w * xv[1][k];
      }
      out[base + j] = xv[0][j] - sm * a0;
      out[cstride + base + j] = xv[1][j] - sm * a1;
    }
  } else {
    const int lane = tid & 63, wave = tid >> 6;
    for (int j = wave; j < n; j += kThreads / 64) {
      const double* row = S + int64_t(j) * n;
      double a0 = 0.0, a1 = 0.0;
      for (int k = lane; k < n; k += 64) {
        const double w = row[k];
        a0 += w * xv[0][k];
        a1 += w * xv[1][k];
      }
      for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off, 64);
        a1 += __shfl_xor(a1, off, 64);
      }
      if (lane == 0) {
        out[base + j] = xv[0][j] - sm * a0;
        out[cstride + base + j] = xv[1][j] - sm * a1;
      }
    }
  }
}

// The per-center steps of the node field, fields f = blockIdx.y, centers i:
// mode 0 (gather): slots[c][f nslots + s] = v[c][f nc + slot_pt[s]];
// mode 1 (blend): out[c][f nc + i] = sum_k coef[4 i + k] slots[c][f nslots
//   + slot_k] / norm[i] (the order of k_tps_blend);
// mode 2 (blend transposed): slots[c][f nslots + slot_k] = coef[4 i + k] /
//   norm[i] g[c][f nc + i];
// mode 3 (gather transposed): out[c][f nc + i] = sum_k slots[c][f nslots +
//   slot_k].
// Modes 0 run over slots s < nused, the others over centers i < nc.
__global__ void k_tps_node_steps(int mode, int nc, int64_t nused, const int32_t* __restrict__ slot_pt,
                                 const int32_t* __restrict__ c_slot,
                                 const double* __restrict__ c_coef,
                                 const double* __restrict__ c_norm, int64_t nslots,
                                 int64_t scstride, int64_t ncstride,
                                 const double* __restrict__ src, double* __restrict__ dst) {
  const int f = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (mode == 0) {
    if (i >= nused) return;
    const int64_t q = int64_t(f) * nc + slot_pt[i];
    dst[int64_t(f) * nslots + i] = src[q];
    dst[scstride + int64_t(f) * nslots + i] = src[ncstride + q];
    return;
  }
  if (i >= nc) return;
  const int64_t pi = int64_t(f) * nc + i;
  const int64_t fs = int64_t(f) * nslots;
  if (mode == 1) {
    double a0 = 0.0, a1 = 0.0;
    for (int k = 0; k < kMaxSlots; ++k) {
      const int sl = c_slot[kMaxSlots * i + k];
      if (sl < 0) break;
      const double c = c_coef[kMaxSlots * i + k];
      a0 += c * src[fs + sl];
      a1 += c * src[scstride + fs + sl];
    }
    dst[pi] = a0 / c_norm[i];
    dst[ncstride + pi] = a1 / c_norm[i];
  } else if (mode == 2) {
    const double g0 = src[pi] / c_norm[i], g1 = src[ncstride + pi] / c_norm[i];
    for (int k = 0; k < kMaxSlots; ++k) {
      const int sl = c_slot[kMaxSlots * i + k];
      if (sl < 0) break;
      const double c = c_coef[kMaxSlots * i + k];
      dst[fs + sl] = c * g0;
      dst[scstride + fs + sl] = c * g1;
    }
  } else {
    double a0 = 0.0, a1 = 0.0;
    for (int k = 0; k < kMaxSlots; ++k) {
      const int sl = c_slot[kMaxSlots * i + k];
      if (sl < 0) break;
      a0 += src[fs + sl];
      a1 += src[scstride + fs + sl];
    }
    dst[pi] = a0;
    dst[ncstride + pi] = a1;
  }
}

// ---------------------------------------------------------------------------
// 5. blend of the smoothed values, 6. evaluation at new positions
// ---------------------------------------------------------------------------
__global__ void k_tps_blend(int npts, const int32_t* __restrict__ slots,
                            const double* __restrict__ coef, const double* __restrict__ norm,
                            const double* __restrict__ sm, int64_t nslots,
                            double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  double a0 = 0.0, a1 = 0.0;
  for (int k = 0; k < kMaxSlots; ++k) {
    const int s = slots[kMaxSlots * i + k];
    if (s < 0) break;
    const double c = coef[kMaxSlots * i + k];
    a0 += c * sm[s];
    a1 += c * sm[nslots + s];
  }
  out[i] = a0 / norm[i];
  out[npts + i] = a1 / norm[i];
}

// Evaluation by subdomain: the (point, slot) pairs e = 4 p + k of the
// points to evaluate are sorted by their subdomain (count, exclusive scan,
// fill: a counting sort), then one workgroup per subdomain holds its
// centers and weights in LDS and evaluates its points, one thread per
// point, the partial values going to partial[c][e]; the last kernel blends
// the partials of each point in slot order. Workspace (int32 then double,
// offsets in piv_tps_interpolate_workspace): cnt[nd + 1], cursor[nd],
// list[4 npts], partial[2][4 npts].
__global__ void k_tps_list_zero(int nd, int32_t* __restrict__ cnt,
                                int32_t* __restrict__ cursor) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= nd) cnt[i] = 0;
  if (i < nd) cursor[i] = 0;
}

// ctr[key] += (number of pending lanes with this key) for every key of the
// wavefront's pending lanes, one atomic per distinct key; each pending lane
// gets its own position in [old ctr[key], new ctr[key]) (-1 if not pending).
// Called by every lane of the wavefront that reached it.
__device__ __forceinline__ int wave_aggregated_add(int32_t* ctr, int key, bool pending) {
  int pos = -1;
  const int lane = __lane_id();
  for (;;) {
    const unsigned long long act = __ballot(pending);
    if (act == 0ull) break;
    const int leader = __ffsll(act) - 1;
    const int lk = __shfl(key, leader);
    const bool mine = pending && key == lk;
    const unsigned long long same = __ballot(mine);
    int base = 0;
    if (lane == leader) base = atomicAdd(&ctr[lk], int(__popcll(same)));
    base = __shfl(base, leader);
    if (mine) {
      pos = base + int(__popcll(same & ((1ull << lane) - 1ull)));
      pending = false;
    }
  }
  return pos;
}

// one thread per point, its slots in turn (neighbouring points share their
// subdomains, so a wavefront issues few atomics)
__global__ void k_tps_list_count(int npts, const int32_t* __restrict__ doms,
                                 const uint8_t* __restrict__ skip, int32_t* __restrict__ cnt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = p < npts && !(skip != nullptr && skip[p < npts ? p : 0]);
  for (int k = 0; k < kMaxSlots; ++k) {
    const int d = live ? doms[int64_t(kMaxSlots) * p + k] : -1;
    wave_aggregated_add(cnt, d, d >= 0);
  }
}

// exclusive prefix sum of cnt[0..nd) in place (cnt[nd] = total), one block
__global__ __launch_bounds__(1024) void k_tps_list_scan(int nd, int32_t* __restrict__ cnt) {
  __shared__ int tmp[32];
  int carry = 0;
  for (int b = 0; b < nd; b += 1024) {
    const int i = b + int(threadIdx.x);
    const int v = i < nd ? cnt[i] : 0;
    int total;
    const int inc = block_scan_1024(v, tmp, total);
    if (i < nd) cnt[i] = carry + inc - v;
    carry += total;
  }
  if (threadIdx.x == 0) cnt[nd] = carry;
}

__global__ void k_tps_list_fill(int npts, const int32_t* __restrict__ doms,
                                const uint8_t* __restrict__ skip,
                                const int32_t* __restrict__ off, int32_t* __restrict__ cursor,
                                int32_t* __restrict__ list) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = p < npts && !(skip != nullptr && skip[p < npts ? p : 0]);
  for (int k = 0; k < kMaxSlots; ++k) {
    const int64_t e = int64_t(kMaxSlots) * p + k;
    const int d = live ? doms[e] : -1;
    const int pos = wave_aggregated_add(cursor, d, d >= 0);
    if (pos >= 0) list[off[d] + pos] = int32_t(e);
  }
}

// sum_j w_j phi(|p - c_j|) plus the linear part, for the points listed
// for subdomain blockIdx.x (the list split over gridDim.y workgroups);
// centers and weights in LDS (x, y, w0, w1 per center), one point per
// thread, the centers in order
__global__ __launch_bounds__(256) void k_tps_eval_dom(
    int npts, int nmax, const int32_t* __restrict__ off, const int32_t* __restrict__ list,
    const double* __restrict__ px, const double* __restrict__ py,
    const double* __restrict__ slot_xy, int64_t nslots, const int64_t* __restrict__ dom_off,
    const double* __restrict__ weights, int64_t wstride, double* __restrict__ partial,
    PhiLattice lat) {
  const int d = blockIdx.x;
  const int e0 = off[d], e1 = off[d + 1];
  const int first = e0 + int(blockIdx.y * blockDim.x);
  const int step = int(gridDim.y * blockDim.x);
  if (first >= e1) return;
  const int64_t s0 = dom_off[d];
  const int n = int(dom_off[d + 1] - s0);
  const int64_t stride = int64_t(kMaxSlots) * npts;
  if (n > nmax) {
    // more centers than the caller's bound (the LDS holds nmax): NaN values
    for (int t = first + int(threadIdx.x); t < e1; t += step) {
      partial[list[t]] = __builtin_nan("");
      partial[stride + list[t]] = __builtin_nan("");
    }
    return;
  }
  const double* W0 = weights + s0 + 3 * int64_t(d);
  const double* W1 = W0 + wstride;
  extern __shared__ double2 cen[];  // (x, y) then (w0, w1) of center j at 2 j, 2 j + 1
  __shared__ double box[4][kThreads / 64];  // bounding box of the centers, by wavefront
  const double* ltab = stage_lattice(lat, reinterpret_cast<double*>(cen + 2 * nmax));
  // with a table: are all centers on its lattice, and where are they
  bool on = ltab != nullptr;
  double bx0 = INFINITY, bx1 = -INFINITY, by0 = INFINITY, by1 = -INFINITY;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const double cx = slot_xy[s0 + j], cy = slot_xy[nslots + s0 + j];
    cen[2 * j] = make_double2(cx, cy);
    cen[2 * j + 1] = make_double2(W0[j], W1[j]);
    on = on && on_lattice(cx, lat) && on_lattice(cy, lat);
    bx0 = fmin(bx0, cx);
    bx1 = fmax(bx1, cx);
    by0 = fmin(by0, cy);
    by1 = fmax(by1, cy);
  }
  for (int off = 32; off > 0; off >>= 1) {
    bx0 = fmin(bx0, __shfl_xor(bx0, off, 64));
    bx1 = fmax(bx1, __shfl_xor(bx1, off, 64));
    by0 = fmin(by0, __shfl_xor(by0, off, 64));
    by1 = fmax(by1, __shfl_xor(by1, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    box[0][threadIdx.x >> 6] = bx0;
    box[1][threadIdx.x >> 6] = bx1;
    box[2][threadIdx.x >> 6] = by0;
    box[3][threadIdx.x >> 6] = by1;
  }
  const double a0 = W0[n], a1 = W0[n + 1], a2 = W0[n + 2];
  const double b0 = W1[n], b1 = W1[n + 1], b2 = W1[n + 2];
  const bool all_on = __syncthreads_and(on);
  for (int k = 0; k < kThreads / 64; ++k) {
    bx0 = fmin(bx0, box[0][k]);
    bx1 = fmax(bx1, box[1][k]);
    by0 = fmin(by0, box[2][k]);
    by1 = fmax(by1, box[3][k]);
  }
  const double reach = all_on ? double(lat.a_max) * lat.step : -1.0;
  for (int t = first + int(threadIdx.x); t < e1; t += step) {
    const int e = list[t];
    const int p = e / kMaxSlots;
    const double x = px[p], y = py[p];
    double v0 = 0.0, v1 = 0.0;
    // a point on the lattice within the table's reach of every center: the
    // table entry of each difference without the check (the entry the check
    // would find), else the checked lookup
    if (all_on && on_lattice(x, lat) && on_lattice(y, lat) && x - bx0 <= reach &&
        bx1 - x <= reach && y - by0 <= reach && by1 - y <= reach) {
      for (int j = 0; j < n; ++j) {
        const double2 c = cen[2 * j];
        const double2 w = cen[2 * j + 1];
        const double ph = tps_phi_on_lattice(x - c.x, y - c.y, lat, ltab);
        v0 += w.x * ph;
        v1 += w.y * ph;
      }
    } else {
      for (int j = 0; j < n; ++j) {
        const double2 c = cen[2 * j];
        const double2 w = cen[2 * j + 1];
        const double ph = tps_phi_at(x - c.x, y - c.y, lat, ltab);
        v0 += w.x * ph;
        v1 += w.y * ph;
      }
    }
    v0 += a0 + a1 * y + a2 * x;
    v1 += b0 + b1 * y + b2 * x;
    partial[e] = v0;
    partial[stride + e] = v1;
  }
}

__global__ void k_tps_blend_partials(int npts, const int32_t* __restrict__ doms,
                                     const double* __restrict__ coef,
                                     const double* __restrict__ norm,
                                     const uint8_t* __restrict__ skip,
                                     const double* __restrict__ partial,
                                     double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts || (skip != nullptr && skip[p])) return;
  const int64_t stride = int64_t(kMaxSlots) * npts;
  double acc0 = 0.0, acc1 = 0.0;
  for (int k = 0; k < kMaxSlots; ++k) {
    const int64_t e = int64_t(kMaxSlots) * p + k;
    if (doms[e] < 0) break;
    const double cf = coef[e];
    acc0 += cf * partial[e];
    acc1 += cf * partial[stride + e];
  }
  out[p] = acc0 / norm[p];
  out[npts + p] = acc1 / norm[p];
}

// workspace layout of piv_tps_interpolate
struct InterpWork {
  int32_t* cnt;
  int32_t* cursor;
  int32_t* list;
  double* partial;
  size_t bytes;
};

InterpWork interp_work(void* base, int npts, int nd) {
  InterpWork w;
  char* b = static_cast<char*>(base);
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t at = o;
    o += (n + 255) & ~size_t(255);
    return b + at;
  };
  w.cnt = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * (size_t(nd) + 1)));
  w.cursor = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * size_t(nd)));
  w.list = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * size_t(kMaxSlots) * npts));
  w.partial = reinterpret_cast<double*>(take(sizeof(double) * 2 * size_t(kMaxSlots) * npts));
  w.bytes = o;
  return w;
}

std::atomic<int> g_chol_threads{0};
std::atomic<int> g_shared_solve{0};

// compute units of the current device (cached per device)
int cu_count() {
  static std::atomic<int> cached[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int v = cached[dev].load(std::memory_order_relaxed);
  if (v == 0) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v <= 0)
      v = 256;
    cached[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

int tps_fail(int code, const std::string& msg) {
  fia::last_error() = msg;
  return code;
}

int tps_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tps_fail(PIV_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return PIV_OK;
}

// Scratch of piv_tps_interpolate_noalloc: the evaluation's point lists and
// partial sums (the workspace of piv_tps_interpolate) in a fixed device array,
// one copy per GPU, held until the process ends. Host code never passes the
// array's name to a kernel (on the host it is only a stand-in): the address of
// the current device's copy comes from hipGetSymbolAddress, once per device.
constexpr size_t kNoAllocScratchBytes = size_t(PIV_TPS_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Address of the scratch array on the current device (looked up on the first
// call on that device and kept).
int scratch_on_current_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return tps_fail(PIV_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return tps_fail(PIV_EUNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return tps_fail(PIV_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                    hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return PIV_OK;
}

int tps_interpolate_impl(const char* name, int32_t npts, const double* px, const double* py,
                         const int32_t* doms,
                        const double* coef, const double* norm, const double* slot_xy,
                        int64_t nslots, int32_t nd, int32_t n_max, const int64_t* dom_off,
                        const double* weights, int64_t wstride, const uint8_t* skip,
                        double* out, void* work, const piv_tps_lattice* lattice,
                        void* stream) {
  if (npts <= 0 || nd <= 0) return PIV_OK;
  if (n_max <= 0 || n_max > kMaxN)
    return tps_fail(PIV_EINVAL, "n_max of piv_tps_interpolate must be in [1, 2048]");
  if (int64_t(npts) * kMaxSlots > INT32_MAX)
    return tps_fail(PIV_EUNSUPPORTED, "too many TPS evaluation points");
  if (!px || !py || !doms || !coef || !norm || !slot_xy || !dom_off || !weights || !out ||
      !work)
    return tps_fail(PIV_EINVAL, std::string("NULL argument of ") + name);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const InterpWork w = interp_work(work, npts, nd);
  const unsigned ge = unsigned((npts + 255) / 256);
  hipLaunchKernelGGL(k_tps_list_zero, dim3((nd + 256) / 256), dim3(256), 0, s, int(nd), w.cnt,
                     w.cursor);
  hipLaunchKernelGGL(k_tps_list_count, dim3(ge), dim3(256), 0, s, int(npts), doms, skip, w.cnt);
  hipLaunchKernelGGL(k_tps_list_scan, dim3(1), dim3(1024), 0, s, int(nd), w.cnt);
  hipLaunchKernelGGL(k_tps_list_fill, dim3(ge), dim3(256), 0, s, int(npts), doms, skip, w.cnt,
                     w.cursor, w.list);
  // about 8 workgroups per CU in all: few subdomains split their lists
  const int splits = std::min(64, std::max(1, (8 * cu_count() + int(nd) - 1) / int(nd)));
  const PhiLattice lat = make_lattice(lattice);
  hipLaunchKernelGGL(k_tps_eval_dom, dim3(nd, splits), dim3(256),
                     size_t(2) * n_max * sizeof(double2) + size_t(lat.lds_words) * sizeof(double),
                     s, int(npts), int(n_max), w.cnt, w.list, px, py, slot_xy, nslots, dom_off, weights,
                     wstride, w.partial, lat);
  hipLaunchKernelGGL(k_tps_blend_partials, dim3((npts + 255) / 256), dim3(256), 0, s, int(npts),
                     doms, coef, norm, skip, w.partial, out);
  return tps_check(name);
}

// matrix build and factorisation of nd subdomains (prep + Cholesky); grid:
// the workgroups of each launch, striding over the subdomains (0: nd, one
// each; fewer where an order lists few of them, the rest -1)
void launch_factor(int32_t nd, int32_t n_max, const double* xs, const double* ys,
                   const int64_t* dom_off, const int32_t* slot_pt, const int64_t* mat_off,
                   const double* smoothing, double* mats, double* refl, double* vwork,
                   double* slot_xy, int64_t nslots, int32_t* status, const int32_t* order,
                   const PhiLattice& lat, bool gen, hipStream_t s, int grid = 0) {
  const dim3 g(unsigned(grid > 0 ? std::min(grid, int(nd)) : nd));
  const bool stride = grid > 0;
  // up to one subdomain per CU: 16 wavefronts per matrix build, else 4
  const size_t prep_lds = size_t(8) * n_max * sizeof(double) + size_t(lat.lds_words) * sizeof(double);
  const bool wide = nd <= cu_count();
#define PIV_PREP(NT, GEN, ST)                                                                \
  hipLaunchKernelGGL((k_tps_prep<NT, GEN, ST>), g, dim3(NT), prep_lds, s, xs, ys, dom_off,  \
                     slot_pt, mat_off, smoothing, mats, refl, vwork, slot_xy, nslots, status,   \
                     int(n_max), lat, order, int(nd))
  if (stride) {
    if (wide && gen) PIV_PREP(1024, true, true);
    else if (wide) PIV_PREP(1024, false, true);
    else if (gen) PIV_PREP(kThreads, true, true);
    else PIV_PREP(kThreads, false, true);
  } else {
    if (wide && gen) PIV_PREP(1024, true, false);
    else if (wide) PIV_PREP(1024, false, false);
    else if (gen) PIV_PREP(kThreads, true, false);
    else PIV_PREP(kThreads, false, false);
  }
#undef PIV_PREP
  // 8 wavefronts per subdomain (one workgroup per CU) or 4 (two per CU).
  // Up to two subdomains per CU the launch is one round either way if the
  // width fits: 8 wavefronts up to one subdomain per CU (and for the strided
  // launches, over orders that list few subdomains), 4 up to two; beyond,
  // the tuned width (default 4)
  const int cus = cu_count();
  int ct = fia::tps_chol_threads();
  if (nd <= cus || stride) ct = 512;
  else if (nd <= 2 * cus) ct = 256;
  else if (ct == 0) ct = 256;
  const size_t chol_lds = size_t(lat.lds_words) * sizeof(double);
#define PIV_CHOL(NT, ST)                                                                       \
  hipLaunchKernelGGL((k_tps_chol<NT, ST>), g, dim3(NT), chol_lds, s, dom_off, mat_off, mats,  \
                     vwork, status, slot_xy, nslots, smoothing, lat, gen, order, int(nd))
  if (ct == 512 && stride) PIV_CHOL(512, true);
  else if (ct == 512) PIV_CHOL(512, false);
  else if (stride) PIV_CHOL(256, true);
  else PIV_CHOL(256, false);
#undef PIV_CHOL
}

void launch_solve(int32_t nd, int32_t n_max, const float* vx, const float* vy,
                  const int64_t* dom_off, const int32_t* slot_pt, const int64_t* mat_off,
                  const double* mats, const double* refl, const double* vwork,
                  const double* smoothing, double threshold, double* smooth_slots,
                  int64_t nslots, double* weights, int64_t wstride, double* stats,
                  const int32_t* status, const int32_t* order, hipStream_t s, int grid = 0) {
  const dim3 g(unsigned(grid > 0 ? std::min(grid, int(nd)) : nd));
  const size_t lds = size_t(3) * ((n_max + 1) & ~1) * sizeof(double);
  if (grid > 0)
    hipLaunchKernelGGL(k_tps_solve<true>, g, dim3(kThreads), lds, s, vx, vy, dom_off, slot_pt,
                       mat_off, mats, refl, vwork, smoothing, threshold, smooth_slots, nslots,
                       weights, wstride, stats, status, (n_max + 1) & ~1, order, int(nd));
  else
    hipLaunchKernelGGL(k_tps_solve<false>, g, dim3(kThreads), lds, s, vx, vy, dom_off, slot_pt,
                       mat_off, mats, refl, vwork, smoothing, threshold, smooth_slots, nslots,
                       weights, wstride, stats, status, (n_max + 1) & ~1, order, int(nd));
}

// workspace of piv_tps_smooth_shared (offsets in piv_tps_shared_workspace)
struct SharedWork {
  int32_t *dmode, *dpat, *dbox, *pat_n, *pat_box, *pslot, *pstatus, *order_direct, *order_shared,
      *order_sorted, *groups, *plist;
  int64_t *pdom_off, *pmat_off, *pxo;
  double *pat_smooth, *pxy, *pmats, *pinv, *pac, *prefl, *pvwork, *pslot_xy, *gmat, *hfac, *gpoly;
  int64_t pxy_stride;
  size_t bytes;
};

SharedWork shared_work(void* base, int nfields, int dom_cap, int npat_cap, int n_max, int cap_r) {
  SharedWork w;
  char* b = static_cast<char*>(base);
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t at = o;
    o += (n + 255) & ~size_t(255);
    return b + at;
  };
  const size_t nd = size_t(nfields) * dom_cap, np = size_t(nfields) * npat_cap;
  const size_t nm = size_t((n_max + 1) & ~1);
  const size_t pts = np * size_t(n_max);  // pattern centers, at most n_max each
  w.dmode = reinterpret_cast<int32_t*>(take(4 * nd));
  w.dpat = reinterpret_cast<int32_t*>(take(4 * nd));
  w.dbox = reinterpret_cast<int32_t*>(take(16 * nd));
  w.order_direct = reinterpret_cast<int32_t*>(take(4 * nd));
  w.order_shared = reinterpret_cast<int32_t*>(take(4 * (nd + 1)));  // + their count
  w.order_sorted = reinterpret_cast<int32_t*>(take(4 * (nd + 1)));
  w.groups = reinterpret_cast<int32_t*>(take(4 * 2 * (nd + 1)));
  w.pat_n = reinterpret_cast<int32_t*>(take(4 * np));
  w.pat_box = reinterpret_cast<int32_t*>(take(16 * np));
  w.pstatus = reinterpret_cast<int32_t*>(take(4 * np));
  w.pslot = reinterpret_cast<int32_t*>(take(4 * pts));
  w.pdom_off = reinterpret_cast<int64_t*>(take(8 * (np + 1)));
  w.pmat_off = reinterpret_cast<int64_t*>(take(8 * (np + 1)));
  w.pxo = reinterpret_cast<int64_t*>(take(8 * (np + 1)));
  w.plist = reinterpret_cast<int32_t*>(take(4 * (np + 1)));
  w.pat_smooth = reinterpret_cast<double*>(take(8 * np));
  w.pxy_stride = int64_t(pts);
  w.pxy = reinterpret_cast<double*>(take(8 * 2 * pts));
  w.pslot_xy = reinterpret_cast<double*>(take(8 * 2 * pts));
  w.prefl = reinterpret_cast<double*>(take(8 * 16 * np));
  w.pvwork = reinterpret_cast<double*>(take(8 * 6 * pts));
  w.pmats = reinterpret_cast<double*>(take(8 * np * size_t(tps_mat_size(n_max))));
  // the patterns' Linv (matrix layout), then their X and Y^T
  w.pinv = reinterpret_cast<double*>(
      take(8 * np * size_t(std::max(tps_mat_size(n_max), tps_x_size(n_max)))));
  w.pac = reinterpret_cast<double*>(take(8 * 6 * pts));
  // capacitance columns of the substitution path (at most kCapMaxR), the
  // factors of H (kCapMaxR^2) or H^-1 (kCapX^2) per subdomain
  w.gmat = reinterpret_cast<double*>(
      take(8 * nd * nm * size_t(std::max(1, std::min(cap_r, kCapMaxR)))));
  w.hfac = reinterpret_cast<double*>(take(8 * nd * kCapX * kCapX));
  w.gpoly = reinterpret_cast<double*>(take(8 * nd * kCapMaxR * 3));
  w.bytes = o;
  return w;
}

__global__ void k_tps_zero_i32(int n, int32_t* a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

}  // namespace

namespace fia {
int tps_chol_threads() { return g_chol_threads.load(std::memory_order_relaxed); }
void set_tps_chol_threads(int threads) { g_chol_threads.store(threads, std::memory_order_relaxed); }
int tps_shared_solve() { return g_shared_solve.load(std::memory_order_relaxed); }
void set_tps_shared_solve(int mode) { g_shared_solve.store(mode, std::memory_order_relaxed); }
}  // namespace fia

extern "C" {

int piv_tps_smooth(int32_t nd, int32_t n_max, const double* xs, const double* ys,
                   const float* vx, const float* vy, const int64_t* dom_off,
                   const int32_t* slot_pt, const int64_t* mat_off, const double* smoothing,
                   double threshold, double* mats, double* refl, double* vwork, double* slot_xy,
                   double* smooth_slots, int64_t nslots, double* weights, int64_t wstride,
                   double* stats, int32_t* status, const int32_t* order,
                   const piv_tps_lattice* lattice, int32_t generated, void* stream) {
  if (nd <= 0) return PIV_OK;
  if (n_max > kMaxN)
    return tps_fail(PIV_EUNSUPPORTED, "more than 2048 centers in a TPS subdomain");
  if (!xs || !ys || !vx || !vy || !dom_off || !slot_pt || !mat_off || !smoothing || !mats ||
      !refl || !vwork || !slot_xy || !smooth_slots || !weights || !stats || !status)
    return tps_fail(PIV_EINVAL, "NULL argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const PhiLattice lat = make_lattice(lattice);
  const bool gen = generated != 0 && lat.table != nullptr;
  launch_factor(nd, n_max, xs, ys, dom_off, slot_pt, mat_off, smoothing, mats, refl, vwork,
                slot_xy, nslots, status, order, lat, gen, s);
  launch_solve(nd, n_max, vx, vy, dom_off, slot_pt, mat_off, mats, refl, vwork, smoothing,
               threshold, smooth_slots, nslots, weights, wstride, stats, status, order, s);
  return tps_check("piv_tps_smooth");
}

int64_t piv_tps_shared_workspace(int32_t nfields, int32_t dom_cap, int32_t npat_cap,
                                 int32_t n_max, int32_t cap_r) {
  if (nfields <= 0 || dom_cap <= 0 || npat_cap <= 0 || n_max <= 0 || cap_r < 0) return -1;
  return int64_t(shared_work(nullptr, nfields, dom_cap, npat_cap, n_max, cap_r).bytes);
}

int piv_tps_smooth_shared(int32_t nfields, int32_t dom_cap, const piv_tps_field* fields,
                          const double* lim, const int32_t* fstat, const int32_t* dom_field,
                          const int32_t* counts, const double* gx, int32_t ngx, int32_t irr_x,
                          const double* gy, int32_t ngy, int32_t irr_y, int32_t npat_cap,
                          int32_t cap_r, int32_t n_max, const double* xs, const double* ys,
                          const int32_t* idx, const float* vx, const float* vy,
                          const int64_t* dom_off, const int32_t* slot_pt, const int64_t* mat_off,
                          const double* smoothing, double threshold, double* mats, double* refl,
                          double* vwork, double* slot_xy, double* smooth_slots, int64_t nslots,
                          double* weights, int64_t wstride, double* stats, int32_t* status,
                          const int32_t* order, const piv_tps_lattice* lattice,
                          int32_t generated, int32_t parts, void* work, void* stream) {
  if (nfields <= 0 || dom_cap <= 0) return PIV_OK;
  if (parts <= 0 || parts > PIV_TPS_PART_ALL) return tps_fail(PIV_EINVAL, "bad parts");
  if (n_max < 4 || n_max > kMaxN) return tps_fail(PIV_EINVAL, "n_max must be in [4, 2048]");
  if (npat_cap <= 0 || npat_cap > kPatMaxPerField)
    return tps_fail(PIV_EINVAL, "npat_cap must be in [1, 64]");
  if (cap_r < 0 || cap_r > kCapX) return tps_fail(PIV_EINVAL, "cap_r must be in [0, 32]");
  // the substitution path corrects for at most kCapMaxR missing nodes
  const bool subst = g_shared_solve.load(std::memory_order_relaxed) == 1;
  if (subst) cap_r = std::min(cap_r, kCapMaxR);
  if (shared_columns_lds(kSharedNV, (n_max + 1) & ~1) > 120 * 1024)
    return tps_fail(PIV_EUNSUPPORTED, "n_max too large for the shared solve's LDS");
  if (int64_t(nfields) * dom_cap > INT32_MAX / 4 || int64_t(nfields) * npat_cap > 1 << 20)
    return tps_fail(PIV_EUNSUPPORTED, "TPS batch too large");
  if (!fields || !lim || !fstat || !dom_field || !counts || !gx || !gy || !xs || !ys || !idx ||
      !vx || !vy || !dom_off || !slot_pt || !mat_off || !smoothing || !mats || !refl ||
      !vwork || !slot_xy || !smooth_slots || !weights || !stats || !status || !order || !work)
    return tps_fail(PIV_EINVAL, "NULL argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nd = nfields * dom_cap;
  const int npat = nfields * npat_cap;
  const int nm = (n_max + 1) & ~1;
  const SharedWork w = shared_work(work, nfields, dom_cap, npat_cap, n_max, cap_r);
  const PhiLattice lat = make_lattice(lattice);
  const bool gen = generated != 0 && lat.table != nullptr;
  if (parts & PIV_TPS_PART_PREP) {
    hipLaunchKernelGGL(k_tps_patterns, dim3(nfields), dim3(256), 0, s, int(dom_cap), fields,
                       lim, fstat, dom_field, counts, smoothing, gx, int(ngx), int(irr_x), gy,
                       int(ngy), int(irr_y), int(cap_r), int(n_max), int(npat_cap), w.dmode,
                       w.dpat, w.dbox, w.pat_n, w.pat_box, w.pat_smooth);
    hipLaunchKernelGGL(k_tps_pattern_layout, dim3(1), dim3(1024), 0, s, npat, w.pat_n, nd, order,
                       w.dmode, w.pdom_off, w.pmat_off, w.pxo, w.order_direct, w.order_shared,
                       w.plist);
    hipLaunchKernelGGL(k_tps_pattern_nodes, dim3(unsigned(std::min(npat, 2 * cu_count()))),
                       dim3(256), 0, s, npat, w.plist, w.pat_n, w.pat_box, w.pdom_off, gx, gy,
                       w.pxy, w.pslot, w.pxy_stride);
    hipLaunchKernelGGL(k_tps_zero_i32, dim3((npat + 255) / 256), dim3(256), 0, s, npat,
                       w.pstatus);
  }
  // the subdomains factorised alone
  if (parts & PIV_TPS_PART_DIRECT) {
    // few of the slots: workgroups striding over the direct order
    const int grid = 4 * cu_count();
    launch_factor(nd, n_max, xs, ys, dom_off, slot_pt, mat_off, smoothing, mats, refl, vwork,
                  slot_xy, nslots, status, w.order_direct, lat, gen, s, grid);
    launch_solve(nd, n_max, vx, vy, dom_off, slot_pt, mat_off, mats, refl, vwork, smoothing,
                 threshold, smooth_slots, nslots, weights, wstride, stats, status,
                 w.order_direct, s, grid);
  }
  if (!(parts & PIV_TPS_PART_SHARED)) return tps_check("piv_tps_smooth_shared");
  // the patterns' factorisations, then the subdomains that share them
  launch_factor(npat, n_max, w.pxy, w.pxy + w.pxy_stride, w.pdom_off, w.pslot, w.pmat_off,
                w.pat_smooth, w.pmats, w.prefl, w.pvwork, w.pslot_xy, w.pxy_stride, w.pstatus,
                w.plist, lat, gen, s, 2 * cu_count());
  if (!subst) {
    // the explicit inverses of the patterns, then the shared subdomains in
    // groups of one pattern
    hipLaunchKernelGGL(k_tps_shared_sort, dim3(nfields), dim3(256), 0, s, int(dom_cap),
                       int(npat_cap), w.pat_n, w.dpat, w.order_shared, nd, w.order_sorted);
    hipLaunchKernelGGL(k_tps_shared_groups, dim3(1), dim3(1024), 0, s, nd, w.dpat,
                       w.order_sorted, shared_x_group(nm), w.groups);
    const int nblk = (n_max - 3 + kNB - 1) / kNB;
    const int64_t nitem = int64_t(npat) * nblk * (nblk + 1) / 2;
    auto grid = [&](int64_t per_cu) { return dim3(unsigned(std::min<int64_t>(nitem, 256 * per_cu))); };
    hipLaunchKernelGGL(k_tps_tri_inverse, grid(4), dim3(kThreads), 0, s, w.plist, npat, nblk,
                       w.pat_n, w.pmat_off, w.pmats, w.pstatus, w.pinv);
    const int nt = (n_max - 3 + 63) / 64;
    hipLaunchKernelGGL(k_tps_inv_gram, grid(4), dim3(kThreads), 0, s, w.plist, npat,
                       nt * (nt + 1) / 2, w.pat_n, w.pmat_off, w.pmats, w.pstatus, w.pinv);
    hipLaunchKernelGGL(k_tps_xprep, dim3(unsigned(std::min(npat, 256))), dim3(kThreads),
                       size_t(6) * nm * sizeof(double), s, w.plist, npat, w.pat_n, w.pmat_off,
                       w.pmats, w.pdom_off, w.prefl, w.pvwork, w.pstatus, w.pxo, w.pinv, w.pac);
    hipLaunchKernelGGL(k_tps_xfill, grid(8), dim3(kThreads), 0, s, w.plist, npat,
                       (n_max + 7) / 8, w.pat_n, w.pmat_off, w.pmats, w.pdom_off, w.pstatus, w.pxo,
                       w.pac, w.pinv);
#define FIA_SHARED_X(G)                                                                           \
  hipLaunchKernelGGL(k_tps_shared_x<G>, dim3(std::max(8, std::min(nd, 2048))), dim3(kThreads),    \
                     shared_x_lds(G, nm), s, nd, w.hfac, w.order_sorted, w.groups, w.dpat, w.dbox,\
                     w.pat_n, w.pstatus, w.pxo, w.pinv,                                           \
                     gx, gy, int(ngx), xs, ys, idx, vx, vy, dom_off, slot_pt, smoothing,          \
                     threshold, smooth_slots, nslots, weights, wstride, stats, status, slot_xy,   \
                     int(cap_r), nm)
    switch (shared_x_group(nm)) {
      case 4: FIA_SHARED_X(4); break;
      case 2: FIA_SHARED_X(2); break;
      default: FIA_SHARED_X(1); break;
    }
#undef FIA_SHARED_X
    return tps_check("piv_tps_smooth_shared");
  }
  // by blocked substitution with the patterns' factors: subdomains missing
  // at most 4 nodes (and every check), then the others
  hipLaunchKernelGGL(k_tps_shared_columns<4>, dim3(nd), dim3(kThreads),
                     shared_columns_lds(4, nm), s, w.order_shared, w.dpat, w.dbox, w.pat_n,
                     w.pdom_off, w.pmat_off, w.pmats, w.prefl, w.pvwork, w.pstatus, gx, gy,
                     int(ngx), xs, ys, idx, dom_off, slot_pt, status, w.gmat, w.hfac, w.gpoly,
                     int(cap_r), nm);
  hipLaunchKernelGGL(k_tps_shared_columns<kSharedNV>, dim3(nd), dim3(kThreads),
                     shared_columns_lds(kSharedNV, nm), s, w.order_shared, w.dpat, w.dbox,
                     w.pat_n, w.pdom_off, w.pmat_off, w.pmats, w.prefl, w.pvwork, w.pstatus, gx,
                     gy, int(ngx), xs, ys, idx, dom_off, slot_pt, status, w.gmat, w.hfac,
                     w.gpoly, int(cap_r), nm);
  hipLaunchKernelGGL(k_tps_shared_iter, dim3(nd), dim3(kThreads), shared_iter_lds(kIterNV, nm),
                     s, w.order_shared, w.dpat, w.dbox, w.pat_n, w.pdom_off, w.pmat_off,
                     w.pmats, w.prefl, w.pvwork, gx, gy, int(ngx), xs, ys, idx, vx, vy, dom_off,
                     slot_pt, smoothing, threshold, smooth_slots, nslots, weights, wstride, stats,
                     status, slot_xy, w.gmat, w.hfac, w.gpoly, int(cap_r), nm);
  return tps_check("piv_tps_smooth_shared");
}

int piv_tps_influence(int32_t nd, int32_t n_max, const int64_t* dom_off,
                      const int64_t* mat_off, const double* mats, const double* refl,
                      const double* vwork, const int32_t* status, double* infl, int64_t stride,
                      void* stream) {
  if (nd <= 0) return PIV_OK;
  if (n_max < 4 || n_max > kMaxN) return tps_fail(PIV_EINVAL, "n_max must be in [4, 2048]");
  if (stride < int64_t(n_max) * n_max) return tps_fail(PIV_EINVAL, "stride < n_max^2");
  if (!dom_off || !mat_off || !mats || !refl || !vwork || !status || !infl)
    return tps_fail(PIV_EINVAL, "NULL argument");
  const int nm = (n_max + 1) & ~1;
  const size_t lds = (size_t(8) * nm + kSolveScratch) * sizeof(double);
  if (lds > 120 * 1024) return tps_fail(PIV_EUNSUPPORTED, "n_max too large for the LDS");
  hipLaunchKernelGGL(k_tps_influence, dim3(nd), dim3(kThreads), lds,
                     static_cast<hipStream_t>(stream), dom_off, mat_off, mats, refl, vwork,
                     status, nm, infl, stride);
  return tps_check("piv_tps_influence");
}

int piv_tps_node_field(int32_t transpose, int32_t nfields, int32_t nc, int32_t nd,
                       int32_t n_max, const int64_t* dom_off, const int32_t* slot_pt,
                       int64_t nused, const int32_t* status, const double* smoothing,
                       const double* infl, int64_t stride, const int32_t* c_slot,
                       const double* c_coef, const double* c_norm, int64_t nslots,
                       const double* in, double* out, double* work, void* stream) {
  if (nfields <= 0 || nc <= 0 || nd <= 0) return PIV_OK;
  if (n_max < 4 || n_max > kMaxN) return tps_fail(PIV_EINVAL, "n_max must be in [4, 2048]");
  if (nused < 0 || nused > nslots) return tps_fail(PIV_EINVAL, "nused must be in [0, nslots]");
  if (nfields > 65535 || nd > 65535) return tps_fail(PIV_EUNSUPPORTED, "too many fields or subdomains");
  if (!dom_off || !slot_pt || !status || !smoothing || !infl || !c_slot || !c_coef || !c_norm ||
      !in || !out || !work)
    return tps_fail(PIV_EINVAL, "NULL argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t sc = int64_t(nfields) * nslots;  // component stride of the slot buffers
  const int64_t ncs = int64_t(nfields) * nc;    // component stride of in / out
  double* w0 = work;
  double* w1 = work + 2 * sc;
  const dim3 gc((nc + 255) / 256, nfields);
  const size_t lds = size_t(2) * n_max * sizeof(double);
  if (!transpose) {
    const dim3 gs(unsigned((nused + 255) / 256), nfields);
    if (nused > 0)
      hipLaunchKernelGGL(k_tps_node_steps, gs, dim3(256), 0, s, 0, int(nc), nused, slot_pt,
                         c_slot, c_coef, c_norm, nslots, sc, ncs, in, w0);
    hipLaunchKernelGGL(k_tps_node_apply, dim3(nfields, nd), dim3(kThreads), lds, s, dom_off,
                       status, smoothing, infl, stride, w0, w1, nslots, sc, 0);
    hipLaunchKernelGGL(k_tps_node_steps, gc, dim3(256), 0, s, 1, int(nc), nused, slot_pt, c_slot,
                       c_coef, c_norm, nslots, sc, ncs, w1, out);
  } else {
    hipLaunchKernelGGL(k_tps_node_steps, gc, dim3(256), 0, s, 2, int(nc), nused, slot_pt, c_slot,
                       c_coef, c_norm, nslots, sc, ncs, in, w0);
    hipLaunchKernelGGL(k_tps_node_apply, dim3(nfields, nd), dim3(kThreads), lds, s, dom_off,
                       status, smoothing, infl, stride, w0, w1, nslots, sc, 1);
    hipLaunchKernelGGL(k_tps_node_steps, gc, dim3(256), 0, s, 3, int(nc), nused, slot_pt, c_slot,
                       c_coef, c_norm, nslots, sc, ncs, w1, out);
  }
  return tps_check("piv_tps_node_field");
}

int piv_tps_select(int32_t nfields, int32_t n, const float* dx, const float* dy,
                   const double* xs, const double* ys, double* cx, double* cy, float* vx,
                   float* vy, int32_t* idx, double* info, double lattice_step, void* stream) {
  if (nfields <= 0 || n <= 0) return tps_fail(PIV_EINVAL, "no vector");
  if (!dx || !dy || !xs || !ys || !cx || !cy || !vx || !vy || !idx || !info)
    return tps_fail(PIV_EINVAL, "NULL argument");
  hipLaunchKernelGGL(k_tps_select, dim3(nfields), dim3(1024), 0,
                     static_cast<hipStream_t>(stream), n, dx, dy, xs, ys, cx, cy, vx, vy, idx,
                     info, lattice_step);
  return tps_check("piv_tps_select");
}

int piv_tps_count(int32_t nd, const piv_tps_field* fields, const int32_t* dom_field,
                  const double* lim, const double* cx, const double* cy, int32_t* counts,
                  void* stream) {
  if (nd <= 0) return tps_fail(PIV_EINVAL, "empty subdomain grid");
  if (!fields || !dom_field || !lim || !cx || !cy || !counts)
    return tps_fail(PIV_EINVAL, "NULL argument");
  hipLaunchKernelGGL(k_tps_count, dim3(nd), dim3(256), 0, static_cast<hipStream_t>(stream),
                     fields, dom_field, lim, cx, cy, counts);
  return tps_check("piv_tps_count");
}

int piv_tps_fill(int32_t nd, const piv_tps_field* fields, const int32_t* dom_field,
                 const double* lim, const double* cx, const double* cy, const int64_t* dom_off,
                 int32_t* slot_pt, void* stream) {
  if (nd <= 0) return tps_fail(PIV_EINVAL, "empty subdomain grid");
  if (!fields || !dom_field || !lim || !cx || !cy || !dom_off || !slot_pt)
    return tps_fail(PIV_EINVAL, "NULL argument");
  hipLaunchKernelGGL(k_tps_fill, dim3(nd), dim3(256), 0, static_cast<hipStream_t>(stream),
                     fields, dom_field, lim, cx, cy, dom_off, slot_pt);
  return tps_check("piv_tps_fill");
}

int piv_tps_point_slots(int32_t nfields, const piv_tps_field* fields, int32_t max_points,
                        const double* lim, const double* px, const double* py,
                        const int64_t* dom_off, const int32_t* slot_pt, int32_t* out_slot,
                        double* out_coef, double* out_norm, int32_t* status, void* stream) {
  if (nfields <= 0 || max_points <= 0) return PIV_OK;
  if (!fields || !lim || !px || !py || !out_slot || !out_coef || !out_norm || !status)
    return tps_fail(PIV_EINVAL, "NULL argument");
  if ((dom_off == nullptr) != (slot_pt == nullptr))
    return tps_fail(PIV_EINVAL, "dom_off and slot_pt go together");
  hipLaunchKernelGGL(k_tps_point_slots, dim3((max_points + 255) / 256, nfields), dim3(256), 0,
                     static_cast<hipStream_t>(stream), fields, lim, max_points, px, py, dom_off,
                     slot_pt, out_slot, out_coef, out_norm, status);
  return tps_check("piv_tps_point_slots");
}

int piv_tps_grid(int32_t nfields, int32_t n, const double* info, double subdom_size,
                 double smoothing_coef, double coef_buffer, int32_t dom_cap, double gen_width,
                 int32_t require_lattice, piv_tps_field* fields, double* lim,
                 int32_t* dom_field, double* smoothing, int32_t* fstat, void* stream) {
  if (nfields <= 0 || n <= 0 || dom_cap <= 0) return tps_fail(PIV_EINVAL, "empty TPS batch");
  if (nfields > kOffsetsMaxFields)
    return tps_fail(PIV_EUNSUPPORTED, "more than 2048 fields in a device-sized TPS batch");
  if (int64_t(nfields) * n > INT32_MAX || int64_t(nfields) * dom_cap * 4 > INT32_MAX)
    return tps_fail(PIV_EUNSUPPORTED, "device-sized TPS batch too large");
  if (!(subdom_size > 0.0)) return tps_fail(PIV_EINVAL, "subdom_size must be positive");
  if (!info || !fields || !lim || !dom_field || !smoothing || !fstat)
    return tps_fail(PIV_EINVAL, "NULL argument");
  hipLaunchKernelGGL(k_tps_grid, dim3(nfields), dim3(256), 0, static_cast<hipStream_t>(stream),
                     info, int(n), subdom_size, smoothing_coef, coef_buffer, int(dom_cap),
                     gen_width, int(require_lattice), fields, lim, dom_field, smoothing, fstat);
  return tps_check("piv_tps_grid");
}

int piv_tps_offsets(int32_t nfields, int32_t dom_cap, piv_tps_field* fields,
                    const int32_t* counts, const int32_t* dom_field, int64_t slots_cap,
                    int64_t mats_cap, int32_t n_max_cap, int32_t* fstat, int64_t* dom_off,
                    int64_t* mat_off, int32_t* order, int64_t* totals, void* stream) {
  if (nfields <= 0 || dom_cap <= 0) return tps_fail(PIV_EINVAL, "empty TPS batch");
  if (nfields > kOffsetsMaxFields)
    return tps_fail(PIV_EUNSUPPORTED, "more than 2048 fields in a device-sized TPS batch");
  if (n_max_cap < 4 || n_max_cap > kMaxN)
    return tps_fail(PIV_EINVAL, "n_max_cap must be in [4, 2048]");
  if (!fields || !counts || !dom_field || !fstat || !dom_off || !mat_off || !order || !totals)
    return tps_fail(PIV_EINVAL, "NULL argument");
  hipLaunchKernelGGL(k_tps_offsets, dim3(1), dim3(1024), 0, static_cast<hipStream_t>(stream),
                     int(nfields), int(dom_cap), fields, counts, dom_field, slots_cap, mats_cap,
                     int(n_max_cap), fstat, dom_off, mat_off, order, totals);
  return tps_check("piv_tps_offsets");
}

int piv_tps_at_centers(int32_t nfields, int32_t n, const double* info, const int32_t* idx,
                       const double* xs, const double* ys, const double* new_x,
                       const double* new_y, const double* smoothed, double* out,
                       uint8_t* skip, void* stream) {
  if (nfields <= 0 || n <= 0) return PIV_OK;
  if (nfields > 65535) return tps_fail(PIV_EINVAL, "too many fields for one launch");
  if (!info || !idx || !xs || !ys || !new_x || !new_y || !smoothed || !out || !skip)
    return tps_fail(PIV_EINVAL, "NULL argument");
  hipLaunchKernelGGL(k_tps_at_centers, dim3((n + 255) / 256, nfields), dim3(256), 0,
                     static_cast<hipStream_t>(stream), int(n), info, idx, xs, ys, new_x, new_y,
                     smoothed, int64_t(nfields) * n, out, skip);
  return tps_check("piv_tps_at_centers");
}

int piv_tps_blend(int32_t npts, const int32_t* slots, const double* coef, const double* norm,
                  const double* smooth_slots, int64_t nslots, double* out, void* stream) {
  if (npts <= 0) return PIV_OK;
  if (!slots || !coef || !norm || !smooth_slots || !out)
    return tps_fail(PIV_EINVAL, "NULL argument");
  hipLaunchKernelGGL(k_tps_blend, dim3((npts + 255) / 256), dim3(256), 0,
                     static_cast<hipStream_t>(stream), npts, slots, coef, norm, smooth_slots,
                     nslots, out);
  return tps_check("piv_tps_blend");
}

int64_t piv_tps_matrix_size(int32_t n) {
  if (n < 0 || n > kMaxN) return -1;
  return tps_mat_size(n);
}

int64_t piv_tps_interpolate_workspace(int32_t npts, int32_t nd) {
  if (npts < 0 || nd < 0) return -1;
  return int64_t(interp_work(nullptr, npts, nd).bytes);
}

int piv_tps_interpolate(int32_t npts, const double* px, const double* py, const int32_t* doms,
                        const double* coef, const double* norm, const double* slot_xy,
                        int64_t nslots, int32_t nd, int32_t n_max, const int64_t* dom_off,
                        const double* weights, int64_t wstride, const uint8_t* skip,
                        double* out, void* work, const piv_tps_lattice* lattice,
                        void* stream) {
  return tps_interpolate_impl("piv_tps_interpolate", npts, px, py, doms, coef, norm, slot_xy,
                              nslots, nd, n_max, dom_off, weights, wstride, skip, out, work,
                              lattice, stream);
}

int piv_tps_interpolate_noalloc(int32_t npts, const double* px, const double* py,
                                const int32_t* doms, const double* coef, const double* norm,
                                const double* slot_xy, int64_t nslots, int32_t nd,
                                int32_t n_max, const int64_t* dom_off, const double* weights,
                                int64_t wstride, const uint8_t* skip, double* out,
                                const piv_tps_lattice* lattice, void* stream) {
  if (npts <= 0 || nd <= 0) return PIV_OK;
  if (npts > INT32_MAX / kMaxSlots)
    return tps_fail(PIV_EUNSUPPORTED, "too many TPS evaluation points");
  const size_t need = interp_work(nullptr, npts, nd).bytes;
  if (need > kNoAllocScratchBytes)
    return tps_fail(PIV_EUNSUPPORTED,
                    "piv_tps_interpolate_noalloc: " + std::to_string(npts) + " points and " +
                        std::to_string(nd) + " subdomains need " + std::to_string(need) +
                        " bytes of scratch, more than its " +
                        std::to_string(kNoAllocScratchBytes) +
                        " (use piv_tps_interpolate with a workspace)");
  void* work = nullptr;
  const int st = scratch_on_current_device(&work);
  if (st != PIV_OK) return st;
  return tps_interpolate_impl("piv_tps_interpolate_noalloc", npts, px, py, doms, coef, norm,
                              slot_xy, nslots, nd, n_max, dom_off, weights, wstride, skip, out,
                              work, lattice, stream);
}

int piv_tps_phi_table(double step, int32_t a_max, double* table, void* stream) {
  if (!(step > 0.0) || a_max < 0 || a_max > kMaxLatticeA)
    return tps_fail(PIV_EINVAL, "lattice step must be positive and a_max in [0, 4095]");
  if (!table) return tps_fail(PIV_EINVAL, "NULL argument");
  const int w = a_max + 1;
  hipLaunchKernelGGL(k_tps_phi_table, dim3((w * w + 255) / 256), dim3(256), 0,
                     static_cast<hipStream_t>(stream), step, int(a_max), table);
  return tps_check("piv_tps_phi_table");
}

}  // extern "C"

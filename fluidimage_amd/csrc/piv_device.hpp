// Device-side building blocks of the PIV correlation path (gfx950).
//
//   * window load from HBM with the reference's zero padding and min offset
//     (src/fluidimage/works/piv/singlepass.py:196-208,391-407);
//   * peak search with the reference's tie rule, error rules and second
//     chance (src/fluidimage/calcul/correl.py:88-164,226-275);
//   * 2d_gaussian2 / centroid sub-pixel fit (src/fluidimage/calcul/subpix.py:16-218).
//
// A window is served by a group of T lanes (T a power of two dividing the
// 64-lane wavefront), so every cross-lane reduction is a __shfl_xor within
// the group and never needs a workgroup barrier.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/piv_hip.h"

namespace fia {

constexpr int kMaxPeaks = 7;  // nb_peaks_to_search limit

constexpr int lanes_for(int n) {
  int t = 8;
  while (t < n) t *= 2;
  return t;
}

template <int DT>
__device__ __forceinline__ float fetch_px(const void* data, int64_t idx) {
  if constexpr (DT == PIV_U8) return float(static_cast<const uint8_t*>(data)[idx]);
  else if constexpr (DT == PIV_U16) return float(static_cast<const uint16_t*>(data)[idx]);
  else return static_cast<const float*>(data)[idx];
}

template <class V>
__device__ __forceinline__ V group_sum(V v, int T) {
  for (int off = T / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, T);
  return v;
}

// numpy bounds of a[start:stop] on an axis of length n
__device__ __forceinline__ void py_slice(int start, int stop, int n, int& lo, int& hi) {
  if (start < 0) {
    start += n;
    if (start < 0) start = 0;
  } else if (start > n) {
    start = n;
  }
  if (stop < 0) {
    stop += n;
    if (stop < 0) stop = 0;
  } else if (stop > n) {
    stop = n;
  }
  lo = start;
  hi = stop;
}

struct PlanArgs {
  const uint8_t* mask;  // where_large_displacement (n0*n1 bytes) or nullptr
  int b0s, b0e, b1s, b1e;
  int subpix_method;
  int nsubpix;
  int particle_radius;
  int nb_peaks;
  int has_dmax;
  // the maps the search and fit see are map_scale^-1 times the reference's
  // (a power of two; 0 reads as 1): the '2d_gaussian' fit, which is not
  // scale invariant (values <= 0 become 1e-6), scales them back
  float map_scale;
};

// Peak search of one window on an LDS search copy of the map in which the
// masked elements (where_large_displacement) and the NaN squares written so
// far (the reference NaNs (2r+1)^2 squares before a second attempt and
// between secondary peaks, correl.py:234-237,256-259) are NaN.
template <int T>
struct Search {
  float* s;
  int n0, n1;  // map shape
  int b0s, b0e, b1s, b1e;

  // nan_indices_max (correl.py:88-132): last maximal non-NaN element of the
  // box in row-major order; returns the PIV_WIN_* code of the peak.
  __device__ __forceinline__ int find(int lane, int& oi0, int& oi1) const {
    float best = 0.f;
    int blin = -1;
    for (int c = b1s + lane; c < b1e; c += T) {
#pragma unroll 4
      for (int r = b0s; r < b0e; ++r) {
        const float v = s[r * n1 + c];
        if (v != v) continue;
        // rows ascend within a lane: ">=" keeps the last maximum
        if (blin < 0 || v >= best) {
          best = v;
          blin = r * n1 + c;
        }
      }
    }
    for (int off = T / 2; off > 0; off >>= 1) {
      const float ob = __shfl_xor(best, off, T);
      const int ol = __shfl_xor(blin, off, T);
      if (ol >= 0 && (blin < 0 || ob > best || (ob == best && ol > blin))) {
        best = ob;
        blin = ol;
      }
    }
    if (blin < 0) {
      oi0 = oi1 = 0;
      return PIV_WIN_NO_PEAK;
    }
    const int i0 = blin / n1, i1 = blin % n1;
    oi0 = i0;
    oi1 = i1;
    if (i0 == 0 || i0 == n0 - 1 || i1 == 0 || i1 == n1 - 1) return PIV_WIN_PEAK_BOUNDARY;
    bool nan = false;
#pragma unroll
    for (int d0 = -1; d0 <= 1; ++d0)
#pragma unroll
      for (int d1 = -1; d1 <= 1; ++d1) {
        const float v = s[(i0 + d0) * n1 + i1 + d1];
        nan |= v != v;
      }
    return nan ? PIV_WIN_PEAK_NAN : PIV_WIN_OK;
  }

  // NaN the (2r+1)^2 square around (i0, i1) with numpy slice bounds; the
  // window's lanes share one wavefront, so a wavefront fence orders the
  // writes before the next search.
  __device__ __forceinline__ void nan_square(int lane, int i0, int i1, int r, int* q) {
    int a0, a1, c0, c1;
    py_slice(i0 - r, i0 + r + 1, n0, a0, a1);
    py_slice(i1 - r, i1 + r + 1, n1, c0, c1);
    const int h = a1 > a0 ? a1 - a0 : 0, w = c1 > c0 ? c1 - c0 : 0;
    for (int k = lane; k < h * w; k += T) s[(a0 + k / w) * n1 + c0 + k % w] = __builtin_nanf("");
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (q != nullptr) {
      q[0] = a0;
      q[1] = a1;
      q[2] = c0;
      q[3] = c1;
    }
  }
};

// The same search on the map itself: an element is skipped when it is NaN,
// when bit r of scol[c] is set (where_large_displacement, scol == nullptr
// for none) or when it lies in one of the squares "NaN-ed" so far, which are
// kept as bounds (wavefront-uniform values) rather than written. Saves the
// search copy (n0 n1 floats of LDS per window).
constexpr int kMaxSquares = kMaxPeaks + 1;

template <int T>
struct MaskedSearch {
  const float* m;
  const uint64_t* scol;
  int n0, n1;
  int b0s, b0e, b1s, b1e;
  int nsq = 0;
  int sa0[kMaxSquares], sa1[kMaxSquares], sc0[kMaxSquares], sc1[kMaxSquares];

  __device__ __forceinline__ bool in_squares(int r, int c) const {
    bool in = false;
#pragma unroll
    for (int k = 0; k < kMaxSquares; ++k)
      in |= k < nsq && r >= sa0[k] && r < sa1[k] && c >= sc0[k] && c < sc1[k];
    return in;
  }

  __device__ __forceinline__ bool skipped(int r, int c, float v) const {
    if (v != v) return true;
    if (scol != nullptr && ((scol[c] >> r) & 1)) return true;
    return nsq > 0 && in_squares(r, c);
  }

  __device__ __forceinline__ int find(int lane, int& oi0, int& oi1) const {
    float best = 0.f;
    int blin = -1;
    if (nsq == 0) {
      // with best = -inf, "v >= best" is the non-NaN test and the tie rule
      // (the last maximum of the lane's column) at once
      best = -__builtin_inff();
      for (int c = b1s + lane; c < b1e; c += T) {
        const float* col = m + c;
        int br = -1;
        if (scol == nullptr) {
#pragma unroll 8
          for (int r = b0s; r < b0e; ++r) {
            const float v = col[r * n1];
            if (v >= best) {
              best = v;
              br = r;
            }
          }
        } else {
          const uint64_t bits = scol[c];
          const uint32_t lo = uint32_t(bits), hi = uint32_t(bits >> 32);
#pragma unroll 4
          for (int r = b0s; r < b0e; ++r) {
            const float v = col[r * n1];
            const uint32_t masked = ((r < 32 ? lo : hi) >> (r & 31)) & 1u;
            if (!masked && v >= best) {
              best = v;
              br = r;
            }
          }
        }
        if (br >= 0) blin = br * n1 + c;
      }
    } else {
      for (int c = b1s + lane; c < b1e; c += T) {
        const uint64_t bits = scol != nullptr ? scol[c] : uint64_t(0);
        for (int r = b0s; r < b0e; ++r) {
          const float v = m[r * n1 + c];
          if (v != v || ((bits >> r) & 1) || in_squares(r, c)) continue;
          if (blin < 0 || v >= best) {
            best = v;
            blin = r * n1 + c;
          }
        }
      }
    }
    for (int off = T / 2; off > 0; off >>= 1) {
      const float ob = __shfl_xor(best, off, T);
      const int ol = __shfl_xor(blin, off, T);
      if (ol >= 0 && (blin < 0 || ob > best || (ob == best && ol > blin))) {
        best = ob;
        blin = ol;
      }
    }
    if (blin < 0) {
      oi0 = oi1 = 0;
      return PIV_WIN_NO_PEAK;
    }
    const int i0 = blin / n1, i1 = blin % n1;
    oi0 = i0;
    oi1 = i1;
    if (i0 == 0 || i0 == n0 - 1 || i1 == 0 || i1 == n1 - 1) return PIV_WIN_PEAK_BOUNDARY;
    bool nan = false;
#pragma unroll
    for (int d0 = -1; d0 <= 1; ++d0)
#pragma unroll
      for (int d1 = -1; d1 <= 1; ++d1)
        nan |= skipped(i0 + d0, i1 + d1, m[(i0 + d0) * n1 + i1 + d1]);
    return nan ? PIV_WIN_PEAK_NAN : PIV_WIN_OK;
  }

  __device__ __forceinline__ void nan_square(int lane, int i0, int i1, int r, int* q) {
    int a0, a1, c0, c1;
    py_slice(i0 - r, i0 + r + 1, n0, a0, a1);
    py_slice(i1 - r, i1 + r + 1, n1, c0, c1);
#pragma unroll
    for (int k = 0; k < kMaxSquares; ++k)
      if (k == nsq) {
        sa0[k] = a0;
        sa1[k] = a1;
        sc0[k] = c0;
        sc1[k] = c1;
      }
    if (nsq < kMaxSquares) ++nsq;
    if (q != nullptr) {
      q[0] = a0;
      q[1] = a1;
      q[2] = c0;
      q[3] = c1;
    }
  }
};

// correl_max = correl[iy, ix] / norm with numpy-2 float32 promotion
// (correl.py:153-157); 0 when norm == 0.
__device__ __forceinline__ float peak_value(float v, double norm) {
  return norm == 0.0 ? 0.f : v / float(norm);
}

// the log-quadratic fit of compute_subpix_2d_gaussian2 (subpix.py:45-61)
// from the sums of x l, y l, x y l, (3 x^2 - 2) l, (3 y^2 - 2) l over the
// 3x3 crop (l the logarithms, x, y in -1..1)
__device__ __forceinline__ void gaussian2_fit(double c10, double c01, double c11, double c20,
                                              double c02, double& dpx, double& dpy) {
  c10 /= 6;
  c01 /= 6;
  c11 /= 4;
  c20 /= 6;
  c02 /= 6;
  double den = 4 * c20 * c02 - c11 * c11;
  dpx = (c11 * c01 - 2 * c10 * c02) / den;
  dpy = (c11 * c10 - 2 * c01 * c20) / den;
}

// compute_subpix_2d_gaussian2 (subpix.py:16-61): float32 normalisation of
// the 3x3 crop, then the log-quadratic fit.
__device__ __forceinline__ void gaussian2(const float* m, int n1, int ix, int iy, double& dpx,
                                          double& dpy) {
  float c[9];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) c[3 * a + b] = m[(iy - 1 + a) * n1 + (ix - 1 + b)];
  float mn = c[0];
  for (int k = 1; k < 9; ++k) mn = fminf(mn, c[k]);
  for (int k = 0; k < 9; ++k) c[k] -= mn;
  float mx = c[0];
  for (int k = 1; k < 9; ++k) mx = fmaxf(mx, c[k]);
  for (int k = 0; k < 9; ++k) c[k] /= mx;  // float32 division, as the reference's /=
  double c10 = 0, c01 = 0, c11 = 0, c20 = 0, c02 = 0;
  for (int a = 0; a < 3; ++a) {
    for (int b = 0; b < 3; ++b) {
      float v = c[3 * a + b];
      if (v == 0.f) v = 1e-8f;
      const double l = double(logf(v));
      int x = b - 1, y = a - 1;
      c10 += x * l;
      c01 += y * l;
      c11 += x * y * l;
      c20 += (3 * x * x - 2) * l;
      c02 += (3 * y * y - 2) * l;
    }
  }
  gaussian2_fit(c10, c01, c11, c20, c02, dpx, dpy);
}

// gaussian2 with the lanes of one window (T >= 16 of one wavefront, all
// active): lane k < 9 takes element k = 3 a + b of the 3x3 crop through the
// normalisation and the logarithm, the other lanes of the first 16 hold
// element 0 (neutral for the minimum and the maximum, which are exact in any
// order), and every lane then forms the five sums in gaussian2's order. The
// coefficients x = b - 1, y = a - 1, x y, 3 x^2 - 2 and 3 y^2 - 2 are 0, +-1
// or -2, so each product is exact and each term one rounded addition; the
// zero terms, which add 0 to a sum of finite logarithms, are left out. The
// result is gaussian2's bit for bit.
template <int T>
__device__ __forceinline__ void gaussian2_lanes(const float* m, int n1, int ix, int iy, int lane,
                                                double& dpx, double& dpy) {
  static_assert(T >= 16 && T <= 64, "one wavefront, at least 16 lanes per window");
  const int k = lane < 9 ? lane : 0;
  const int a = k / 3, b = k - 3 * a;
  float c = m[(iy - 1 + a) * n1 + (ix - 1 + b)];
  float mn = c;
  for (int off = 8; off > 0; off >>= 1) mn = fminf(mn, __shfl_xor(mn, off, T));
  c -= mn;
  float mx = c;
  for (int off = 8; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, T));
  c /= mx;
  if (c == 0.f) c = 1e-8f;
  const float lf = logf(c);
  double l[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) l[j] = double(__shfl(lf, j, T));
  double c10 = -l[0];
  c10 += l[2];
  c10 -= l[3];
  c10 += l[5];
  c10 -= l[6];
  c10 += l[8];
  double c01 = -l[0];
  c01 -= l[1];
  c01 -= l[2];
  c01 += l[6];
  c01 += l[7];
  c01 += l[8];
  double c11 = l[0];
  c11 -= l[2];
  c11 -= l[6];
  c11 += l[8];
  double c20 = l[0];
  c20 += -2.0 * l[1];
  c20 += l[2];
  c20 += l[3];
  c20 += -2.0 * l[4];
  c20 += l[5];
  c20 += l[6];
  c20 += -2.0 * l[7];
  c20 += l[8];
  double c02 = l[0];
  c02 += l[1];
  c02 += l[2];
  c02 += -2.0 * l[3];
  c02 += -2.0 * l[4];
  c02 += -2.0 * l[5];
  c02 += l[6];
  c02 += l[7];
  c02 += l[8];
  gaussian2_fit(c10, c01, c11, c20, c02, dpx, dpy);
}

__device__ __forceinline__ void centroid(const float* m, int n1, int ix, int iy, int n,
                                         double& dpx, double& dpy) {
  double s = 0, sx = 0, sy = 0;
  for (int a = -n; a <= n; ++a)
    for (int b = -n; b <= n; ++b) {
      double v = m[(iy + a) * n1 + (ix + b)];
      s += v;
      sx += b * v;
      sy += a * v;
    }
  dpx = sx / s;
  dpy = sy / s;
}

// The '2d_gaussian' fit of SubPix.compute_subpix (subpix.py:157-181): least
// squares of log c = a x^2 + b y^2 + c x + d y + e on the (2n+1)^2 crop
// (values <= 0 taken as 1e-6, float32 log). On the symmetric grid the normal
// equations decouple (the odd columns x, y are orthogonal to the even ones):
// with N = 2n + 1, S2 = sum x^2, S4 = sum x^4 over -n..n, u = sum x^2 l,
// v = sum y^2 l, w = sum l: a = (u - S2 w / N) / (N S4 - S2^2), the same for
// b with v, c = sum x l / (N S2), d = sum y l / (N S2) -- the reference's
// pseudo-inverse of M = [X^2, Y^2, X, Y, 1] applied to the logs. Returns
// false where the reference falls back to the centroid (a > 0 or b > 0, or
// a NaN displacement); sigma^2 = (1 / sqrt(-2 a))^2 as the reference rounds it.
__device__ __forceinline__ bool gaussian_ls(const float* m, int n1, int ix, int iy, int n,
                                            float scale, double& dpx, double& dpy) {
  double u = 0, v = 0, w = 0, sx = 0, sy = 0;
  for (int a = -n; a <= n; ++a)
    for (int b = -n; b <= n; ++b) {
      float c = m[(iy + a) * n1 + (ix + b)] * scale;
      if (c <= 0.f) c = 1e-6f;
      const double l = double(logf(c));
      u += double(b * b) * l;
      v += double(a * a) * l;
      w += l;
      sx += double(b) * l;
      sy += double(a) * l;
    }
  const double N = 2 * n + 1;
  const double S2 = double(n) * (n + 1) * (2 * n + 1) / 3.0;
  double S4 = 0.0;
  for (int k = 1; k <= n; ++k) S4 += 2.0 * double(k) * k * k * k;
  const double den = N * S4 - S2 * S2;
  const double ca = (u - S2 * w / N) / den;
  const double cb = (v - S2 * w / N) / den;
  if (ca > 0.0 || cb > 0.0) return false;
  const double sgx = 1.0 / sqrt(-2.0 * ca), sgy = 1.0 / sqrt(-2.0 * cb);
  dpx = sx / (N * S2) * (sgx * sgx);
  dpy = sy / (N * S2) * (sgy * sgy);
  return !(isnan(dpx) || isnan(dpy));
}

// SubPix.compute_subpix (subpix.py:107-218) -> PIV_WIN_* code; map n0 x n1
__device__ __forceinline__ int subpix(const float* m, int n0, int n1, int method, int nsub,
                                      int ix, int iy, double& dpx, double& dpy,
                                      float map_scale = 0.f) {
  dpx = dpy = 0.0;
  if (iy - nsub < 0 || iy + nsub + 1 > n0 || ix - nsub < 0 || ix + nsub + 1 > n1)
    return PIV_WIN_SUBPIX_BOUNDARY;
  double thr = 2.0 * (0.5 + nsub) * (0.5 + nsub);
  if (method == PIV_SUBPIX_NONE) return PIV_WIN_OK;
  if (method == PIV_SUBPIX_2D_GAUSSIAN) {
    // a too large displacement of this fit is an error, not a fallback
    if (gaussian_ls(m, n1, ix, iy, nsub, map_scale > 0.f ? map_scale : 1.f, dpx, dpy))
      return dpx * dpx + dpy * dpy > thr ? PIV_WIN_SUBPIX_WRONG : PIV_WIN_OK;
  }
  if (method == PIV_SUBPIX_2D_GAUSSIAN2) {
    gaussian2(m, n1, ix, iy, dpx, dpy);
    if (dpx * dpx + dpy * dpy > thr) method = PIV_SUBPIX_CENTROID;
    else return PIV_WIN_OK;
  }
  centroid(m, n1, ix, iy, nsub, dpx, dpy);
  if (dpx * dpx + dpy * dpy > thr) return PIV_WIN_SUBPIX_WRONG;
  return PIV_WIN_OK;
}

// subpix() with the lanes of one window (all T lanes active, every lane
// gets the result): the '2d_gaussian2' fit by gaussian2_lanes when T lanes
// of one wavefront are at least 16, the serial fit on every lane otherwise
template <int T>
__device__ __forceinline__ int subpix_lanes(const float* m, int n0, int n1, int method, int nsub,
                                            int ix, int iy, int lane, double& dpx, double& dpy,
                                            float map_scale) {
  if constexpr (T >= 16 && T <= 64) {
    if (method == PIV_SUBPIX_2D_GAUSSIAN2) {
      dpx = dpy = 0.0;
      if (iy - nsub < 0 || iy + nsub + 1 > n0 || ix - nsub < 0 || ix + nsub + 1 > n1)
        return PIV_WIN_SUBPIX_BOUNDARY;
      const double thr = 2.0 * (0.5 + nsub) * (0.5 + nsub);
      gaussian2_lanes<T>(m, n1, ix, iy, lane, dpx, dpy);
      if (!(dpx * dpx + dpy * dpy > thr)) return PIV_WIN_OK;
      centroid(m, n1, ix, iy, nsub, dpx, dpy);
      if (dpx * dpx + dpy * dpy > thr) return PIV_WIN_SUBPIX_WRONG;
      return PIV_WIN_OK;
    }
  }
  return subpix(m, n0, n1, method, nsub, ix, iy, dpx, dpy, map_scale);
}

// Everything that follows the correlation map of one window: peak search
// with second chance, secondary peaks, sub-pixel fit and the result
// records. `m` is the LDS map (reference layout), `srch` its LDS search copy
// (masked elements NaN), `win` the window index. `sq` receives the
// second-chance square (rows [sq0, sq1), cols [sq2, sq3)), empty if none,
// for the map side effect with displacement_max None. The map is n0 x n1
// with the zero displacement at (iy0, ix0): (n0 // 2 - 1, n1 // 2 - 1) for
// the FFT maps (correl.py:189), (dmax, dmax) for the direct maps
// (correl.py:409-410,444-445).
template <int T, class S>
__device__ __forceinline__ void finish_window_with(S& s, const float* m, int n0, int n1, int iy0,
                                                   int ix0, const PlanArgs& pa, int lane,
                                                   int win, double norm, int apply_subpix,
                                                   int dx_in, int dy_in, const piv_results& res,
                                                   int* sq);

template <int T>
__device__ __forceinline__ void finish_window(const float* m, float* srch, int n0, int n1,
                                              int iy0, int ix0, const PlanArgs& pa, int lane,
                                              int win, double norm, int apply_subpix, int dx_in,
                                              int dy_in, const piv_results& res, int* sq) {
  Search<T> s;
  s.s = srch;
  s.n0 = n0;
  s.n1 = n1;
  s.b0s = pa.b0s;
  s.b0e = pa.b0e;
  s.b1s = pa.b1s;
  s.b1e = pa.b1e;
  finish_window_with<T>(s, m, n0, n1, iy0, ix0, pa, lane, win, norm, apply_subpix, dx_in, dy_in,
                        res, sq);
}

// finish_window without a search copy: the mask as column bits (bit i0 of
// scol[i1], nullptr for none)
template <int T>
__device__ __forceinline__ void finish_window_masked(const float* m, const uint64_t* scol, int n0,
                                                     int n1, int iy0, int ix0, const PlanArgs& pa,
                                                     int lane, int win, double norm,
                                                     int apply_subpix, int dx_in, int dy_in,
                                                     const piv_results& res, int* sq) {
  MaskedSearch<T> s;
  s.m = m;
  s.scol = scol;
  s.n0 = n0;
  s.n1 = n1;
  s.b0s = pa.b0s;
  s.b0e = pa.b0e;
  s.b1s = pa.b1s;
  s.b1e = pa.b1e;
  finish_window_with<T>(s, m, n0, n1, iy0, ix0, pa, lane, win, norm, apply_subpix, dx_in, dy_in,
                        res, sq);
}

template <int T, class S>
__device__ __forceinline__ void finish_window_with(S& s, const float* m, int n0, int n1, int iy0,
                                                   int ix0, const PlanArgs& pa, int lane,
                                                   int win, double norm, int apply_subpix,
                                                   int dx_in, int dy_in, const piv_results& res,
                                                   int* sq) {
  sq[0] = sq[1] = sq[2] = sq[3] = 0;
  int i0, i1;
  int code = s.find(lane, i0, i1);
  float cmax = code == PIV_WIN_NO_PEAK ? __builtin_nanf("") : peak_value(m[i0 * n1 + i1], norm);
  if (code != PIV_WIN_OK && code != PIV_WIN_NO_PEAK) {
    // second chance (correl.py:231-246)
    s.nan_square(lane, i0, i1, pa.particle_radius, sq);
    int j0, j1;
    const int code2 = s.find(lane, j0, j1);
    if (code2 == PIV_WIN_OK) {
      i0 = j0;
      i1 = j1;
      cmax = peak_value(m[i0 * n1 + i1], norm);
      code = PIV_WIN_OK;
    } else {
      code = code2;
    }
  }
  const bool writer = lane == 0;
  if (code != PIV_WIN_OK) {
    if (writer) {
      const float dx = code == PIV_WIN_NO_PEAK ? __builtin_nanf("") : float(ix0 - i1);
      const float dy = code == PIV_WIN_NO_PEAK ? __builtin_nanf("") : float(iy0 - i0);
      res.dx[win] = dx + float(dx_in);
      res.dy[win] = dy + float(dy_in);
      res.correl_max[win] = cmax;
      res.err[win] = code;
      if (res.ix) res.ix[win] = i1;
      if (res.iy) res.iy[win] = i0;
      if (res.nsec) res.nsec[win] = 0;
    }
    return;
  }
  // secondary peaks (correl.py:250-270)
  int nsec = 0;
  if (pa.nb_peaks > 1) {
    int c0 = i0, c1 = i1;
    for (int k = 0; k < pa.nb_peaks - 1; ++k) {
      s.nan_square(lane, c0, c1, pa.particle_radius, nullptr);
      int j0, j1;
      const int c = s.find(lane, j0, j1);
      if (c != PIV_WIN_OK) break;
      c0 = j0;
      c1 = j1;
      if (writer) {
        const size_t o = size_t(win) * (pa.nb_peaks - 1) + k;
        res.sec[3 * o + 0] = float(ix0 - j1);
        res.sec[3 * o + 1] = float(iy0 - j0);
        res.sec[3 * o + 2] = peak_value(m[j0 * n1 + j1], norm);
        if (res.sec_sub) {
          double px, py;
          const int e =
              subpix(m, n0, n1, pa.subpix_method, pa.nsubpix, j1, j0, px, py, pa.map_scale);
          res.sec_sub[2 * o + 0] = e ? __builtin_nanf("") : float(ix0 - (j1 + px));
          res.sec_sub[2 * o + 1] = e ? __builtin_nanf("") : float(iy0 - (j0 + py));
        }
      }
      ++nsec;
    }
  }
  // the fit on the window's lanes, the records by the writer
  double px = 0.0, py = 0.0;
  int e = PIV_WIN_OK;
  if (apply_subpix)
    e = subpix_lanes<T>(m, n0, n1, pa.subpix_method, pa.nsubpix, i1, i0, lane, px, py,
                        pa.map_scale);
  if (writer) {
    float dx = float(ix0 - i1), dy = float(iy0 - i0);
    int err = PIV_WIN_OK;
    if (apply_subpix) {
      if (e == PIV_WIN_OK) {
        dx = float(ix0 - (i1 + px));
        dy = float(iy0 - (i0 + py));
      } else {
        err = e;
      }
    }
    res.dx[win] = dx + float(dx_in);
    res.dy[win] = dy + float(dy_in);
    res.correl_max[win] = cmax;
    res.err[win] = err;
    if (res.ix) res.ix[win] = i1;
    if (res.iy) res.iy[win] = i0;
    if (res.nsec) res.nsec[win] = nsec;
  }
}

}  // namespace fia

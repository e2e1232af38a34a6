// Direct (spatial) cross-correlation of interrogation windows (gfx950), the
// FFT-free backend of the reference's CorrelPythran / CorrelPyCuda
// (src/fluidimage/calcul/correl.py:284-451, correl_pycuda.py:26-240), and the
// peak search / sub-pixel fit on maps of any shape.
//
// For the displacement (dy, dx), |dy|, |dx| <= dmax, the map value is the sum
// of im1[y, x] im0[y + oy, x + ox] over the overlap of the two windows, with
// the reference's overlap bounds: for a non-positive displacement
// d (= dy or dx) on an axis of sizes n0 (im0) and n1 (im1), with
// c = n0 // 2 - n1 // 2 + d:
//   length n1 + min(c, 0), start in im1 -min(c, 0), start in im0 max(0, c);
// for a positive one:
//   length n1 - max(n0 // 2 + n1 // 2 + d - n0, 0), start in im1 0,
//   start in im0 n0 // 2 - n1 // 2 + d.
// Map element [dmax + dy, dmax + dx]. Normalisation "pythran" (correl_numpy):
// sum / (overlap area) * (n1 rows * n1 cols of im1), norm
// sqrt(sum im1^2 sum im0^2); "pycuda": the plain sum, norm
// sqrt(sum im1^2 sum im0^2) * size(im0). Sums and norms in FP64, the map
// value rounded once to float32.
//
// One workgroup per window; both windows sit in LDS; the threads share out
// the (2 dmax + 1)^2 map elements, each summing its overlap row by row.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/piv_hip.h"
#include "last_error.hpp"
#include "piv_device.hpp"

using namespace fia;

namespace {

constexpr int kDirectThreads = 256;
constexpr int kMaxDirectLdsFloats = 16384;  // both windows: 64 KB of LDS
constexpr int kMaxAnyMap = 127;             // maps of any shape up to 127 x 127

int direct_fail(int code, const std::string& msg) {
  fia::last_error() = msg;
  return code;
}

int direct_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return direct_fail(PIV_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return PIV_OK;
}

// overlap on one axis for the displacement d (reference bounds, see top)
__host__ __device__ __forceinline__ void overlap(int n0, int n1, int d, int& len, int& s1,
                                                 int& s0) {
  if (d <= 0) {
    const int c = n0 / 2 - n1 / 2 + d;
    len = n1 + (c < 0 ? c : 0);
    s1 = c < 0 ? -c : 0;
    s0 = c > 0 ? c : 0;
  } else {
    const int e = n0 / 2 + n1 / 2 + d - n0;
    len = n1 - (e > 0 ? e : 0);
    s1 = 0;
    s0 = n0 / 2 - n1 / 2 + d;
  }
}

__device__ double block_sum_d(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < kDirectThreads / 64; ++k) s += red[k];
  __syncthreads();
  return s;
}

__global__ __launch_bounds__(kDirectThreads) void k_correlate_direct(
    const float* __restrict__ im0, int n0y, int n0x, const float* __restrict__ im1, int n1y,
    int n1x, int dmax, int normalisation, float* __restrict__ maps, double* __restrict__ norms) {
  extern __shared__ float win[];  // im0 (n0y x n0x) then im1 (n1y x n1x)
  __shared__ double red[kDirectThreads / 64];
  const int w = blockIdx.x;
  const int s0 = n0y * n0x, s1 = n1y * n1x;
  float* a = win;
  float* b = win + s0;
  const float* g0 = im0 + size_t(w) * s0;
  const float* g1 = im1 + size_t(w) * s1;
  double e0 = 0.0, e1 = 0.0;
  for (int k = threadIdx.x; k < s0; k += kDirectThreads) {
    const float v = g0[k];
    a[k] = v;
    e0 += double(v) * double(v);
  }
  for (int k = threadIdx.x; k < s1; k += kDirectThreads) {
    const float v = g1[k];
    b[k] = v;
    e1 += double(v) * double(v);
  }
  e0 = block_sum_d(e0, red);  // (its barriers also publish the windows)
  e1 = block_sum_d(e1, red);
  const double rootn = sqrt(e1 * e0);
  if (threadIdx.x == 0)
    norms[w] = normalisation == PIV_DIRECT_PYCUDA ? rootn * double(s0) : rootn;
  const int nm = 2 * dmax + 1;
  float* out = maps + size_t(w) * nm * nm;
  for (int e = threadIdx.x; e < nm * nm; e += kDirectThreads) {
    const int dy = e / nm - dmax, dx = e % nm - dmax;
    int ly, y1, y0, lx, x1, x0;
    overlap(n0y, n1y, dy, ly, y1, y0);
    overlap(n0x, n1x, dx, lx, x1, x0);
    double acc = 0.0;
    for (int iy = 0; iy < ly; ++iy) {
      const float* rb = b + (y1 + iy) * n1x + x1;
      const float* ra = a + (y0 + iy) * n0x + x0;
      for (int ix = 0; ix < lx; ++ix) acc += double(rb[ix]) * double(ra[ix]);
    }
    const double v = normalisation == PIV_DIRECT_PYCUDA
                         ? acc
                         : acc / (double(lx) * double(ly)) * double(s1);
    out[e] = float(v);
  }
}

// peak search with second chance, secondary peaks and sub-pixel fit on
// n0 x n1 maps with the zero displacement at (iy0, ix0); no search mask, so
// the second-chance NaN square is written back into the map, as the
// reference's CorrelBase.compute_displacements_from_correl does to the
// caller's array (correl.py:226-246). One wavefront per map.
__global__ __launch_bounds__(64) void k_peaks_any(float* __restrict__ maps,
                                                  const double* __restrict__ norms, int n0,
                                                  int n1, int iy0, int ix0, int apply_subpix,
                                                  piv_results res, PlanArgs pa) {
  extern __shared__ float sm[];  // map then its search copy
  const int nn = n0 * n1;
  float* m = sm;
  float* srch = sm + nn;
  const int win = blockIdx.x;
  const int lane = threadIdx.x;
  float* g = maps + size_t(win) * nn;
  for (int k = lane; k < nn; k += 64) {
    const float v = g[k];
    m[k] = v;
    srch[k] = v;
  }
  __syncthreads();
  int sq[4];
  finish_window<64>(m, srch, n0, n1, iy0, ix0, pa, lane, win, norms[win], apply_subpix, 0, 0,
                    res, sq);
  for (int k = lane; k < nn; k += 64) {
    const int i0 = k / n1, i1 = k % n1;
    if (i0 >= sq[0] && i0 < sq[1] && i1 >= sq[2] && i1 < sq[3]) g[k] = __builtin_nanf("");
  }
}

__global__ void k_subpix_any(const float* __restrict__ maps, int n0, int n1,
                             const int32_t* __restrict__ ix, const int32_t* __restrict__ iy,
                             int nwin, int method, int nsub, float* out_x, float* out_y,
                             int32_t* err) {
  const int win = blockIdx.x * blockDim.x + threadIdx.x;
  if (win >= nwin) return;
  double px, py;
  const int x = ix[win], y = iy[win];
  const int e = subpix(maps + size_t(win) * n0 * n1, n0, n1, method, nsub, x, y, px, py);
  err[win] = e;
  out_x[win] = e ? float(x) : float(x + px);
  out_y[win] = e ? float(y) : float(y + py);
}

bool direct_shapes_ok(int n0y, int n0x, int n1y, int n1x, int dmax) {
  for (int d = -dmax; d <= dmax; ++d) {
    int l, s1, s0;
    overlap(n0y, n1y, d, l, s1, s0);
    if (l <= 0 || s1 < 0 || s0 < 0 || s1 + l > n1y || s0 + l > n0y) return false;
    overlap(n0x, n1x, d, l, s1, s0);
    if (l <= 0 || s1 < 0 || s0 < 0 || s1 + l > n1x || s0 + l > n0x) return false;
  }
  return true;
}

}  // namespace

extern "C" {

int piv_correlate_direct(int32_t nwin, const float* im0, int32_t n0y, int32_t n0x,
                         const float* im1, int32_t n1y, int32_t n1x, int32_t displacement_max,
                         int32_t normalisation, float* maps, double* norms, void* stream) {
  if (nwin <= 0) return PIV_OK;
  if (!im0 || !im1 || !maps || !norms) return direct_fail(PIV_EINVAL, "NULL argument");
  if (normalisation != PIV_DIRECT_PYTHRAN && normalisation != PIV_DIRECT_PYCUDA)
    return direct_fail(PIV_EINVAL, "unknown normalisation");
  if (displacement_max <= 0)
    return direct_fail(PIV_EINVAL, "displacement_max <= 0 : problem with images shapes?");
  if (2 * displacement_max + 1 > kMaxAnyMap)
    return direct_fail(PIV_EUNSUPPORTED, "direct maps are built up to 127 x 127");
  if (n0y <= 0 || n0x <= 0 || n1y <= 0 || n1x <= 0 ||
      int64_t(n0y) * n0x + int64_t(n1y) * n1x > kMaxDirectLdsFloats)
    return direct_fail(PIV_EUNSUPPORTED, "windows too large for the direct correlation");
  if (!direct_shapes_ok(n0y, n0x, n1y, n1x, displacement_max))
    return direct_fail(PIV_EINVAL,
                       "displacement_max too large for these window shapes (empty or "
                       "out-of-window overlap)");
  const size_t lds = size_t(n0y * n0x + n1y * n1x) * sizeof(float);
  hipLaunchKernelGGL(k_correlate_direct, dim3(nwin), dim3(kDirectThreads), lds,
                     static_cast<hipStream_t>(stream), im0, int(n0y), int(n0x), im1, int(n1y),
                     int(n1x), int(displacement_max), int(normalisation), maps, norms);
  return direct_check("piv_correlate_direct");
}

int piv_peaks_any(int32_t nwin, float* maps, const double* norms, int32_t n0, int32_t n1,
                  int32_t iy0, int32_t ix0, int32_t subpix_method, int32_t nsubpix,
                  int32_t particle_radius, int32_t nb_peaks, int32_t apply_subpix,
                  const piv_results* res, void* stream) {
  if (nwin <= 0) return PIV_OK;
  if (!maps || !norms || !res || !res->dx || !res->dy || !res->correl_max || !res->err)
    return direct_fail(PIV_EINVAL, "NULL argument");
  if (n0 < 3 || n1 < 3 || n0 > kMaxAnyMap || n1 > kMaxAnyMap)
    return direct_fail(PIV_EUNSUPPORTED, "maps of 3 x 3 to 127 x 127 only");
  if (subpix_method < 0 || subpix_method > PIV_SUBPIX_2D_GAUSSIAN || nsubpix < 1 ||
      particle_radius < 0 || nb_peaks < 1)
    return direct_fail(PIV_EINVAL, "bad search parameters");
  if (nb_peaks > 1 && (res->sec == nullptr || res->nsec == nullptr))
    return direct_fail(PIV_EINVAL, "secondary peak arrays required when nb_peaks > 1");
  PlanArgs pa{};
  pa.mask = nullptr;
  pa.b0s = 0;
  pa.b0e = n0;
  pa.b1s = 0;
  pa.b1e = n1;
  pa.subpix_method = subpix_method;
  pa.nsubpix = nsubpix;
  pa.particle_radius = particle_radius;
  pa.nb_peaks = nb_peaks;
  pa.has_dmax = 0;
  const size_t lds = size_t(2) * n0 * n1 * sizeof(float);
  hipLaunchKernelGGL(k_peaks_any, dim3(nwin), dim3(64), lds, static_cast<hipStream_t>(stream),
                     maps, norms, int(n0), int(n1), int(iy0), int(ix0), int(apply_subpix), *res,
                     pa);
  return direct_check("piv_peaks_any");
}

int piv_subpix_any(int32_t nwin, const float* maps, int32_t n0, int32_t n1, const int32_t* ix,
                   const int32_t* iy, int32_t subpix_method, int32_t nsubpix, float* out_x,
                   float* out_y, int32_t* err, void* stream) {
  if (nwin <= 0) return PIV_OK;
  if (!maps || !ix || !iy || !out_x || !out_y || !err)
    return direct_fail(PIV_EINVAL, "NULL argument");
  if (n0 < 1 || n1 < 1 || subpix_method < 0 || subpix_method > PIV_SUBPIX_2D_GAUSSIAN || nsubpix < 1)
    return direct_fail(PIV_EINVAL, "bad sub-pixel parameters");
  hipLaunchKernelGGL(k_subpix_any, dim3((nwin + 255) / 256), dim3(256), 0,
                     static_cast<hipStream_t>(stream), maps, int(n0), int(n1), ix, iy,
                     int(nwin), int(subpix_method), int(nsubpix), out_x, out_y, err);
  return direct_check("piv_subpix_any");
}

}  // extern "C"

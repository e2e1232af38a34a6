// libpiv_hip: MI355X (gfx950) kernels and C ABI of the PIV correlation path.
//
// Fused window kernel (one group of T lanes per interrogation window,
// several windows per workgroup):
//   HBM crop of both images (zero padding + min offset)
//   -> spatial mean removal (== the reference's DC-bin zeroing)
//   -> one complex 2-D FFT of z = im0' + i im1' (column pass in registers,
//      LDS transpose, row pass in registers)
//   -> split F0 / F1 with Z(-k), half-spectrum energies, cross spectrum
//      conj(F0) F1
//   -> inverse 2-D FFT (rows, LDS transpose, columns)
//   -> map in the reference layout (after _like_fftshift) in LDS
//   -> masked peak search, second chance, secondary peaks, sub-pixel fit.
// No global scratch: every intermediate stays in registers or LDS.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/piv_hip.h"
#include "fft_reg.hpp"
#include "last_error.hpp"
#include "piv_device.hpp"
#include "tuning.hpp"

using namespace fia;

namespace {

int fail(int code, const std::string& msg) {
  fia::last_error() = msg;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(PIV_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return PIV_OK;
}

// -------------------------------------------------------------------------
// fused correlation kernel
// -------------------------------------------------------------------------
// pixel fetch through a global-address-space pointer (no flat loads)
// Workgroups are dispatched to the 8 XCDs of the chip round-robin
// (workgroup b to XCD b % 8). Renumbering them so that each XCD runs one
// contiguous range of logical blocks keeps neighbouring windows, which
// share image rows, under the same L2. A bijection of [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  constexpr int kXcd = 8;
  const int x = b % kXcd, j = b / kXcd;
  const int q = nb / kXcd, r = nb % kXcd;
  return x * q + min(x, r) + j;
}

template <int DT>
__device__ __forceinline__ float gfetch(const void* data, int64_t idx) {
  if constexpr (DT == PIV_U8)
    return float(((const __attribute__((address_space(1))) uint8_t*)data)[idx]);
  else if constexpr (DT == PIV_U16)
    return float(((const __attribute__((address_space(1))) uint16_t*)data)[idx]);
  else
    return ((const __attribute__((address_space(1))) float*)data)[idx];
}

template <int DT>
struct PixelType;
template <>
struct PixelType<PIV_U8> {
  using type = uint8_t;
};
template <>
struct PixelType<PIV_U16> {
  using type = uint16_t;
};
template <>
struct PixelType<PIV_F32> {
  using type = float;
};

template <int N0, int N1>
struct Shape {
  static constexpr int T = lanes_for(N0 > N1 ? N0 : N1);
  static constexpr int NN = N0 * N1;
};

// 2 F0(k), 2 F1(k) of the two real windows from Z(k) = a and Z(-k) = (zmx,
// zmy) of their complex FFT: F0 = (Z(k) + conj Z(-k)) / 2, F1 = (Z(k) - conj
// Z(-k)) / 2i; zero at the DC bin (dc). The factor 2 of each spectrum makes
// the cross spectrum, the map and each energy 4 times theirs, and the norm
// too: every product and sum of the scaled values is the scaled result of
// the same operation exactly (powers of two), so the peaks, their ratios to
// the norm and the sub-pixel fits are bit for bit those of the unscaled
// values; the kept map and the norm are scaled back by 1/4 (exact).
struct SplitSpectra {
  float2 f0, f1;
};

__device__ __forceinline__ SplitSpectra split_spectra(float2 a, float zmx, float zmy, bool dc) {
  const float2 b = make_float2(zmx, -zmy);
  const float2 d = make_float2(a.x - b.x, a.y - b.y);
  SplitSpectra s;
  s.f0 = make_float2(a.x + b.x, a.y + b.y);
  s.f1 = make_float2(d.y, -d.x);
  if (dc) {
    s.f0 = make_float2(0.f, 0.f);
    s.f1 = make_float2(0.f, 0.f);
  }
  return s;
}

// element k of a row-major map of N1 columns lies in the square of rows
// [sq0, sq1), columns [sq2, sq3)
template <int N1>
__device__ __forceinline__ bool in_square(int k, const int* sq) {
  const int i0 = k / N1, i1 = k % N1;
  return i0 >= sq[0] && i0 < sq[1] && i1 >= sq[2] && i1 < sq[3];
}

// LDS layout of one component (real or imaginary parts) of the complex
// N0 x N1 buffer: element (r, c) lives at r * LD + c with an odd row stride
// LD. The transposes between the column and the row passes move the real
// parts, then the imaginary parts, so that no more than one full complex
// array per lane is live at a time and the buffer is N0 LD floats. The
// column pass (lanes over c) accesses consecutive elements; in the row pass
// (lanes over r) the odd stride puts the lanes' 32-bit accesses in distinct
// banks. Every access is one base register plus a compile-time offset. The
// buffer then holds the map (N0 N1 floats); the peak search reads the mask
// bits and its NaN squares directly (no search copy). A 64 x 64 window takes
// 17 KB of LDS, so that VGPRs, not LDS, bound the occupancy.
template <int N1>
constexpr int kLd = N1 % 2 == 0 ? N1 + 1 : N1;

// minimum wavefronts per SIMD the register allocation must allow: 2 for
// the 48-point transforms (whose 264 VGPRs would allow 1; 18 spilled
// registers cost less than the halved occupancy)
template <int N0, int N1>
constexpr int correlate_min_waves() {
  return (N0 == 48 || N1 == 48) ? 2 : 1;
}

template <int N0, int N1, int DT, int WPB>
__global__ __launch_bounds__((lanes_for(N0 > N1 ? N0 : N1) * WPB))
__attribute__((amdgpu_waves_per_eu(correlate_min_waves<N0, N1>()))) void k_correlate(
    const piv_image* __restrict__ images, const piv_window* __restrict__ windows, int nwin,
    int options, float* __restrict__ maps, piv_results res, PlanArgs pa) {
  constexpr int T = Shape<N0, N1>::T;
  const int apply_subpix = options & PIV_OPT_SUBPIX;
  constexpr int NN = N0 * N1;
  constexpr int LD = kLd<N1>;
  constexpr int NB = N0 * LD;  // floats of one window's buffer
  static_assert(NB >= NN, "the map (NN floats) fits in the buffer");
  __shared__ __align__(16) float smem[WPB * NB];
  __shared__ uint64_t scol[N1];  // bit i0 of scol[i1]: where_large_displacement[i0, i1]

  const int lane = threadIdx.x & (T - 1);
  const int wl = threadIdx.x / T;
  const int win = xcd_block(blockIdx.x, gridDim.x) * WPB + wl;
  static_assert(N0 <= 64, "column masks hold up to 64 rows");
  if (pa.mask != nullptr) {
    for (int c = threadIdx.x; c < N1; c += T * WPB) {
      uint64_t bits = 0;
      for (int i0 = 0; i0 < N0; ++i0)
        if (pa.mask[i0 * N1 + c]) bits |= uint64_t(1) << i0;
      scol[c] = bits;
    }
  }

  float* buf = smem + wl * NB;
  piv_window w;
  const bool live = win < nwin;
  if (live) {
    w = windows[win];
  } else {
    w.flags = PIV_WINFLAG_INVALID;
    w.dx_in = w.dy_in = 0;
  }
  const bool compute = live && !(w.flags & PIV_WINFLAG_INVALID);

  // ---- 1. crop (column `lane` of both windows) -------------------------
  float2 z[N0];
  float s0 = 0.f, s1 = 0.f;
  if (compute && lane < N1) {
    const piv_image im0 = images[w.image0];
    const piv_image im1 = images[w.image1];
    const float m0 = im0.minval ? *im0.minval : 0.f;
    const float m1 = im1.minval ? *im1.minval : 0.f;
    const bool inside = w.y0 >= 0 && w.y0 + N0 <= im0.height && w.x0 >= 0 &&
                        w.x0 + N1 <= im0.width && w.y1 >= 0 && w.y1 + N0 <= im1.height &&
                        w.x1 >= 0 && w.x1 + N1 <= im1.width;
    if (inside) {
      // interior window: no padding, one pointer walk per image
      using E = typename PixelType<DT>::type;
      const __attribute__((address_space(1))) E* pa0 =
          (const __attribute__((address_space(1))) E*)im0.data + int64_t(w.y0) * im0.pitch +
          w.x0 + lane;
      const __attribute__((address_space(1))) E* pb0 =
          (const __attribute__((address_space(1))) E*)im1.data + int64_t(w.y1) * im1.pitch +
          w.x1 + lane;
#pragma unroll
      for (int r = 0; r < N0; ++r) {
        const float a = float(pa0[int64_t(r) * im0.pitch]) - m0;
        const float b = float(pb0[int64_t(r) * im1.pitch]) - m1;
        z[r] = make_float2(a, b);
        s0 += a;
        s1 += b;
      }
    } else {
      const int xa = w.x0 + lane, xb = w.x1 + lane;
      const bool ina = xa >= 0 && xa < im0.width;
      const bool inb = xb >= 0 && xb < im1.width;
#pragma unroll
      for (int r = 0; r < N0; ++r) {
        const int ya = w.y0 + r, yb = w.y1 + r;
        float a = 0.f, b = 0.f;
        if (ina && ya >= 0 && ya < im0.height)
          a = gfetch<DT>(im0.data, int64_t(ya) * im0.pitch + xa) - m0;
        if (inb && yb >= 0 && yb < im1.height)
          b = gfetch<DT>(im1.data, int64_t(yb) * im1.pitch + xb) - m1;
        z[r] = make_float2(a, b);
        s0 += a;
        s1 += b;
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < N0; ++r) z[r] = make_float2(0.f, 0.f);
  }
  // ---- 2. mean removal ---------------------------------------------------
  s0 = group_sum(s0, T);
  s1 = group_sum(s1, T);
  if (lane < N1) {
    const float mean0 = s0 / float(NN), mean1 = s1 / float(NN);
#pragma unroll
    for (int r = 0; r < N0; ++r) {
      z[r].x -= mean0;
      z[r].y -= mean1;
    }
  }
  // ---- 3. forward 2-D FFT of z = im0' + i im1' -------------------------
  // (in place: the forward transforms leave frequency k at fft_pos(n, k),
  // the inverse ones take it there; the loops below index accordingly)
  fft_dif<N0, -1, 0, 1, N0>(z);
  float2 y[N1];
  const int r = lane;  // row owned in the row pass
  if (r >= N0) {
#pragma unroll
    for (int c = 0; c < N1; ++c) y[c] = make_float2(0.f, 0.f);
  }
  // transpose, real parts then imaginary parts
  sfor<0, 2>([&](auto part_) {
    constexpr int part = decltype(part_)::value;
    if (lane < N1) {
      sfor<0, N0>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        constexpr int pk = fft_pos(N0, k);
        buf[k * LD + lane] = part == 0 ? z[pk].x : z[pk].y;
      });
    }
    __syncthreads();
    if (r < N0) {
      const float* row = buf + r * LD;
#pragma unroll
      for (int c = 0; c < N1; ++c) {
        if constexpr (part == 0) y[c].x = row[c];
        else y[c].y = row[c];
      }
    }
    __syncthreads();
  });
  fft_dif<N1, -1, 0, 1, N1>(y);
  // ---- 4. F0, F1 from Z(k), Z(-k); energies; cross spectrum ---------------
  // Z(-r, -c) is element -c of row N0 - r, held by lane N0 - r of the group:
  // columns c and -c are done together (both read before either is written)
  float e0f = 0.f, e1f = 0.f;
  const int pr = (r == 0 || r >= N0) ? r : N0 - r;
  sfor<0, N1 / 2 + 1>([&](auto c_) {
    constexpr int c = decltype(c_)::value;
    constexpr int cm = c == 0 ? 0 : N1 - c;
    constexpr int pc = fft_pos(N1, c);
    constexpr int pcm = fft_pos(N1, cm);
    // partner of c: element -c of row -r; partner of -c: element c of row -r
    const float ax = __shfl(y[pcm].x, pr, T), ay = __shfl(y[pcm].y, pr, T);
    float bx = 0.f, by = 0.f;
    if constexpr (cm != c) {
      bx = __shfl(y[pc].x, pr, T);
      by = __shfl(y[pc].y, pr, T);
    }
    const SplitSpectra f = split_spectra(y[pc], ax, ay, r == 0 && c == 0);
    const float2 f0 = f.f0, f1 = f.f1;
    {
      const float wgt = (c == 0 || c == N1 / 2) ? 1.f : 2.f;
      e0f += wgt * (f0.x * f0.x + f0.y * f0.y);
      e1f += wgt * (f1.x * f1.x + f1.y * f1.y);
    }
    if constexpr (cm != c) {
      const SplitSpectra g = split_spectra(y[pcm], bx, by, false);
      const float2 g0 = g.f0, g1 = g.f1;
      y[pcm] = make_float2(g0.x * g1.x + g0.y * g1.y, g0.x * g1.y - g0.y * g1.x);
    }
    // conj(f0) * f1
    y[pc] = make_float2(f0.x * f1.x + f0.y * f1.y, f0.x * f1.y - f0.y * f1.x);
  });
  if (r >= N0) e0f = e1f = 0.f;
  const double e0 = group_sum(double(e0f), T);
  const double e1 = group_sum(double(e1f), T);
  const double norm = sqrt(e0 * e1);
  // ---- 5. inverse 2-D FFT (rows, then columns) ------------------------
  fft_dit<N1, +1, 0, 1, N1>(y);
  sfor<0, 2>([&](auto part_) {
    constexpr int part = decltype(part_)::value;
    if (r < N0) {
      float* row = buf + r * LD;
#pragma unroll
      for (int c = 0; c < N1; ++c) row[c] = part == 0 ? y[c].x : y[c].y;
    }
    __syncthreads();
    if (lane < N1) {
      sfor<0, N0>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        constexpr int pk = fft_pos(N0, k);
        if constexpr (part == 0) z[pk].x = buf[k * LD + lane];
        else z[pk].y = buf[k * LD + lane];
      });
    }
    __syncthreads();
  });
  // ---- 6. map in the reference layout: arr[a0, a1] -> out[g(a0), g(a1)],
  //         g(a) = (n - 1 - a + n // 2) mod n  (inverse of _like_fftshift)
  float* mapf = buf;
  const int g1 = (N1 - 1 - lane + N1 / 2) % N1;
  if constexpr (N0 % 2 == 0) {
    // the column transform's output x is real: with M = N0 / 2,
    // x[2m] + i x[2m + 1] = IFFT_M(E + i O)[m], E[k] = X[k] + X[k + M],
    // O[k] = (X[k] - X[k + M]) exp(2 pi i k / N0) (half the work of the
    // N0-point transform)
    constexpr int M = N0 / 2;
    float2 h[M];
    sfor<0, M>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      constexpr int pa = fft_pos(N0, k), pb = fft_pos(N0, k + M), ph = fft_pos(M, k);
      const float2 a = z[pa];
      const float2 b = z[pb];
      const float2 e = cadd(a, b);
      const float2 o = twmul<k, N0, +1>(csub(a, b));
      h[ph] = cadd(e, mul_si<+1>(o));
    });
    fft_dit<M, +1, 0, 1, M>(h);
    if (lane < N1) {
      sfor<0, M>([&](auto m_) {
        constexpr int m = decltype(m_)::value;
        constexpr int ga = (N0 - 1 - 2 * m + N0 / 2) % N0;
        constexpr int gb = (N0 - 2 - 2 * m + N0 / 2) % N0;
        mapf[ga * N1 + g1] = h[m].x;
        mapf[gb * N1 + g1] = h[m].y;
      });
    }
  } else {
    fft_dit<N0, +1, 0, 1, N0>(z);
    if (lane < N1) {
#pragma unroll
      for (int a0 = 0; a0 < N0; ++a0) {
        const int g0 = (N0 - 1 - a0 + N0 / 2) % N0;
        mapf[g0 * N1 + g1] = z[a0].x;
      }
    }
  }
  __syncthreads();
  // ---- 7. peak search, sub-pixel fit, records ---------------------------
  int sq[4] = {0, 0, 0, 0};
  if (compute) {
    if (lane == 0 && res.norm) res.norm[win] = 0.25 * norm;
    PlanArgs pl = pa;
    pl.map_scale = 0.25f;  // the LDS map is 4 times the reference's
    finish_window_masked<T>(mapf, pa.mask != nullptr ? scol : nullptr, N0, N1, N0 / 2 - 1,
                            N1 / 2 - 1, pl, lane, win, norm, apply_subpix, w.dx_in, w.dy_in,
                            res, sq);
  } else if (live && lane == 0) {
    const float nan = __builtin_nanf("");
    res.dx[win] = nan;
    res.dy[win] = nan;
    res.correl_max[win] = nan;
    res.err[win] = PIV_WIN_BAD_CROP;
    if (res.norm) res.norm[win] = nan;
    if (res.nsec) res.nsec[win] = 0;
  }
  if (maps != nullptr && live) {
    // the map, scaled back (x 1/4, exact); four elements per access when the
    // output is 16-byte aligned (the buffer always is: N0 LD % 4 == 0)
    float* out = maps + size_t(win) * NN;
    const bool side = !pa.has_dmax && (options & PIV_OPT_MAP_SQUARE);
    const float nan = __builtin_nanf("");
    if (NN % 4 == 0 && NB % 4 == 0 && (reinterpret_cast<uintptr_t>(maps) & 15) == 0) {
      for (int k = 4 * lane; k < NN; k += 4 * T) {
        float4 v = compute ? *reinterpret_cast<const float4*>(mapf + k) : make_float4(nan, nan, nan, nan);
        v.x *= 0.25f;
        v.y *= 0.25f;
        v.z *= 0.25f;
        v.w *= 0.25f;
        if (side) {
          if (in_square<N1>(k, sq)) v.x = nan;
          if (in_square<N1>(k + 1, sq)) v.y = nan;
          if (in_square<N1>(k + 2, sq)) v.z = nan;
          if (in_square<N1>(k + 3, sq)) v.w = nan;
        }
        *reinterpret_cast<float4*>(out + k) = v;
      }
    } else {
      for (int k = lane; k < NN; k += T) {
        float v = compute ? 0.25f * mapf[k] : nan;
        if (side && in_square<N1>(k, sq)) v = nan;
        out[k] = v;
      }
    }
  }
}

// peak search (+ sub-pixel) on maps given by the caller
template <int N0, int N1, int WPB, int T>
__global__ __launch_bounds__(T * WPB) void k_peaks(float* __restrict__ maps,
                                                   const double* __restrict__ norms,
                                                   int nwin, int apply_subpix,
                                                   piv_results res, PlanArgs pa) {
  constexpr int NN = N0 * N1;
  __shared__ float smem[WPB * 2 * NN];
  const int lane = threadIdx.x & (T - 1);
  const int wl = threadIdx.x / T;
  const int win = blockIdx.x * WPB + wl;
  float* m = smem + wl * 2 * NN;
  float* srch = m + NN;
  const bool live = win < nwin;
  if (live)
    for (int k = lane; k < NN; k += T) {
      const float v = maps[size_t(win) * NN + k];
      m[k] = v;
      srch[k] = (pa.mask != nullptr && pa.mask[k]) ? __builtin_nanf("") : v;
    }
  __syncthreads();
  if (!live) return;
  int sq[4];
  const double norm = norms[win];
  finish_window<T>(m, srch, N0, N1, N0 / 2 - 1, N1 / 2 - 1, pa, lane, win, norm, apply_subpix,
                   0, 0, res, sq);
  if (!pa.has_dmax) {
    for (int k = lane; k < NN; k += T) {
      const int i0 = k / N1, i1 = k % N1;
      if (i0 >= sq[0] && i0 < sq[1] && i1 >= sq[2] && i1 < sq[3])
        maps[size_t(win) * NN + k] = __builtin_nanf("");
    }
  }
}

template <int N0, int N1>
__global__ void k_subpix_at(const float* __restrict__ maps, const int32_t* __restrict__ ix,
                            const int32_t* __restrict__ iy, int nwin, float* out_x,
                            float* out_y, int32_t* err, PlanArgs pa) {
  const int win = blockIdx.x * blockDim.x + threadIdx.x;
  if (win >= nwin) return;
  double px, py;
  const int x = ix[win], y = iy[win];
  const int e = subpix(maps + size_t(win) * N0 * N1, N0, N1, pa.subpix_method, pa.nsubpix, x, y,
                               px, py);
  err[win] = e;
  out_x[win] = e ? float(x) : float(x + px);
  out_y[win] = e ? float(y) : float(y + py);
}

// -------------------------------------------------------------------------
// large windows (128 x 128): LDS-staged FFTs, one workgroup per window.
// A line of N = R1 R2 points is transformed in LDS in two register passes
// (R1-point FFTs of the R2 interleaved subsequences, twiddles, R2-point
// FFTs), which leaves frequency k = k1 + R1 k2 at position k1 R2 + k2; the
// inverse consumes that order and returns the natural one. The spectrum
// separation works on (k, -k) pairs, so it runs in place.
// -------------------------------------------------------------------------
constexpr int kBigThreads = 256;

template <int N>
struct BigFFT {
  static constexpr int R1 = N == 128 ? 16 : 8;
  static constexpr int R2 = N / R1;
  static constexpr int P = N + 1;  // padded row stride (float2) of the LDS buffer
  __device__ static int pos(int k) { return (k % R1) * R2 + k / R1; }
  __device__ static int freq(int p) { return p / R2 + R1 * (p % R2); }
};

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// forward line FFT in place: natural order in, position k1 R2 + k2 holds
// frequency k1 + R1 k2; tw[m] = exp(-2 pi i m / N)
template <int N>
__device__ void big_fft_fwd(float2* L, int st, const float2* tw) {
  constexpr int R1 = BigFFT<N>::R1, R2 = BigFFT<N>::R2;
  for (int j2 = 0; j2 < R2; ++j2) {
    float2 v[R1];
#pragma unroll
    for (int j1 = 0; j1 < R1; ++j1) v[j1] = L[(j1 * R2 + j2) * st];
    fft<R1, -1>(v);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) L[(k1 * R2 + j2) * st] = cmulf(v[k1], tw[j2 * k1]);
  }
  for (int k1 = 0; k1 < R1; ++k1) {
    float2 v[R2];
#pragma unroll
    for (int j2 = 0; j2 < R2; ++j2) v[j2] = L[(k1 * R2 + j2) * st];
    fft<R2, -1>(v);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) L[(k1 * R2 + k2) * st] = v[k2];
  }
}

// unnormalised inverse line FFT in place: permuted order in, natural out
template <int N>
__device__ void big_fft_inv(float2* L, int st, const float2* tw) {
  constexpr int R1 = BigFFT<N>::R1, R2 = BigFFT<N>::R2;
  for (int k1 = 0; k1 < R1; ++k1) {
    float2 v[R2];
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k2] = L[(k1 * R2 + k2) * st];
    fft<R2, +1>(v);
#pragma unroll
    for (int j2 = 0; j2 < R2; ++j2) {
      const float2 w = tw[j2 * k1];
      L[(k1 * R2 + j2) * st] = cmulf(v[j2], make_float2(w.x, -w.y));
    }
  }
  for (int j2 = 0; j2 < R2; ++j2) {
    float2 v[R1];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) v[k1] = L[(k1 * R2 + j2) * st];
    fft<R1, +1>(v);
#pragma unroll
    for (int j1 = 0; j1 < R1; ++j1) L[(j1 * R2 + j2) * st] = v[j1];
  }
}

__device__ __forceinline__ double big_block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < kBigThreads / 64; ++k) t += red[k];
  return t;
}

template <int N, int DT>
__global__ __launch_bounds__(kBigThreads) void k_correlate_big(
    const piv_image* __restrict__ images, const piv_window* __restrict__ windows, int nwin,
    int options, float* __restrict__ maps, piv_results res, PlanArgs pa) {
  using F = BigFFT<N>;
  constexpr int P = F::P, NN = N * N;
  static_assert(size_t(N) * P * sizeof(float2) >= 2 * size_t(NN) * sizeof(float),
                "the map and its search copy reuse the spectrum buffer");
  __shared__ float2 Z[N * P];
  __shared__ float2 tw[N];
  __shared__ double red[kBigThreads / 64];
  __shared__ int sqs[4];
  const int tid = threadIdx.x;
  const int win = xcd_block(blockIdx.x, gridDim.x);  // grid = nwin, one window per block
  const piv_window w = windows[win];
  if (w.flags & PIV_WINFLAG_INVALID) {  // uniform across the block
    const float nan = __builtin_nanf("");
    if (tid == 0) {
      res.dx[win] = nan;
      res.dy[win] = nan;
      res.correl_max[win] = nan;
      res.err[win] = PIV_WIN_BAD_CROP;
      if (res.norm) res.norm[win] = nan;
      if (res.nsec) res.nsec[win] = 0;
    }
    if (maps != nullptr)
      for (int k = tid; k < NN; k += kBigThreads) maps[size_t(win) * NN + k] = nan;
    return;
  }
  for (int m = tid; m < N; m += kBigThreads) {
    double sn, cs;
    sincospi(-2.0 * m / N, &sn, &cs);
    tw[m] = make_float2(float(cs), float(sn));
  }
  // ---- 1. crop (zero outside the image) and mean removal -----------------
  const piv_image im0 = images[w.image0];
  const piv_image im1 = images[w.image1];
  const float m0 = im0.minval ? *im0.minval : 0.f;
  const float m1 = im1.minval ? *im1.minval : 0.f;
  float s0 = 0.f, s1 = 0.f;
  for (int e = tid; e < NN; e += kBigThreads) {
    const int r = e / N, c = e % N;
    const int ya = w.y0 + r, xa = w.x0 + c, yb = w.y1 + r, xb = w.x1 + c;
    float a = 0.f, b = 0.f;
    if (ya >= 0 && ya < im0.height && xa >= 0 && xa < im0.width)
      a = fetch_px<DT>(im0.data, int64_t(ya) * im0.pitch + xa) - m0;
    if (yb >= 0 && yb < im1.height && xb >= 0 && xb < im1.width)
      b = fetch_px<DT>(im1.data, int64_t(yb) * im1.pitch + xb) - m1;
    Z[r * P + c] = make_float2(a, b);
    s0 += a;
    s1 += b;
  }
  const float mean0 = float(big_block_sum(double(s0), red)) / float(NN);
  const float mean1 = float(big_block_sum(double(s1), red)) / float(NN);
  for (int e = tid; e < NN; e += kBigThreads) {
    float2& z = Z[(e / N) * P + e % N];
    z.x -= mean0;
    z.y -= mean1;
  }
  __syncthreads();
  // ---- 2. forward 2-D FFT of im0' + i im1' (columns, then rows) ----------
  if (tid < N) big_fft_fwd<N>(Z + tid, P, tw);
  __syncthreads();
  if (tid < N) big_fft_fwd<N>(Z + tid * P, 1, tw);
  __syncthreads();
  // ---- 3. F0, F1 from Z(k), Z(-k); energies; cross spectrum (in place, by
  //         (k, -k) pairs: cross(-k) = conj(cross(k))) ----------------------
  double e0 = 0.0, e1 = 0.0;
  for (int e = tid; e < NN; e += kBigThreads) {
    const int pr = e / N, pc = e % N;
    const int kr = F::freq(pr), kc = F::freq(pc);
    const int mr = kr == 0 ? 0 : N - kr, mc = kc == 0 ? 0 : N - kc;
    const int qr = F::pos(mr), qc = F::pos(mc);
    const int e2 = qr * N + qc;
    if (e2 < e) continue;
    const float2 a = Z[pr * P + pc];
    const float2 zm = Z[qr * P + qc];
    const float2 b = make_float2(zm.x, -zm.y);
    float2 f0 = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
    const float2 d = make_float2(a.x - b.x, a.y - b.y);
    float2 f1 = make_float2(0.5f * d.y, -0.5f * d.x);
    if (kr == 0 && kc == 0) {
      f0 = make_float2(0.f, 0.f);
      f1 = make_float2(0.f, 0.f);
    }
    const float p0 = f0.x * f0.x + f0.y * f0.y, p1 = f1.x * f1.x + f1.y * f1.y;
    if (kc <= N / 2) {
      const float wgt = (kc == 0 || kc == N / 2) ? 1.f : 2.f;
      e0 += double(wgt * p0);
      e1 += double(wgt * p1);
    }
    const float2 cr = make_float2(f0.x * f1.x + f0.y * f1.y, f0.x * f1.y - f0.y * f1.x);
    Z[pr * P + pc] = cr;
    if (e2 != e) {
      if (mc <= N / 2) {
        const float wgt = (mc == 0 || mc == N / 2) ? 1.f : 2.f;
        e0 += double(wgt * p0);
        e1 += double(wgt * p1);
      }
      Z[qr * P + qc] = make_float2(cr.x, -cr.y);
    }
  }
  e0 = big_block_sum(e0, red);
  e1 = big_block_sum(e1, red);
  const double norm = sqrt(e0 * e1);
  // ---- 4. inverse 2-D FFT (rows, then columns) -----------------------------
  if (tid < N) big_fft_inv<N>(Z + tid * P, 1, tw);
  __syncthreads();
  if (tid < N) big_fft_inv<N>(Z + tid, P, tw);
  __syncthreads();
  // ---- 5. map in the reference layout (g(a) = (n - 1 - a + n // 2) mod n)
  //         and its search copy, in the same LDS --------------------------
  float vals[NN / kBigThreads];
#pragma unroll
  for (int i = 0; i < NN / kBigThreads; ++i) {
    const int e = tid + i * kBigThreads;
    vals[i] = Z[(e / N) * P + e % N].x;
  }
  __syncthreads();
  float* mapf = reinterpret_cast<float*>(Z);
  float* srch = mapf + NN;
#pragma unroll
  for (int i = 0; i < NN / kBigThreads; ++i) {
    const int e = tid + i * kBigThreads;
    const int g0 = (N - 1 - e / N + N / 2) % N, g1 = (N - 1 - e % N + N / 2) % N;
    const int k = g0 * N + g1;
    mapf[k] = vals[i];
    srch[k] = (pa.mask != nullptr && pa.mask[k]) ? __builtin_nanf("") : vals[i];
  }
  __syncthreads();
  // ---- 6. peak search and sub-pixel fit by wavefront 0 -------------------
  if (tid < 64) {
    int sq[4] = {0, 0, 0, 0};
    if (tid == 0 && res.norm) res.norm[win] = norm;
    finish_window<64>(mapf, srch, N, N, N / 2 - 1, N / 2 - 1, pa, tid, win, norm,
                      options & PIV_OPT_SUBPIX, w.dx_in, w.dy_in, res, sq);
    if (tid == 0) {
      sqs[0] = sq[0];
      sqs[1] = sq[1];
      sqs[2] = sq[2];
      sqs[3] = sq[3];
    }
  }
  __syncthreads();
  if (maps != nullptr) {
    float* out = maps + size_t(win) * NN;
    const bool side = !pa.has_dmax && (options & PIV_OPT_MAP_SQUARE);
    for (int k = tid; k < NN; k += kBigThreads) {
      float v = mapf[k];
      if (side) {
        const int i0 = k / N, i1 = k % N;
        if (i0 >= sqs[0] && i0 < sqs[1] && i1 >= sqs[2] && i1 < sqs[3]) v = __builtin_nanf("");
      }
      out[k] = v;
    }
  }
}

// -------------------------------------------------------------------------
// windows of any other shape
// -------------------------------------------------------------------------
// The reference takes any window shape through FFTW (fft.py:205-244,
// correl.py:596-662). Shapes without a register-FFT kernel run here: one
// workgroup of 256 threads per window, the window's complex buffer
// z = im0' + i im1' in LDS (8 n0 n1 bytes), transformed by separable DFTs
// (each output a sum over its line with a double-precision twiddle table,
// accumulated in double, stored as float like an FFT stage), then the same
// steps as the fused kernel: F0 / F1 from Z(k) and Z(-k), DC bins zero,
// half-spectrum energies with the reference's column weights, conj(F0) F1,
// inverse DFTs, the map in the _like_fftshift layout, masked peak search,
// second chance, secondary peaks and sub-pixel fit. Values are unscaled (no
// factor 2 on the spectra).
constexpr int kAnyThreads = 256;
constexpr int kAnyMaxLen = 256;       // rows or columns of a window
constexpr int kAnyMaxElems = 16384;   // n0 n1: 128 KB of complex buffer
constexpr int kAnyChunk = 8;          // outputs per thread in one DFT step

// shapes served by k_correlate_any
bool any_shape_ok(int n0, int n1) {
  return n0 >= 2 && n1 >= 2 && n0 <= kAnyMaxLen && n1 <= kAnyMaxLen && n0 * n1 <= kAnyMaxElems;
}

// the complex buffer, then the twiddle tables at a 16-byte boundary
size_t any_tw_offset(int n0, int n1) { return (size_t(8) * n0 * n1 + 15) / 16 * 16; }
size_t any_lds_bytes(int n0, int n1) { return any_tw_offset(n0, n1) + size_t(16) * (n0 + n1); }

// tw[k] = (cos, sin)(2 pi k / n)
__device__ __forceinline__ void any_twiddles(double2* tw, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    double s, c;
    sincospi(2.0 * double(k) / double(n), &s, &c);
    tw[k] = make_double2(c, s);
  }
}

// In-place DFTs of `nlines` lines of `len` elements (element e of line l at
// buf[l * lstride + e * estride]): out[k] = sum_e x[e] exp(sign 2 pi i e k /
// len). Lines are done in chunks of at most kAnyThreads * kAnyChunk
// outputs whose inputs are all read before any output is written.
__device__ void any_dft_lines(float2* buf, int nlines, int len, int lstride, int estride,
                              const double2* tw, double sign) {
  const int tid = threadIdx.x;
  const int lines_per_chunk = max(1, (kAnyThreads * kAnyChunk) / len);
  for (int l0 = 0; l0 < nlines; l0 += lines_per_chunk) {
    const int nout = (min(nlines, l0 + lines_per_chunk) - l0) * len;
    float2 out[kAnyChunk];
#pragma unroll
    for (int j = 0; j < kAnyChunk; ++j) {
      const int o = tid + j * kAnyThreads;
      out[j] = make_float2(0.f, 0.f);
      if (o < nout) {
        const int line = l0 + o / len, k = o % len;
        const float2* x = buf + line * lstride;
        double ar = 0.0, ai = 0.0;
        int idx = 0;
        for (int e = 0; e < len; ++e) {
          const float2 v = x[e * estride];
          const double2 w = tw[idx];
          const double wi = sign * w.y;
          ar += double(v.x) * w.x - double(v.y) * wi;
          ai += double(v.x) * wi + double(v.y) * w.x;
          idx += k;
          if (idx >= len) idx -= len;
        }
        out[j] = make_float2(float(ar), float(ai));
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kAnyChunk; ++j) {
      const int o = tid + j * kAnyThreads;
      if (o < nout) buf[(l0 + o / len) * lstride + (o % len) * estride] = out[j];
    }
    __syncthreads();
  }
}

__device__ __forceinline__ double any_block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kAnyThreads / 64; ++w) s += red[w];
  return s;
}

template <int DT>
__global__ __launch_bounds__(kAnyThreads) void k_correlate_any(
    int n0, int n1, const piv_image* __restrict__ images, const piv_window* __restrict__ windows,
    int nwin, int options, float* __restrict__ maps, piv_results res, PlanArgs pa) {
  extern __shared__ __align__(16) unsigned char any_smem[];
  const int nn = n0 * n1;
  float2* buf = reinterpret_cast<float2*>(any_smem);
  double2* tw0 = reinterpret_cast<double2*>(any_smem + (size_t(8) * nn + 15) / 16 * 16);
  double2* tw1 = tw0 + n0;
  __shared__ double red[kAnyThreads / 64];
  __shared__ int sqs[4];
  const int tid = threadIdx.x;
  const int win = blockIdx.x;
  if (win >= nwin) return;
  const piv_window w = windows[win];
  const bool compute = !(w.flags & PIV_WINFLAG_INVALID);
  const int apply_subpix = options & PIV_OPT_SUBPIX;
  float* fl = reinterpret_cast<float*>(any_smem);
  float* mapf = fl + nn;  // the map in the reference layout (after the DFTs)
  float* srch = fl;       // its search copy, masked elements NaN
  double norm = 0.0;
  if (compute) {
    any_twiddles(tw0, n0);
    any_twiddles(tw1, n1);
    // ---- crop (zero padding, min offset) and float32 mean removal ----
    const piv_image im0 = images[w.image0];
    const piv_image im1 = images[w.image1];
    const float m0 = im0.minval ? *im0.minval : 0.f;
    const float m1 = im1.minval ? *im1.minval : 0.f;
    float s0 = 0.f, s1 = 0.f;
    for (int k = tid; k < nn; k += kAnyThreads) {
      const int r = k / n1, c = k % n1;
      const int ya = w.y0 + r, xa = w.x0 + c, yb = w.y1 + r, xb = w.x1 + c;
      float a = 0.f, b = 0.f;
      if (ya >= 0 && ya < im0.height && xa >= 0 && xa < im0.width)
        a = gfetch<DT>(im0.data, int64_t(ya) * im0.pitch + xa) - m0;
      if (yb >= 0 && yb < im1.height && xb >= 0 && xb < im1.width)
        b = gfetch<DT>(im1.data, int64_t(yb) * im1.pitch + xb) - m1;
      buf[k] = make_float2(a, b);
      s0 += a;
      s1 += b;
    }
    const float mean0 = float(any_block_sum(double(s0), red)) / float(nn);
    const float mean1 = float(any_block_sum(double(s1), red)) / float(nn);
    for (int k = tid; k < nn; k += kAnyThreads) {
      buf[k].x -= mean0;
      buf[k].y -= mean1;
    }
    __syncthreads();
    // ---- forward 2-D DFT of z (rows, then columns) ----
    any_dft_lines(buf, n0, n1, n1, 1, tw1, -1.0);
    any_dft_lines(buf, n1, n0, 1, n1, tw0, -1.0);
    // ---- F0, F1 from Z(k), Z(-k); energies; cross spectrum ----
    double e0 = 0.0, e1 = 0.0;
    const int half = n1 / 2;
    for (int k = tid; k < nn; k += kAnyThreads) {
      const int k0 = k / n1, k1 = k % n1;
      const int q0 = k0 ? n0 - k0 : 0, q1 = k1 ? n1 - k1 : 0;
      const int km = q0 * n1 + q1;
      if (km < k) continue;  // done with its partner
      const float2 a = buf[k], b = buf[km];
      double f0r = 0.5 * (double(a.x) + double(b.x)), f0i = 0.5 * (double(a.y) - double(b.y));
      double f1r = 0.5 * (double(a.y) + double(b.y)), f1i = -0.5 * (double(a.x) - double(b.x));
      if (k == 0) f0r = f0i = f1r = f1i = 0.0;  // fft_im[0, 0] = 0 (correl.py:655,657)
      const double p0 = f0r * f0r + f0i * f0i, p1 = f1r * f1r + f1i * f1i;
      // rfft2 half spectrum: columns 0..n1 // 2, the first and last weighted 1
      // (fft.py:72-90), |F(-k)| = |F(k)|
      if (k1 <= half) {
        const double wgt = (k1 == 0 || k1 == half) ? 1.0 : 2.0;
        e0 += wgt * p0;
        e1 += wgt * p1;
      }
      if (km != k && q1 <= half) {
        const double wgt = (q1 == 0 || q1 == half) ? 1.0 : 2.0;
        e0 += wgt * p0;
        e1 += wgt * p1;
      }
      const double xr = f0r * f1r + f0i * f1i, xi = f0r * f1i - f0i * f1r;
      buf[k] = make_float2(float(xr), float(xi));
      if (km != k) buf[km] = make_float2(float(xr), float(-xi));
    }
    e0 = any_block_sum(e0, red);
    e1 = any_block_sum(e1, red);
    norm = sqrt(e0 * e1);
    __syncthreads();
    // ---- inverse 2-D DFT (rows, then columns): real map ----
    any_dft_lines(buf, n0, n1, n1, 1, tw1, +1.0);
    any_dft_lines(buf, n1, n0, 1, n1, tw0, +1.0);
    // real parts to floats [0, nn) (chunks read before written: float k
    // comes from float 2k >= k)
    for (int k0 = 0; k0 < nn; k0 += kAnyThreads * kAnyChunk) {
      float v[kAnyChunk];
#pragma unroll
      for (int j = 0; j < kAnyChunk; ++j) {
        const int k = k0 + tid + j * kAnyThreads;
        v[j] = k < nn ? buf[k].x : 0.f;
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kAnyChunk; ++j) {
        const int k = k0 + tid + j * kAnyThreads;
        if (k < nn) fl[k] = v[j];
      }
      __syncthreads();
    }
    // arr[a0, a1] -> out[g(a0), g(a1)], g(a) = (n - 1 - a + n // 2) mod n
    for (int k = tid; k < nn; k += kAnyThreads) {
      const int a0 = k / n1, a1 = k % n1;
      const int g0 = (n0 - 1 - a0 + n0 / 2) % n0, g1 = (n1 - 1 - a1 + n1 / 2) % n1;
      mapf[g0 * n1 + g1] = fl[k];
    }
    __syncthreads();
    for (int k = tid; k < nn; k += kAnyThreads)
      srch[k] = (pa.mask != nullptr && pa.mask[k]) ? __builtin_nanf("") : mapf[k];
    __syncthreads();
  }
  // ---- peak search, sub-pixel fit, records (one wavefront) ----
  if (tid < 64) {
    int sq[4] = {0, 0, 0, 0};
    if (compute) {
      if (tid == 0 && res.norm) res.norm[win] = norm;
      finish_window<64>(mapf, srch, n0, n1, n0 / 2 - 1, n1 / 2 - 1, pa, tid, win, norm,
                        apply_subpix, w.dx_in, w.dy_in, res, sq);
    } else if (tid == 0) {
      const float nan = __builtin_nanf("");
      res.dx[win] = nan;
      res.dy[win] = nan;
      res.correl_max[win] = nan;
      res.err[win] = PIV_WIN_BAD_CROP;
      if (res.norm) res.norm[win] = nan;
      if (res.nsec) res.nsec[win] = 0;
    }
    if (tid == 0)
      for (int i = 0; i < 4; ++i) sqs[i] = sq[i];
  }
  __syncthreads();
  if (maps != nullptr) {
    float* out = maps + size_t(win) * nn;
    const bool side = !pa.has_dmax && (options & PIV_OPT_MAP_SQUARE);
    for (int k = tid; k < nn; k += kAnyThreads) {
      float v = compute ? mapf[k] : __builtin_nanf("");
      if (side) {
        const int i0 = k / n1, i1 = k % n1;
        if (i0 >= sqs[0] && i0 < sqs[1] && i1 >= sqs[2] && i1 < sqs[3]) v = __builtin_nanf("");
      }
      out[k] = v;
    }
  }
}

// peak search (+ sub-pixel) on caller maps of any shape: one wavefront per
// map, the map and its masked search copy in LDS
__global__ __launch_bounds__(64) void k_peaks_generic(float* __restrict__ maps,
                                                      const double* __restrict__ norms, int n0,
                                                      int n1, int nwin, int apply_subpix,
                                                      piv_results res, PlanArgs pa) {
  extern __shared__ __align__(16) unsigned char any_smem[];
  const int nn = n0 * n1;
  float* m = reinterpret_cast<float*>(any_smem);
  float* srch = m + nn;
  const int win = blockIdx.x;
  const int lane = threadIdx.x;
  if (win >= nwin) return;
  float* g = maps + size_t(win) * nn;
  for (int k = lane; k < nn; k += 64) {
    const float v = g[k];
    m[k] = v;
    srch[k] = (pa.mask != nullptr && pa.mask[k]) ? __builtin_nanf("") : v;
  }
  __syncthreads();
  int sq[4];
  finish_window<64>(m, srch, n0, n1, n0 / 2 - 1, n1 / 2 - 1, pa, lane, win, norms[win],
                    apply_subpix, 0, 0, res, sq);
  if (!pa.has_dmax) {
    for (int k = lane; k < nn; k += 64) {
      const int i0 = k / n1, i1 = k % n1;
      if (i0 >= sq[0] && i0 < sq[1] && i1 >= sq[2] && i1 < sq[3]) g[k] = __builtin_nanf("");
    }
  }
}

__global__ void k_subpix_generic(const float* __restrict__ maps, int n0, int n1,
                                 const int32_t* __restrict__ ix, const int32_t* __restrict__ iy,
                                 int nwin, float* out_x, float* out_y, int32_t* err,
                                 PlanArgs pa) {
  const int win = blockIdx.x * blockDim.x + threadIdx.x;
  if (win >= nwin) return;
  double px, py;
  const int x = ix[win], y = iy[win];
  const int e = subpix(maps + size_t(win) * n0 * n1, n0, n1, pa.subpix_method, pa.nsubpix, x, y,
                       px, py);
  err[win] = e;
  out_x[win] = e ? float(x) : float(x + px);
  out_y[win] = e ? float(y) : float(y + py);
}

// -------------------------------------------------------------------------
// shape dispatch
// -------------------------------------------------------------------------
// A shape's fused correlation kernel comes in up to kMaxVariants launch
// variants that differ only in the windows per workgroup (WPB): every
// window is computed by the same T lanes with the same instructions, so the
// variant changes the speed, not the results. The variant in use is a
// tunable (piv_tune_set), picked per shape by the autotuner.
constexpr int kMaxVariants = 4;

using CorrelateFn = void (*)(dim3, dim3, hipStream_t, const piv_image*, const piv_window*, int,
                             int, float*, piv_results, PlanArgs);

struct CorrVariant {
  int wpb, threads;  // windows and threads per workgroup (wpb 0: no variant)
  CorrelateFn correlate[3];
};

struct Launchers {
  int n0, n1;
  CorrVariant var[kMaxVariants];  // var[0] is the default
  void (*peaks)(dim3, dim3, hipStream_t, float*, const double*, int, int, piv_results, PlanArgs);
  void (*subpix)(dim3, dim3, hipStream_t, const float*, const int32_t*, const int32_t*, int,
                 float*, float*, int32_t*, PlanArgs);
  int pwpb, pthreads;  // peaks from maps: windows and threads per workgroup
};

template <int N0, int N1>
constexpr int default_wpb() {
  constexpr int T = Shape<N0, N1>::T;
  int w = 256 / T;
  while (w > 1 && w * N0 * N1 * 8 > 65536) --w;
  return w;
}

// the other WPB candidates of a tuned shape: a quarter, half and twice the
// default, within 1024 threads and 80 KB of LDS per workgroup (two
// workgroups per CU); 0 when there is none (or it repeats another)
template <int N0, int N1>
constexpr int wpb_half() {
  return default_wpb<N0, N1>() > 1 ? default_wpb<N0, N1>() / 2 : 0;
}

template <int N0, int N1>
constexpr int wpb_quarter() {
  return default_wpb<N0, N1>() >= 4 ? default_wpb<N0, N1>() / 4 : 0;
}

template <int N0, int N1>
constexpr int wpb_double() {
  constexpr int w = 2 * default_wpb<N0, N1>();
  return (Shape<N0, N1>::T * w <= 1024 && w * N0 * N1 * 8 <= 80 * 1024) ? w : 0;
}

template <int N0, int N1, int DT, int WPB>
void launch_correlate(dim3 g, dim3 b, hipStream_t s, const piv_image* im, const piv_window* wi,
                      int n, int sp, float* maps, piv_results r, PlanArgs pa) {
  hipLaunchKernelGGL((k_correlate<N0, N1, DT, WPB>), g, b, 0, s, im, wi, n, sp, maps, r, pa);
}

template <int N0, int N1, int WPB>
constexpr CorrVariant make_variant() {
  if constexpr (WPB == 0) {
    return CorrVariant{0, 0, {nullptr, nullptr, nullptr}};
  } else {
    return CorrVariant{WPB,
                       Shape<N0, N1>::T * WPB,
                       {&launch_correlate<N0, N1, PIV_U8, WPB>,
                        &launch_correlate<N0, N1, PIV_U16, WPB>,
                        &launch_correlate<N0, N1, PIV_F32, WPB>}};
  }
}

template <int N, int DT>
void launch_correlate_big(dim3 g, dim3 b, hipStream_t s, const piv_image* im,
                          const piv_window* wi, int n, int sp, float* maps, piv_results r,
                          PlanArgs pa) {
  hipLaunchKernelGGL((k_correlate_big<N, DT>), g, b, 0, s, im, wi, n, sp, maps, r, pa);
}

// the search of one window runs on at most one wavefront
template <int N0, int N1>
constexpr int search_lanes() {
  return Shape<N0, N1>::T < 64 ? Shape<N0, N1>::T : 64;
}

template <int N0, int N1>
constexpr int peaks_wpb() {
  constexpr int T = search_lanes<N0, N1>();
  int w = 256 / T;
  while (w > 1 && w * N0 * N1 * 8 > 65536) --w;
  return w;
}

template <int N0, int N1>
void launch_peaks(dim3 g, dim3 b, hipStream_t s, float* maps, const double* norms, int n, int sp,
                  piv_results r, PlanArgs pa) {
  hipLaunchKernelGGL((k_peaks<N0, N1, peaks_wpb<N0, N1>(), search_lanes<N0, N1>()>), g, b, 0, s,
                     maps, norms, n, sp, r, pa);
}

template <int N0, int N1>
void launch_subpix(dim3 g, dim3 b, hipStream_t s, const float* maps, const int32_t* ix,
                   const int32_t* iy, int n, float* ox, float* oy, int32_t* e, PlanArgs pa) {
  hipLaunchKernelGGL((k_subpix_at<N0, N1>), g, b, 0, s, maps, ix, iy, n, ox, oy, e, pa);
}

// Tuned: the window sizes of the multipass schedules (16, 24, 32, 48, 64)
// get the extra WPB variants; the others only the default
template <int N0, int N1, bool Tuned = false>
constexpr Launchers make_launchers() {
  return Launchers{N0,
                   N1,
                   {make_variant<N0, N1, default_wpb<N0, N1>()>(),
                    make_variant<N0, N1, Tuned ? wpb_half<N0, N1>() : 0>(),
                    make_variant<N0, N1, Tuned ? wpb_double<N0, N1>() : 0>(),
                    make_variant<N0, N1, Tuned ? wpb_quarter<N0, N1>() : 0>()},
                   &launch_peaks<N0, N1>,
                   &launch_subpix<N0, N1>,
                   peaks_wpb<N0, N1>(),
                   search_lanes<N0, N1>() * peaks_wpb<N0, N1>()};
}

template <int N>
constexpr Launchers make_launchers_big() {
  return Launchers{N,
                   N,
                   {CorrVariant{1,
                                kBigThreads,
                                {&launch_correlate_big<N, PIV_U8>, &launch_correlate_big<N, PIV_U16>,
                                 &launch_correlate_big<N, PIV_F32>}},
                    make_variant<N, N, 0>(), make_variant<N, N, 0>(), make_variant<N, N, 0>()},
                   &launch_peaks<N, N>,
                   &launch_subpix<N, N>,
                   peaks_wpb<N, N>(),
                   search_lanes<N, N>() * peaks_wpb<N, N>()};
}

const Launchers kShapes[] = {
    make_launchers<12, 12>(),       make_launchers<16, 16, true>(), make_launchers<20, 20>(),
    make_launchers<24, 24, true>(), make_launchers<32, 32, true>(), make_launchers<40, 40>(),
    make_launchers<48, 48, true>(), make_launchers<64, 64, true>(), make_launchers<64, 32>(),
    make_launchers<32, 64>(),       make_launchers_big<128>(),
};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

// variant in use per shape (index into Launchers::var)
std::atomic<int> g_variant[kNumShapes];

int shape_index(int n0, int n1) {
  for (int i = 0; i < kNumShapes; ++i)
    if (kShapes[i].n0 == n0 && kShapes[i].n1 == n1) return i;
  return -1;
}

const Launchers* find_shape(int n0, int n1) {
  const int i = shape_index(n0, n1);
  return i < 0 ? nullptr : &kShapes[i];
}

const CorrVariant& variant_of(const Launchers* l) {
  const int v = g_variant[l - kShapes].load(std::memory_order_relaxed);
  return l->var[v];
}

// -------------------------------------------------------------------------
// image minimum (im - im.min(), singlepass.py:206)
// -------------------------------------------------------------------------
__device__ __forceinline__ uint32_t float_key(float f) {
  uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
  uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(b);
}

__global__ void k_min_init(uint32_t* keys, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keys[i] = 0xffffffffu;
}

template <int DT>
__device__ __forceinline__ float row_min(const piv_image& im, int y, float m) {
  const int64_t base = int64_t(y) * im.pitch;
  for (int x = threadIdx.x; x < im.width; x += blockDim.x)
    m = fminf(m, gfetch<DT>(im.data, base + x));
  return m;
}

// rows blockIdx.x, blockIdx.x + gridDim.x, ... of image blockIdx.y
__global__ __launch_bounds__(256) void k_min_reduce(const piv_image* __restrict__ images,
                                                    uint32_t* keys) {
  const piv_image im = images[blockIdx.y];
  float m = INFINITY;
  for (int y = blockIdx.x; y < im.height; y += gridDim.x) {
    if (im.dtype == PIV_U8) m = row_min<PIV_U8>(im, y, m);
    else if (im.dtype == PIV_U16) m = row_min<PIV_U16>(im, y, m);
    else m = row_min<PIV_F32>(im, y, m);
  }
  for (int off = 32; off > 0; off >>= 1) m = fminf(m, __shfl_xor(m, off, 64));
  __shared__ float part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float b = fminf(fminf(part[0], part[1]), fminf(part[2], part[3]));
    atomicMin(&keys[blockIdx.y], float_key(b));
  }
}

// contiguous images whose data are 16-byte aligned: 16-byte loads over the
// flat image (8 uint16, 16 uint8 or 4 float32 per load), four loads in
// flight per thread, blocks blockIdx.x, + gridDim.x, ... of image
// blockIdx.y; the last (size mod 16) bytes by scalar loads
template <int DT>
__device__ __forceinline__ float min16(uint4 q) {
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
  float m = INFINITY;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if constexpr (DT == PIV_U8) {
      const uint32_t a = min(min(w[k] & 0xffu, (w[k] >> 8) & 0xffu),
                             min((w[k] >> 16) & 0xffu, w[k] >> 24));
      m = fminf(m, float(a));
    } else if constexpr (DT == PIV_U16) {
      m = fminf(m, float(min(w[k] & 0xffffu, w[k] >> 16)));
    } else {
      m = fminf(m, __uint_as_float(w[k]));
    }
  }
  return m;
}

template <int DT>
__global__ __launch_bounds__(256) void k_min_reduce_flat(const piv_image* __restrict__ images,
                                                         uint32_t* keys) {
  using E = typename PixelType<DT>::type;
  const piv_image im = images[blockIdx.y];
  const int64_t nbytes = int64_t(im.height) * im.width * int64_t(sizeof(E));
  const int64_t n16 = nbytes / 16;
  const uint4* q = static_cast<const uint4*>(im.data);
  float m0 = INFINITY, m1 = INFINITY, m2 = INFINITY, m3 = INFINITY;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const uint4 a = q[i], b = q[i + stride], c = q[i + 2 * stride], d = q[i + 3 * stride];
    m0 = fminf(m0, min16<DT>(a));
    m1 = fminf(m1, min16<DT>(b));
    m2 = fminf(m2, min16<DT>(c));
    m3 = fminf(m3, min16<DT>(d));
  }
  for (; i < n16; i += stride) m0 = fminf(m0, min16<DT>(q[i]));
  float m = fminf(fminf(m0, m1), fminf(m2, m3));
  if (blockIdx.x == 0) {
    const int64_t first = n16 * 16 / int64_t(sizeof(E));
    const int64_t total = int64_t(im.height) * im.width;
    for (int64_t k = first + threadIdx.x; k < total; k += blockDim.x)
      m = fminf(m, gfetch<DT>(im.data, k));
  }
  for (int off = 32; off > 0; off >>= 1) m = fminf(m, __shfl_xor(m, off, 64));
  __shared__ float part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float b = fminf(fminf(part[0], part[1]), fminf(part[2], part[3]));
    atomicMin(&keys[blockIdx.y], float_key(b));
  }
}

__global__ void k_min_final(uint32_t* keys, float* out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = key_float(keys[i]);
}

// -------------------------------------------------------------------------
// WorkFIX (fix.py:66-133) and mean_neighbors (mean_neighbors.py:5-155)
// -------------------------------------------------------------------------
// pair blockIdx.y of a batch: its n vectors at offset blockIdx.y * n of every array
__global__ void k_fix_mark(float* dx, float* dy, const float* cmax, const int32_t* err, int n,
                           float correl_min, float dmax2, int use_dmax, int32_t* flags) {
  // numpy's deltaxs**2 + deltays**2 (fix.py:99) rounds the squares and the
  // sum separately: no fused multiply-add here
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  i += blockIdx.y * n;
  int f = 0;
  float x = dx[i], y = dy[i];
  if (err[i] != PIV_WIN_OK) f |= 1;
  if (cmax[i] < correl_min) f |= 2;
  // delta2 test on the values after the previous NaN assignments
  float xx = f ? __builtin_nanf("") : x, yy = f ? __builtin_nanf("") : y;
  if (use_dmax && xx * xx + yy * yy > dmax2) f |= 4;
  if (f) {
    dx[i] = __builtin_nanf("");
    dy[i] = __builtin_nanf("");
  }
  flags[i] = f;
}

// sum over the valid neighbours with the reference's order of additions
// (mean_neighbors.py:15-115); NaN vectors count as 0 with weight 0
__device__ __forceinline__ void neigh_sum(const float* f, const float* g, int ny, int nx, int i0,
                                          int i1, float& sa, float& sb) {
  auto A = [&](int p, int q) {
    const float x = f[p * nx + q], y = g[p * nx + q];
    return (x != x || y != y) ? 0.f : x;
  };
  auto B = [&](int p, int q) {
    const float x = f[p * nx + q], y = g[p * nx + q];
    return (x != x || y != y) ? 0.f : 1.f;
  };
  const int n0 = ny, n1 = nx;
  if (i0 == 0 && i1 == 0) {
    sa = A(1, 0) + A(1, 1) + A(0, 1);
    sb = B(1, 0) + B(1, 1) + B(0, 1);
  } else if (i0 == 0 && i1 == n1 - 1) {
    sa = A(1, i1) + A(1, i1 - 1) + A(0, i1 - 1);
    sb = B(1, i1) + B(1, i1 - 1) + B(0, i1 - 1);
  } else if (i0 == n0 - 1 && i1 == 0) {
    sa = A(i0 - 1, 0) + A(i0 - 1, 1) + A(i0, 1);
    sb = B(i0 - 1, 0) + B(i0 - 1, 1) + B(i0, 1);
  } else if (i0 == n0 - 1 && i1 == n1 - 1) {
    sa = A(i0 - 1, i1) + A(i0 - 1, i1 - 1) + A(i0, i1 - 1);
    sb = B(i0 - 1, i1) + B(i0 - 1, i1 - 1) + B(i0, i1 - 1);
  } else if (i0 == 0 || i0 == n0 - 1) {
    const int o = i0 == 0 ? 1 : -1;
    sa = A(i0, i1 - 1) + A(i0, i1 + 1);
    sb = B(i0, i1 - 1) + B(i0, i1 + 1);
    for (int idx = -1; idx < 2; ++idx) {
      sa += A(i0 + o, i1 - idx);
      sb += B(i0 + o, i1 - idx);
    }
  } else if (i1 == 0 || i1 == n1 - 1) {
    const int o = i1 == 0 ? 1 : -1;
    sa = A(i0 - 1, i1) + A(i0 + 1, i1);
    sb = B(i0 - 1, i1) + B(i0 + 1, i1);
    for (int idx = -1; idx < 2; ++idx) {
      sa += A(i0 + idx, i1 + o);
      sb += B(i0 + idx, i1 + o);
    }
  } else {
    sa = A(i0, i1 - 1) + A(i0, i1 + 1);
    sb = B(i0, i1 - 1) + B(i0, i1 + 1);
    for (int idx = -1; idx < 2; ++idx) {
      sa += A(i0 - 1, i1 - idx) + A(i0 + 1, i1 - idx);
      sb += B(i0 - 1, i1 - idx) + B(i0 + 1, i1 - idx);
    }
  }
}

// neighbour means, squared differences and the threshold flag; reads the
// fields as left by k_fix_mark and writes nothing they are read from
__global__ void k_fix_neigh(const float* dx, const float* dy, int ny, int nx, float thr2,
                            int use_thr, int32_t* flags, float* diff2, float* mean_x,
                            float* mean_y) {
  // diff_x**2 + diff_y**2 (fix.py:121-123): separately rounded, no FMA
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ny * nx) return;
  const int i0 = i / nx, i1 = i % nx;
  {
    const int64_t o = int64_t(blockIdx.y) * ny * nx;
    dx += o;
    dy += o;
    flags += o;
    if (diff2) diff2 += o;
    if (mean_x) mean_x += o;
    if (mean_y) mean_y += o;
  }
  const float x = dx[i], y = dy[i];
  float mx = __builtin_nanf(""), my = __builtin_nanf("");
  if (!(x != x || y != y) && ny > 1 && nx > 1) {
    float sx, sy, cx, cy;
    neigh_sum(dx, dy, ny, nx, i0, i1, sx, cx);
    neigh_sum(dy, dx, ny, nx, i0, i1, sy, cy);
    if (cx != 0.f) {
      mx = sx / cx;
      my = sy / cy;
    }
  }
  const float ddx = mx - x, ddy = my - y;
  const float d2 = ddx * ddx + ddy * ddy;
  if (diff2) diff2[i] = d2;
  if (mean_x) mean_x[i] = mx;
  if (mean_y) mean_y[i] = my;
  if (use_thr && d2 > thr2) flags[i] |= 8;
}

__global__ void k_fix_apply(float* dx, float* dy, const int32_t* flags, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  i += blockIdx.y * n;
  if (flags[i] & 8) {
    dx[i] = __builtin_nanf("");
    dy[i] = __builtin_nanf("");
  }
}

// -------------------------------------------------------------------------
// window records of one pass (singlepass.py:210-260 first pass,
// :752-753 predictor rounding, :787-870 later passes, :391-407 crop bounds)
// -------------------------------------------------------------------------
struct CropGeom {
  int s0y, s0x, e0y, e0x;  // _start_for_crop0 / _stop_for_crop0
  int s1y, s1x, e1y, e1x;  // _start_for_crop1 / _stop_for_crop1
};

__global__ void k_make_windows(int ny, int nx, const int32_t* __restrict__ iyvecs,
                               const int32_t* __restrict__ ixvecs,
                               const double* __restrict__ dxa, const double* __restrict__ dya,
                               int height, int width, int npad, CropGeom g, int image0,
                               int image1, piv_window* __restrict__ out, double* xs, double* ys,
                               int32_t* dxin_out, int32_t* dyin_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ny * nx) return;
  const int gx = ixvecs[i % nx], gy = iyvecs[i / nx];
  // pair blockIdx.y of a batch: images image0 + 2 k, image1 + 2 k, its
  // vectors at offset k ny nx of the predictor and output arrays
  image0 += 2 * blockIdx.y;
  image1 += 2 * blockIdx.y;
  i += blockIdx.y * ny * nx;
  int dxi = 0, dyi = 0;
  int ixs0 = gx, iys0 = gy, ixs1 = gx, iys1 = gy;
  if (dxa != nullptr) {
    // np.round (half to even) then astype(int32)
    dxi = int(rint(dxa[i]));
    dyi = int(rint(dya[i]));
    ixs0 = gx - (dxi >> 1);  // floor division by 2
    iys0 = gy - (dyi >> 1);
    ixs1 = ixs0 + dxi;
    iys1 = iys0 + dyi;
    const bool out1 = ixs1 > width || ixs1 < 0 || iys1 > height || iys1 < 0;
    const bool out0 = ixs0 > width || ixs0 < 0 || iys0 > height || iys0 < 0;
    if (out1) {
      ixs0 = gx - dxi;
      iys0 = gy - dyi;
      ixs1 = gx;
      iys1 = gy;
    } else if (out0) {
      ixs0 = gx;
      iys0 = gy;
      ixs1 = gx + dxi;
      iys1 = gy + dyi;
    }
  }
  if (xs) xs[i] = dxa != nullptr ? (double(ixs0) + double(ixs1)) / 2.0 : double(gx);
  if (ys) ys[i] = dxa != nullptr ? (double(iys0) + double(iys1)) / 2.0 : double(gy);
  if (dxin_out) dxin_out[i] = dxi;
  if (dyin_out) dyin_out[i] = dyi;
  // crop rows [c - start, c + stop) of the padded image must lie in
  // [0, H + 2 npad], else the slice has the wrong shape ("Bad im_crop shape.")
  const int hp = height + 2 * npad, wp = width + 2 * npad;
  const int py0 = iys0 + npad, px0 = ixs0 + npad, py1 = iys1 + npad, px1 = ixs1 + npad;
  const bool ok = py0 - g.s0y >= 0 && py0 + g.e0y <= hp && px0 - g.s0x >= 0 &&
                  px0 + g.e0x <= wp && py1 - g.s1y >= 0 && py1 + g.e1y <= hp &&
                  px1 - g.s1x >= 0 && px1 + g.e1x <= wp;
  piv_window w;
  w.image0 = image0;
  w.image1 = image1;
  w.y0 = iys0 - g.s0y;
  w.x0 = ixs0 - g.s0x;
  w.y1 = iys1 - g.s1y;
  w.x1 = ixs1 - g.s1x;
  w.dx_in = dxi;
  w.dy_in = dyi;
  w.flags = ok ? 0 : PIV_WINFLAG_INVALID;
  w.reserved = 0;
  out[i] = w;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
struct piv_plan {
  int n0, n1;
  const Launchers* l;
  uint8_t* d_mask;
  PlanArgs pa;
};

extern "C" {

const char* piv_version(void) { return "fluidimage_amd piv_hip 0.1 (gfx950)"; }

const char* piv_last_error(void) { return fia::last_error().c_str(); }

int piv_shape_supported(int32_t n0, int32_t n1) {
  return find_shape(n0, n1) != nullptr || any_shape_ok(n0, n1);
}

int piv_tune_candidates(int32_t param, int32_t n0, int32_t n1, int32_t* values, int32_t cap) {
  if (values == nullptr && cap > 0) return -PIV_EINVAL;
  int count = 0;
  auto put = [&](int v) {
    if (count < cap) values[count] = v;
    ++count;
  };
  if (param == PIV_TUNE_CORRELATE_WPB) {
    const Launchers* l = find_shape(n0, n1);
    if (l == nullptr) return -fail(PIV_EUNSUPPORTED, "window shape not built");
    for (const CorrVariant& v : l->var)
      if (v.wpb > 0) put(v.wpb);
    return count;
  }
  if (param == PIV_TUNE_TPS_CHOL_THREADS) {
    put(0);
    put(256);
    put(512);
    return count;
  }
  if (param == PIV_TUNE_TPS_SHARED_SOLVE) {
    put(0);
    put(1);
    return count;
  }
  return -fail(PIV_EINVAL, "unknown tuning parameter");
}

int piv_tune_get(int32_t param, int32_t n0, int32_t n1) {
  if (param == PIV_TUNE_CORRELATE_WPB) {
    const Launchers* l = find_shape(n0, n1);
    if (l == nullptr) return -fail(PIV_EUNSUPPORTED, "window shape not built");
    return variant_of(l).wpb;
  }
  if (param == PIV_TUNE_TPS_CHOL_THREADS) return fia::tps_chol_threads();
  if (param == PIV_TUNE_TPS_SHARED_SOLVE) return fia::tps_shared_solve();
  return -fail(PIV_EINVAL, "unknown tuning parameter");
}

int piv_tune_set(int32_t param, int32_t n0, int32_t n1, int32_t value) {
  if (param == PIV_TUNE_CORRELATE_WPB) {
    const int i = shape_index(n0, n1);
    if (i < 0) return fail(PIV_EUNSUPPORTED, "window shape not built");
    for (int k = 0; k < kMaxVariants; ++k)
      if (kShapes[i].var[k].wpb > 0 && kShapes[i].var[k].wpb == value) {
        g_variant[i].store(k, std::memory_order_relaxed);
        return PIV_OK;
      }
    return fail(PIV_EINVAL, "windows per workgroup not among the candidates of this shape");
  }
  if (param == PIV_TUNE_TPS_CHOL_THREADS) {
    if (value != 0 && value != 256 && value != 512)
      return fail(PIV_EINVAL, "TPS Cholesky threads must be 0, 256 or 512");
    fia::set_tps_chol_threads(value);
    return PIV_OK;
  }
  if (param == PIV_TUNE_TPS_SHARED_SOLVE) {
    if (value != 0 && value != 1) return fail(PIV_EINVAL, "TPS shared solve must be 0 or 1");
    fia::set_tps_shared_solve(value);
    return PIV_OK;
  }
  return fail(PIV_EINVAL, "unknown tuning parameter");
}

int piv_plan_create(piv_plan** plan, int32_t n0, int32_t n1, double displacement_max,
                    const uint8_t* mask, const int32_t* box, int32_t subpix_method,
                    int32_t nsubpix, int32_t particle_radius, int32_t nb_peaks_to_search) {
  if (plan == nullptr) return fail(PIV_EINVAL, "plan is NULL");
  *plan = nullptr;
  // a register-FFT kernel of the shape, else the DFT kernel of any shape
  const Launchers* l = find_shape(n0, n1);
  if (l == nullptr && !any_shape_ok(n0, n1))
    return fail(PIV_EUNSUPPORTED,
                "window shape " + std::to_string(n0) + "x" + std::to_string(n1) +
                    " is not supported by libpiv_hip (rows and columns 2 to " +
                    std::to_string(kAnyMaxLen) + ", at most " + std::to_string(kAnyMaxElems) +
                    " elements)");
  if (subpix_method < 0 || subpix_method > PIV_SUBPIX_2D_GAUSSIAN)
    return fail(PIV_EINVAL, "bad subpix method");
  if (nsubpix < 1 || 2 * nsubpix + 1 > (n0 < n1 ? n0 : n1))
    return fail(PIV_EINVAL, "bad nsubpix");
  // SubPix.__init__ (subpix.py:76-81): the 3x3 fit takes no nsubpix
  if (subpix_method == PIV_SUBPIX_2D_GAUSSIAN2 && nsubpix != 1)
    return fail(PIV_EINVAL,
                "Subpixel method '2d_gaussian2' doesn't require nsubpix. "
                "In this case, nsubpix has to be equal to None.");
  if (nb_peaks_to_search < 1 || nb_peaks_to_search > kMaxPeaks)
    return fail(PIV_EUNSUPPORTED, "nb_peaks_to_search must be in [1, 7]");
  if (particle_radius < 0) return fail(PIV_EINVAL, "bad particle_radius");
  piv_plan* p = new piv_plan();
  p->n0 = n0;
  p->n1 = n1;
  p->l = l;
  p->d_mask = nullptr;
  p->pa.mask = nullptr;
  p->pa.has_dmax = displacement_max >= 0.0;
  if (p->pa.has_dmax) {
    if (mask == nullptr || box == nullptr) {
      delete p;
      return fail(PIV_EINVAL, "mask and box are required with displacement_max");
    }
    if (hipMalloc(&p->d_mask, size_t(n0) * n1) != hipSuccess ||
        hipMemcpy(p->d_mask, mask, size_t(n0) * n1, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(p->d_mask);
      delete p;
      return fail(PIV_EHIP, "cannot upload the search mask");
    }
    p->pa.mask = p->d_mask;
    p->pa.b0s = box[0];
    p->pa.b0e = box[1];
    p->pa.b1s = box[2];
    p->pa.b1e = box[3];
    if (box[0] < 0 || box[1] > n0 || box[2] < 0 || box[3] > n1) {
      (void)hipFree(p->d_mask);
      delete p;
      return fail(PIV_EINVAL, "search box outside the window");
    }
  } else {
    p->pa.b0s = 0;
    p->pa.b0e = n0;
    p->pa.b1s = 0;
    p->pa.b1e = n1;
  }
  p->pa.subpix_method = subpix_method;
  p->pa.nsubpix = nsubpix;
  p->pa.particle_radius = particle_radius;
  p->pa.nb_peaks = nb_peaks_to_search;
  *plan = p;
  return PIV_OK;
}

int piv_plan_destroy(piv_plan* plan) {
  if (plan == nullptr) return PIV_OK;
  if (plan->d_mask) (void)hipFree(plan->d_mask);
  delete plan;
  return PIV_OK;
}

int piv_images_min(const piv_image* images, int32_t nimages, float* out_min,
                   int32_t flat_dtype, void* stream) {
  if (nimages <= 0) return PIV_OK;
  if (flat_dtype > PIV_F32) return fail(PIV_EINVAL, "bad dtype");
  if (images == nullptr || out_min == nullptr) return fail(PIV_EINVAL, "NULL argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // keys live in the output buffer (same size, reinterpreted)
  uint32_t* keys = reinterpret_cast<uint32_t*>(out_min);
  hipLaunchKernelGGL(k_min_init, dim3((nimages + 255) / 256), dim3(256), 0, s, keys, nimages);
  // the flat kernel when the caller says so: every image contiguous
  // (pitch == width), of one dtype, with 16-byte aligned data
  if (flat_dtype >= 0) {
    const dim3 g(64, nimages), b(256);
    if (flat_dtype == PIV_U8)
      hipLaunchKernelGGL(k_min_reduce_flat<PIV_U8>, g, b, 0, s, images, keys);
    else if (flat_dtype == PIV_U16)
      hipLaunchKernelGGL(k_min_reduce_flat<PIV_U16>, g, b, 0, s, images, keys);
    else
      hipLaunchKernelGGL(k_min_reduce_flat<PIV_F32>, g, b, 0, s, images, keys);
  } else {
    hipLaunchKernelGGL(k_min_reduce, dim3(256, nimages), dim3(256), 0, s, images, keys);
  }
  hipLaunchKernelGGL(k_min_final, dim3((nimages + 255) / 256), dim3(256), 0, s, keys, out_min,
                     nimages);
  return check_launch("piv_images_min");
}

int piv_correlate_batch(piv_plan* plan, int32_t dtype, const piv_image* images,
                           const piv_window* windows, int32_t nwin, int32_t options,
                           float* maps, const piv_results* res, void* stream) {
  if (plan == nullptr || res == nullptr) return fail(PIV_EINVAL, "NULL plan or results");
  if (nwin <= 0) return PIV_OK;
  if (dtype < 0 || dtype > PIV_F32) return fail(PIV_EINVAL, "bad dtype");
  if (images == nullptr || windows == nullptr || res->dx == nullptr || res->dy == nullptr ||
      res->correl_max == nullptr || res->err == nullptr)
    return fail(PIV_EINVAL, "NULL image, window or result array");
  if (plan->pa.nb_peaks > 1 && (res->sec == nullptr || res->nsec == nullptr))
    return fail(PIV_EINVAL, "secondary peak arrays required when nb_peaks_to_search > 1");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (plan->l == nullptr) {
    const size_t lds = any_lds_bytes(plan->n0, plan->n1);
    const dim3 grid(nwin), block(kAnyThreads);
    if (dtype == PIV_U8)
      hipLaunchKernelGGL(k_correlate_any<PIV_U8>, grid, block, lds, s, plan->n0, plan->n1,
                         images, windows, int(nwin), int(options), maps, *res, plan->pa);
    else if (dtype == PIV_U16)
      hipLaunchKernelGGL(k_correlate_any<PIV_U16>, grid, block, lds, s, plan->n0, plan->n1,
                         images, windows, int(nwin), int(options), maps, *res, plan->pa);
    else
      hipLaunchKernelGGL(k_correlate_any<PIV_F32>, grid, block, lds, s, plan->n0, plan->n1,
                         images, windows, int(nwin), int(options), maps, *res, plan->pa);
    return check_launch("piv_correlate_batch");
  }
  const CorrVariant& v = variant_of(plan->l);
  dim3 grid((nwin + v.wpb - 1) / v.wpb), block(v.threads);
  v.correlate[dtype](grid, block, s, images, windows, nwin, options, maps, *res, plan->pa);
  return check_launch("piv_correlate_batch");
}

int piv_peaks_from_maps(piv_plan* plan, float* maps, const double* norms, int32_t nwin,
                        int32_t apply_subpix, const piv_results* res, void* stream) {
  if (plan == nullptr || res == nullptr) return fail(PIV_EINVAL, "NULL plan or results");
  if (nwin <= 0) return PIV_OK;
  if (maps == nullptr || norms == nullptr || res->dx == nullptr || res->dy == nullptr ||
      res->correl_max == nullptr || res->err == nullptr)
    return fail(PIV_EINVAL, "NULL map, norm or result array");
  if (plan->pa.nb_peaks > 1 && (res->sec == nullptr || res->nsec == nullptr))
    return fail(PIV_EINVAL, "secondary peak arrays required when nb_peaks_to_search > 1");
  const Launchers* l = plan->l;
  if (l == nullptr) {
    const size_t lds = size_t(8) * plan->n0 * plan->n1;
    hipLaunchKernelGGL(k_peaks_generic, dim3(nwin), dim3(64), lds,
                       static_cast<hipStream_t>(stream), maps, norms, plan->n0, plan->n1,
                       int(nwin), int(apply_subpix), *res, plan->pa);
    return check_launch("piv_peaks_from_maps");
  }
  dim3 grid((nwin + l->pwpb - 1) / l->pwpb), block(l->pthreads);
  l->peaks(grid, block, static_cast<hipStream_t>(stream), maps, norms, nwin, apply_subpix, *res,
           plan->pa);
  return check_launch("piv_peaks_from_maps");
}

int piv_subpix_at(piv_plan* plan, const float* maps, const int32_t* ix, const int32_t* iy,
                  int32_t nwin, float* out_x, float* out_y, int32_t* err, void* stream) {
  if (plan == nullptr) return fail(PIV_EINVAL, "NULL plan");
  if (nwin <= 0) return PIV_OK;
  if (!maps || !ix || !iy || !out_x || !out_y || !err) return fail(PIV_EINVAL, "NULL argument");
  if (plan->l == nullptr) {
    hipLaunchKernelGGL(k_subpix_generic, dim3((nwin + 63) / 64), dim3(64), 0,
                       static_cast<hipStream_t>(stream), maps, plan->n0, plan->n1, ix, iy,
                       int(nwin), out_x, out_y, err, plan->pa);
    return check_launch("piv_subpix_at");
  }
  plan->l->subpix(dim3((nwin + 63) / 64), dim3(64), static_cast<hipStream_t>(stream), maps, ix,
                  iy, nwin, out_x, out_y, err, plan->pa);
  return check_launch("piv_subpix_at");
}

int piv_fix(int32_t npairs, float* dx, float* dy, const float* correl_max, const int32_t* err,
            int32_t ny, int32_t nx, float correl_min, float displacement_max, float threshold,
            int32_t* flags, float* diff2, float* mean_x, float* mean_y, void* stream) {
  const int n = ny * nx;
  if (n <= 0 || npairs <= 0) return PIV_OK;
  if (npairs > 65535) return fail(PIV_EINVAL, "too many pairs for one launch");
  if (!dx || !dy || !correl_max || !err || !flags) return fail(PIV_EINVAL, "NULL argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 nb((n + 255) / 256, npairs);
  const float dmax2 = float(double(displacement_max) * displacement_max);
  hipLaunchKernelGGL(k_fix_mark, dim3(nb), dim3(256), 0, s, dx, dy, correl_max, err, n,
                     correl_min, dmax2, displacement_max > 0.f, flags);
  const float thr2 = float(double(threshold) * threshold);
  hipLaunchKernelGGL(k_fix_neigh, dim3(nb), dim3(256), 0, s, dx, dy, ny, nx, thr2,
                     threshold >= 0.f, flags, diff2, mean_x, mean_y);
  if (threshold >= 0.f)
    hipLaunchKernelGGL(k_fix_apply, dim3(nb), dim3(256), 0, s, dx, dy, flags, n);
  return check_launch("piv_fix");
}

int piv_make_windows(int32_t npairs, int32_t ny, int32_t nx, const int32_t* iyvecs,
                     const int32_t* ixvecs, const double* dx_approx, const double* dy_approx,
                     int32_t height, int32_t width, int32_t npad, const int32_t* crop,
                     int32_t image0, int32_t image1, piv_window* windows, double* xs, double* ys,
                     int32_t* dx_in, int32_t* dy_in, void* stream) {
  const int n = ny * nx;
  if (n <= 0 || npairs <= 0) return PIV_OK;
  if (npairs > 65535) return fail(PIV_EINVAL, "too many pairs for one launch");
  if (!iyvecs || !ixvecs || !crop || !windows) return fail(PIV_EINVAL, "NULL argument");
  if ((dx_approx == nullptr) != (dy_approx == nullptr))
    return fail(PIV_EINVAL, "dx_approx and dy_approx must both be given or both NULL");
  CropGeom g{crop[0], crop[1], crop[2], crop[3], crop[4], crop[5], crop[6], crop[7]};
  hipLaunchKernelGGL(k_make_windows, dim3((n + 255) / 256, npairs), dim3(256), 0,
                     static_cast<hipStream_t>(stream), ny, nx, iyvecs, ixvecs, dx_approx,
                     dy_approx, height, width, npad, g, image0, image1, windows, xs, ys, dx_in,
                     dy_in);
  return check_launch("piv_make_windows");
}

}  // extern "C"

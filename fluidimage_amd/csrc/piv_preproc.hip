// Sliding-window preprocessing of PIV images (gfx950): the reference's
// sliding_minima, sliding_median and sliding_percentile
// (src/fluidimage/works/preproc/_toolbox_py.py:66-165):
//
//   out = img - weight * F(img),  out[out < 0] = 0,
//
// F the scipy.ndimage minimum / rank filter over a wy x wx box (offsets
// -(w // 2) .. w - 1 - w // 2, scipy's origin 0) with scipy's boundary
// modes. Integer images give float64 outputs (numpy promotes the float
// weight), float32 images float32 outputs (a Python float is weak against
// a float32 array), each with numpy's roundings (no contraction).
//
// Both kernels stage a tile of the image with its halo in LDS (boundary
// modes applied while loading). The minimum is separable: row minima of
// the tile rows, then column minima. The rank filter gives one output to
// each lane: an exact radix selection of the rank-th smallest key of its
// window, one bit at a time from the highest bit in which the window's
// smallest and largest keys differ (keys: the values themselves for
// integer images, order-preserving bit patterns for float32).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/piv_hip.h"
#include "last_error.hpp"

namespace {

int pfail(int code, const std::string& msg) {
  fia::last_error() = msg;
  return code;
}

int pcheck(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return pfail(PIV_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return PIV_OK;
}

constexpr int kTX = 64;       // output tile columns (one wavefront per row)
constexpr int kMinTY = 16;    // output tile rows of the minimum filter
constexpr int kRankTY = 4;    // output tile rows of the rank filter
constexpr int kThreads = 256;
constexpr int kMaxWin = 128;  // window rows / columns

// index of the sample that stands for i in [0, n) under a scipy.ndimage
// boundary mode; -1 for 'constant' outside
__device__ __forceinline__ int boundary_index(int i, int n, int mode) {
  if (i >= 0 && i < n) return i;
  switch (mode) {
    case PIV_BOUNDARY_REFLECT: {  // d c b a | a b c d | d c b a
      const int p = 2 * n;
      i %= p;
      if (i < 0) i += p;
      return i < n ? i : p - 1 - i;
    }
    case PIV_BOUNDARY_MIRROR: {  // d c b | a b c d | c b a
      if (n == 1) return 0;
      const int p = 2 * n - 2;
      i %= p;
      if (i < 0) i += p;
      return i < n ? i : p - i;
    }
    case PIV_BOUNDARY_NEAREST:
      return i < 0 ? 0 : n - 1;
    case PIV_BOUNDARY_WRAP: {
      i %= n;
      return i < 0 ? i + n : i;
    }
    default:
      return -1;
  }
}

template <int DT>
__device__ __forceinline__ float load_px(const piv_image& im, int y, int x) {
  const int64_t o = int64_t(y) * im.pitch + x;
  if constexpr (DT == PIV_U8) return float(static_cast<const uint8_t*>(im.data)[o]);
  else if constexpr (DT == PIV_U16) return float(static_cast<const uint16_t*>(im.data)[o]);
  else return static_cast<const float*>(im.data)[o];
}

// the tile value at image (y, x) under the boundary mode (cval outside for
// 'constant', cast to the image type as scipy casts it)
template <int DT>
__device__ __forceinline__ float tile_value(const piv_image& im, int y, int x, int mode,
                                            float cval) {
  const int yy = boundary_index(y, im.height, mode), xx = boundary_index(x, im.width, mode);
  if (yy < 0 || xx < 0) return cval;
  return load_px<DT>(im, yy, xx);
}

// out = img - weight * f, negatives to 0, with numpy's types and roundings
template <int DT>
__device__ __forceinline__ void store_out(void* out, int64_t o, float img, float f,
                                          double weight) {
  // numpy rounds the product, then the difference: no fused multiply-add
#pragma clang fp contract(off)
  if constexpr (DT == PIV_F32) {
    const float t = float(weight) * f;
    float v = img - t;
    if (v < 0.f) v = 0.f;
    static_cast<float*>(out)[o] = v;
  } else {
    const double t = weight * double(f);
    double v = double(img) - t;
    if (v < 0.0) v = 0.0;
    static_cast<double*>(out)[o] = v;
  }
}

template <int DT>
__global__ __launch_bounds__(kThreads) void k_sliding_min(const piv_image* __restrict__ images,
                                                          int wy, int wx, int mode, float cval,
                                                          double weight, void* __restrict__ out) {
  extern __shared__ float pre_smem[];
  const piv_image im = images[blockIdx.z];
  const int H = im.height, W = im.width;
  const int ty0 = blockIdx.y * kMinTY, tx0 = blockIdx.x * kTX;
  const int oy = wy / 2, ox = wx / 2;  // window offsets -o .. w - 1 - o
  const int IR = kMinTY + wy - 1, IC = kTX + wx - 1;
  float* in = pre_smem;          // IR x IC
  float* rmin = in + IR * IC;    // IR x kTX
  for (int k = threadIdx.x; k < IR * IC; k += kThreads) {
    const int r = k / IC, c = k % IC;
    in[k] = tile_value<DT>(im, ty0 - oy + r, tx0 - ox + c, mode, cval);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < IR * kTX; k += kThreads) {
    const int r = k / kTX, c = k % kTX;
    const float* p = in + r * IC + c;
    float m = p[0];
    for (int d = 1; d < wx; ++d) m = fminf(m, p[d]);
    rmin[k] = m;
  }
  __syncthreads();
  const int64_t base = int64_t(blockIdx.z) * H * W;
  for (int k = threadIdx.x; k < kMinTY * kTX; k += kThreads) {
    const int r = k / kTX, c = k % kTX;
    const int y = ty0 + r, x = tx0 + c;
    if (y >= H || x >= W) continue;
    const float* p = rmin + r * kTX + c;
    float m = p[0];
    for (int d = 1; d < wy; ++d) m = fminf(m, p[d * kTX]);
    store_out<DT>(out, base + int64_t(y) * W + x, in[(r + oy) * IC + c + ox], m, weight);
  }
}

// order-preserving keys of the values
template <int DT>
__device__ __forceinline__ uint32_t value_key(float v) {
  if constexpr (DT == PIV_F32) {
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  } else {
    return uint32_t(v);
  }
}

template <int DT>
__device__ __forceinline__ float key_value(uint32_t k) {
  if constexpr (DT == PIV_F32) {
    const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(b);
  } else {
    return float(k);
  }
}

template <int DT>
__global__ __launch_bounds__(kThreads) void k_sliding_rank(const piv_image* __restrict__ images,
                                                           int wy, int wx, int rank, int mode,
                                                           float cval, double weight,
                                                           void* __restrict__ out) {
  extern __shared__ uint32_t pre_keys[];
  const piv_image im = images[blockIdx.z];
  const int H = im.height, W = im.width;
  const int ty0 = blockIdx.y * kRankTY, tx0 = blockIdx.x * kTX;
  const int oy = wy / 2, ox = wx / 2;
  const int IR = kRankTY + wy - 1, IC = kTX + wx - 1;
  for (int k = threadIdx.x; k < IR * IC; k += kThreads) {
    const int r = k / IC, c = k % IC;
    pre_keys[k] = value_key<DT>(tile_value<DT>(im, ty0 - oy + r, tx0 - ox + c, mode, cval));
  }
  __syncthreads();
  const int r = threadIdx.x / kTX, c = threadIdx.x % kTX;
  const int y = ty0 + r, x = tx0 + c;
  if (y >= H || x >= W) return;
  const uint32_t* win = pre_keys + r * IC + c;
  // smallest and largest keys: the bits above their highest differing bit
  // are those of the answer
  uint32_t lo = 0xffffffffu, hi = 0u;
  for (int a = 0; a < wy; ++a)
    for (int b = 0; b < wx; ++b) {
      const uint32_t k = win[a * IC + b];
      lo = min(lo, k);
      hi = max(hi, k);
    }
  uint32_t key = lo;
  const uint32_t diff = lo ^ hi;
  if (diff != 0u) {
    const int top = 31 - __clz(int(diff));
    // prefix: the common bits above `top`, zeros below
    uint32_t prefix = top == 31 ? 0u : (lo >> (top + 1)) << (top + 1);
    int r_left = rank;
    for (int bit = top; bit >= 0; --bit) {
      // keys whose bits above `bit` equal the prefix's and whose bit is 0
      const uint32_t want = prefix >> bit;
      int count = 0;
      for (int a = 0; a < wy; ++a) {
        const uint32_t* row = win + a * IC;
        for (int b = 0; b < wx; ++b) count += (row[b] >> bit) == want ? 1 : 0;
      }
      if (r_left >= count) {
        r_left -= count;
        prefix |= 1u << bit;
      }
    }
    key = prefix;
  }
  const float img = key_value<DT>(win[oy * IC + ox]);
  store_out<DT>(out, int64_t(blockIdx.z) * H * W + int64_t(y) * W + x, img, key_value<DT>(key),
                weight);
}

}  // namespace

extern "C" {

int piv_sliding_filter(const piv_image* images, int32_t nimages, int32_t height, int32_t width,
                       int32_t dtype, int32_t wy, int32_t wx, int32_t rank, int32_t mode,
                       double cval, double weight, void* out, void* stream) {
  if (nimages <= 0) return PIV_OK;
  if (images == nullptr || out == nullptr) return pfail(PIV_EINVAL, "NULL argument");
  if (dtype < PIV_U8 || dtype > PIV_F32) return pfail(PIV_EINVAL, "bad dtype");
  if (height <= 0 || width <= 0) return pfail(PIV_EINVAL, "bad image shape");
  if (nimages > 65535) return pfail(PIV_EINVAL, "too many images for one launch");
  if (wy < 1 || wx < 1 || wy > kMaxWin || wx > kMaxWin)
    return pfail(PIV_EUNSUPPORTED, "window sizes from 1 to 128 only");
  if (rank < 0 || rank >= wy * wx) return pfail(PIV_EINVAL, "rank not within filter footprint size");
  if (mode < PIV_BOUNDARY_REFLECT || mode > PIV_BOUNDARY_WRAP)
    return pfail(PIV_EINVAL, "unknown boundary mode");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // scipy casts cval to the image type
  float cv = float(cval);
  if (dtype == PIV_U8 || dtype == PIV_U16) {
    const double m = dtype == PIV_U8 ? 255.0 : 65535.0;
    double c = cval < 0.0 ? 0.0 : (cval > m ? m : cval);
    cv = float(int64_t(c));
  }
  if (rank == 0) {
    const dim3 g((width + kTX - 1) / kTX, (height + kMinTY - 1) / kMinTY, nimages);
    const size_t lds = sizeof(float) * (size_t(kMinTY + wy - 1) * (kTX + wx - 1) +
                                        size_t(kMinTY + wy - 1) * kTX);
    if (dtype == PIV_U8)
      hipLaunchKernelGGL(k_sliding_min<PIV_U8>, g, dim3(kThreads), lds, s, images, wy, wx, mode,
                         cv, weight, out);
    else if (dtype == PIV_U16)
      hipLaunchKernelGGL(k_sliding_min<PIV_U16>, g, dim3(kThreads), lds, s, images, wy, wx, mode,
                         cv, weight, out);
    else
      hipLaunchKernelGGL(k_sliding_min<PIV_F32>, g, dim3(kThreads), lds, s, images, wy, wx, mode,
                         cv, weight, out);
  } else {
    const dim3 g((width + kTX - 1) / kTX, (height + kRankTY - 1) / kRankTY, nimages);
    const size_t lds = sizeof(uint32_t) * size_t(kRankTY + wy - 1) * (kTX + wx - 1);
    if (dtype == PIV_U8)
      hipLaunchKernelGGL(k_sliding_rank<PIV_U8>, g, dim3(kThreads), lds, s, images, wy, wx, rank,
                         mode, cv, weight, out);
    else if (dtype == PIV_U16)
      hipLaunchKernelGGL(k_sliding_rank<PIV_U16>, g, dim3(kThreads), lds, s, images, wy, wx, rank,
                         mode, cv, weight, out);
    else
      hipLaunchKernelGGL(k_sliding_rank<PIV_F32>, g, dim3(kThreads), lds, s, images, wy, wx, rank,
                         mode, cv, weight, out);
  }
  return pcheck("piv_sliding_filter");
}

}  // extern "C"

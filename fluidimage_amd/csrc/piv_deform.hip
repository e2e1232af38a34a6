// Image deformation between passes (gfx950): bicubic resampling of the
// interrogation windows by the predictor field. This has no counterpart in
// the reference, whose later passes shift the windows by the rounded
// predictor (src/fluidimage/works/piv/singlepass.py:752-753,822-825).
//
// Method: central-difference window deformation. Window w of pair k is
// centred on grid node (yc, xc); its pixel (r, c) is at y = yc - n0/2 + r,
// x = xc - n1/2 + c (the reference's crop convention). The predictor (u, v)
// at that pixel is the bilinear interpolation of the pair's field on the
// pass grid (per-pixel cell index and fraction tables, constant outside the
// grid's bounding box). The deformed crops are im0(y - v/2, x - u/2) and
// im1(y + v/2, x + u/2), sampled by Keys cubic convolution (a = -0.5; R. Keys,
// IEEE TASSP 29(6), 1981) on the image minus its minimum, zero outside the
// image; the predictor and the positions are FP64, the weights and the tap
// sums float32 (the images' own precision). A workgroup deforms one row:
// the predictor is first interpolated along y at every grid column (LDS),
// then along x per pixel; a row of 4 taps is one vector load; inside the
// image the minimum is taken out of the sample once, as m times the sum of
// the 16 weights. The pixel's predictor does not depend on the window, so
// the images are deformed once (float32) and the windows are cropped from
// them by the fused correlation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/piv_hip.h"
#include "last_error.hpp"

namespace {

constexpr int kDeformThreads = 256;
constexpr int kMaxGridColumns = 4096;  // the row's predictor in LDS: 64 KB

// a sample position kept within 8 px of the image: every tap of a position
// beyond that is outside the image (zero) either way, and the tap indices
// of the clamped value stay far from integer overflow (NaN goes to -8)
__device__ __forceinline__ double clamp_pos(double p, int n) {
  return fmin(fmax(p, -8.0), double(n) + 8.0);
}

template <int DT>
__device__ __forceinline__ float pixel(const void* data, int64_t idx);
template <>
__device__ __forceinline__ float pixel<PIV_U8>(const void* data, int64_t idx) {
  return float(((const __attribute__((address_space(1))) uint8_t*)data)[idx]);
}
template <>
__device__ __forceinline__ float pixel<PIV_U16>(const void* data, int64_t idx) {
  return float(((const __attribute__((address_space(1))) uint16_t*)data)[idx]);
}
template <>
__device__ __forceinline__ float pixel<PIV_F32>(const void* data, int64_t idx) {
  return ((const __attribute__((address_space(1))) float*)data)[idx];
}

// Keys weights of the taps at offsets -1, 0, 1, 2 for the fraction t in
// [0, 1), a = -0.5: the cubic pieces expanded in t (Catmull-Rom form),
// w0 = (-t^3 + 2 t^2 - t) / 2, w1 = (3 t^3 - 5 t^2 + 2) / 2,
// w2 = (-3 t^3 + 4 t^2 + t) / 2, w3 = (t^3 - t^2) / 2
__device__ __forceinline__ void keys_weights(float t, float w[4]) {
  const float t2 = t * t, t3 = t2 * t;
  w[0] = 0.5f * (2.0f * t2 - t3 - t);
  w[1] = 0.5f * (3.0f * t3 - 5.0f * t2) + 1.0f;
  w[2] = 0.5f * (4.0f * t2 - 3.0f * t3 + t);
  w[3] = 0.5f * (t3 - t2);
}

// the 4 taps of a row of the image starting at element `idx` (in the image,
// element-aligned data)
template <int DT>
__device__ __forceinline__ void row_taps(const void* data, int64_t idx, float p[4]);
// The loads are aligned: a tap row at an arbitrary element is read as the
// aligned dwords that hold it and shifted into place (unaligned multi-dword
// loads cost the memory pipeline several passes). They read up to 3 bytes
// before the taps (in the same dword, so the same page) and up to 4 bytes
// after them (the next row, or kOverread more elements of the last row: the
// caller keeps the last columns of the last row out of this path).
template <int DT>
constexpr int kOverread = DT == PIV_U8 ? 4 : DT == PIV_U16 ? 2 : 0;

template <>
__device__ __forceinline__ void row_taps<PIV_U8>(const void* data, int64_t idx, float p[4]) {
  const uintptr_t byte = reinterpret_cast<uintptr_t>(data) + uintptr_t(idx);
  typedef uint32_t u32x2 __attribute__((ext_vector_type(2), aligned(4)));
  const u32x2 d = *((const __attribute__((address_space(1))) u32x2*)(byte & ~uintptr_t(3)));
  const uint32_t w = __builtin_amdgcn_alignbit(d[1], d[0], uint32_t(byte & 3) * 8);
#pragma unroll
  for (int b = 0; b < 4; ++b) p[b] = float((w >> (8 * b)) & 0xffu);  // 32-bit lanes
}
template <>
__device__ __forceinline__ void row_taps<PIV_U16>(const void* data, int64_t idx, float p[4]) {
  // three aligned dwords holding the 8 bytes, which start at byte 0 or 2 of
  // the first one
  const uintptr_t byte = reinterpret_cast<uintptr_t>(data) + uintptr_t(idx) * 2;
  typedef uint32_t u32x3 __attribute__((ext_vector_type(3), aligned(4)));
  const u32x3 w = *((const __attribute__((address_space(1))) u32x3*)(byte & ~uintptr_t(3)));
  const uint32_t sh = uint32_t(byte & 3) * 8;
  const uint32_t lo = __builtin_amdgcn_alignbit(w[1], w[0], sh);
  const uint32_t hi = __builtin_amdgcn_alignbit(w[2], w[1], sh);
  p[0] = float(lo & 0xffffu);
  p[1] = float(lo >> 16);
  p[2] = float(hi & 0xffffu);
  p[3] = float(hi >> 16);
}
template <>
__device__ __forceinline__ void row_taps<PIV_F32>(const void* data, int64_t idx, float p[4]) {
  typedef float f4 __attribute__((ext_vector_type(4), aligned(4)));
  const f4 w = *((const __attribute__((address_space(1))) f4*)(
      (const __attribute__((address_space(1))) float*)data + idx));
  p[0] = w[0];
  p[1] = w[1];
  p[2] = w[2];
  p[3] = w[3];
}

// The sample with tap origin (iy, ix) = floor(p) - 1 and fractions
// (ty, tx) of the position: float32 weights and 16-tap sum, zero outside
// the image (the position's split into origin and fraction is done in FP64
// by the caller; the fraction is then exact in float32 to 2^-24).
template <int DT>
__device__ float bicubic_taps(const piv_image& im, float m, int iy, int ix, float ty, float tx) {
  float wy[4], wx[4];
  keys_weights(ty, wy);
  keys_weights(tx, wx);
  float acc = 0.0f;
  if (iy >= 0 && iy <= im.height - 4 && ix >= 0 && ix <= im.width - 4 &&
      (iy < im.height - 4 || ix <= im.width - 4 - kOverread<DT>)) {
    // interior: the 16 taps inside the image, sum w p - m sum w
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      float p[4];
      row_taps<DT>(im.data, int64_t(iy + a) * im.pitch + ix, p);
      float ra = 0.0f;
#pragma unroll
      for (int b = 0; b < 4; ++b) ra += wx[b] * p[b];
      acc += wy[a] * ra;
    }
    const float sw = ((wy[0] + wy[1]) + (wy[2] + wy[3])) * ((wx[0] + wx[1]) + (wx[2] + wx[3]));
    return acc - m * sw;
  }
  for (int a = 0; a < 4; ++a) {
    const int yy = iy + a;
    if (yy < 0 || yy >= im.height) continue;
    const int64_t row = int64_t(yy) * im.pitch;
    float ra = 0.0f;
    for (int b = 0; b < 4; ++b) {
      const int xx = ix + b;
      if (xx < 0 || xx >= im.width) continue;
      ra += wx[b] * (pixel<DT>(im.data, row + xx) - m);
    }
    acc += wy[a] * ra;
  }
  return acc;
}

// Workgroups are dispatched to the 8 XCDs of the chip round-robin
// (workgroup b to XCD b % 8); renumbered so that each XCD deforms one
// contiguous band of rows, the source rows that neighbouring output rows
// share (the 4 taps plus the displacement) stay under one L2. A bijection
// of [0, nb).
__device__ __forceinline__ int xcd_band(int b, int nb) {
  constexpr int kXcd = 8;
  const int x = b % kXcd, j = b / kXcd;
  const int q = nb / kXcd, r = nb % kXcd;
  return x * q + min(x, r) + j;
}

// deformed images of one pair and row (logical workgroup pair * height + y):
// the predictor at a pixel does not depend on the window holding it, so the
// deformed crop of a window centred on a grid node is the crop of the
// deformed image there
template <int DT>
__global__ __launch_bounds__(kDeformThreads) void k_deform_images(
    const piv_image* __restrict__ images, piv_deform_grid g, const double* __restrict__ u,
    const double* __restrict__ v, int nvec, float* __restrict__ out, int64_t out_stride,
    int height) {
  const int lb = xcd_band(int(blockIdx.x), int(gridDim.x));
  const int pair = lb / height, y = lb % height;
  const piv_image im0 = images[2 * pair], im1 = images[2 * pair + 1];
  const float m0 = im0.minval ? *im0.minval : 0.0f;
  const float m1 = im1.minval ? *im1.minval : 0.0f;
  const double* U = u + int64_t(pair) * nvec;
  const double* V = v + int64_t(pair) * nvec;
  const int W = im0.width;
  float* out0 = out + int64_t(2 * pair) * out_stride + int64_t(y) * W;
  float* out1 = out + int64_t(2 * pair + 1) * out_stride + int64_t(y) * W;
  const int jy = g.cell_y[y];
  const double ty = g.t_y[y];
  const int jy1 = min(jy + 1, g.ny - 1);
  // the predictor interpolated along y at every grid column of this row
  extern __shared__ double rowuv[];  // u then v, g.nx each
  for (int j = threadIdx.x; j < g.nx; j += kDeformThreads) {
    const int64_t i0 = int64_t(jy) * g.nx + j, i1 = int64_t(jy1) * g.nx + j;
    rowuv[j] = (1.0 - ty) * U[i0] + ty * U[i1];
    rowuv[g.nx + j] = (1.0 - ty) * V[i0] + ty * V[i1];
  }
  __syncthreads();
  // independent pixels: unrolled so that their tap loads are in flight together
#pragma unroll 4
  for (int x = threadIdx.x; x < W; x += kDeformThreads) {
    // bilinear predictor at the pixel: along x between the row's values
    const int jx = g.cell_x[x];
    const double tx = g.t_x[x];
    const int jx1 = min(jx + 1, g.nx - 1);
    const double uu = (1.0 - tx) * rowuv[jx] + tx * rowuv[jx1];
    const double vv = (1.0 - tx) * rowuv[g.nx + jx] + tx * rowuv[g.nx + jx1];
    const double py0 = clamp_pos(y - 0.5 * vv, im0.height), px0 = clamp_pos(x - 0.5 * uu, W);
    const double py1 = clamp_pos(y + 0.5 * vv, im1.height),
                 px1 = clamp_pos(x + 0.5 * uu, im1.width);
    const double fy0 = floor(py0), fx0 = floor(px0), fy1 = floor(py1), fx1 = floor(px1);
    out0[x] = bicubic_taps<DT>(im0, m0, int(fy0) - 1, int(fx0) - 1, float(py0 - fy0),
                               float(px0 - fx0));
    out1[x] = bicubic_taps<DT>(im1, m1, int(fy1) - 1, int(fx1) - 1, float(py1 - fy1),
                               float(px1 - fx1));
  }
}

__global__ void k_add_predictor(int n, float* __restrict__ dx, float* __restrict__ dy,
                                const double* __restrict__ u, const double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dx[i] = float(double(dx[i]) + u[i]);
  dy[i] = float(double(dy[i]) + v[i]);
}

int deform_fail(int code, const std::string& msg) {
  fia::last_error() = msg;
  return code;
}

int deform_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return deform_fail(PIV_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return PIV_OK;
}

}  // namespace

extern "C" {

int piv_deform_images(int32_t dtype, const piv_image* images, int32_t npairs,
                      int32_t height, int32_t width, const piv_deform_grid* grid,
                      const double* u, const double* v, float* out, void* stream) {
  if (!images || !grid || !u || !v || !out) return deform_fail(PIV_EINVAL, "NULL argument");
  if (npairs <= 0 || height <= 0 || width <= 0 || grid->ny <= 0 || grid->nx <= 0)
    return deform_fail(PIV_EINVAL, "empty deformation problem");
  if (int64_t(npairs) * height > INT32_MAX)
    return deform_fail(PIV_EINVAL, "too many pairs or rows for one launch");
  if (!grid->iyvecs || !grid->ixvecs || !grid->cell_y || !grid->t_y || !grid->cell_x ||
      !grid->t_x)
    return deform_fail(PIV_EINVAL, "NULL grid table");
  const int nvec = grid->ny * grid->nx;
  if (grid->nx > kMaxGridColumns)
    return deform_fail(PIV_EINVAL, "deformation grid wider than 4096 nodes");
  const size_t lds = size_t(2) * grid->nx * sizeof(double);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 gr(static_cast<unsigned>(int64_t(height) * npairs));
  const dim3 bl(kDeformThreads);
  const int64_t stride = int64_t(height) * width;
  switch (dtype) {
    case PIV_U8:
      hipLaunchKernelGGL(k_deform_images<PIV_U8>, gr, bl, lds, s, images, *grid, u, v, nvec, out,
                         stride, int(height));
      break;
    case PIV_U16:
      hipLaunchKernelGGL(k_deform_images<PIV_U16>, gr, bl, lds, s, images, *grid, u, v, nvec, out,
                         stride, int(height));
      break;
    case PIV_F32:
      hipLaunchKernelGGL(k_deform_images<PIV_F32>, gr, bl, lds, s, images, *grid, u, v, nvec, out,
                         stride, int(height));
      break;
    default:
      return deform_fail(PIV_EINVAL, "unknown image dtype");
  }
  return deform_check("piv_deform_images");
}

int piv_add_predictor(int32_t n, float* dx, float* dy, const double* u, const double* v,
                      void* stream) {
  if (n <= 0) return PIV_OK;
  if (!dx || !dy || !u || !v) return deform_fail(PIV_EINVAL, "NULL argument");
  hipLaunchKernelGGL(k_add_predictor, dim3((n + 255) / 256), dim3(256), 0,
                     static_cast<hipStream_t>(stream), n, dx, dy, u, v);
  return deform_check("piv_add_predictor");
}

}  // extern "C"

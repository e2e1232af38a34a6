// Differentiable refinement of the final PIV field (gfx950, FP64). No
// counterpart in the reference: it stops after the last multipass
// (src/fluidimage/works/piv/multipass.py:185-188).
//
// Parameters: the displacement (Dx, Dy) at every node of the final vector
// grid (iyvecs x ixvecs, row-major), per pair. Field at a pixel: bilinear
// interpolation of the node values (cell tables of piv_deform_grid).
// Window w of a pair is centred on node w (the reference's crop convention);
// its samples are a_i = im0(p_i - d_i / 2), b_i = im1(p_i + d_i / 2) by Keys
// cubic convolution (a = -0.5) on the image minus its minimum, zero outside.
// Loss of the pair: sum over its windows of 1 - NCC(a, b) (the windowed
// normalised SSD is 2 (1 - NCC)); windows with a constant crop contribute 0.
// The backward pass recomputes each window's samples (nothing per pixel is
// stored), differentiates NCC, the Keys kernel and the bilinear weights, and
// accumulates the node gradients of a window in LDS before adding them to
// the global gradient.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/piv_hip.h"
#include "last_error.hpp"

namespace {

constexpr int kRefThreads = 256;
constexpr int kNodeSpan = 8;  // node rows / columns touched by one window, at most

template <int DT>
__device__ __forceinline__ double pixel(const void* data, int64_t idx);
template <>
__device__ __forceinline__ double pixel<PIV_U8>(const void* data, int64_t idx) {
  return double(static_cast<const uint8_t*>(data)[idx]);
}
template <>
__device__ __forceinline__ double pixel<PIV_U16>(const void* data, int64_t idx) {
  return double(static_cast<const uint16_t*>(data)[idx]);
}
template <>
__device__ __forceinline__ double pixel<PIV_F32>(const void* data, int64_t idx) {
  return double(static_cast<const float*>(data)[idx]);
}

// Keys weights (taps -1, 0, 1, 2) and their derivatives with respect to the
// sample position, for the fraction t in [0, 1)
__device__ __forceinline__ void keys(double t, double w[4], double dw[4]) {
  constexpr double a = -0.5;
  const double d0 = 1.0 + t, d1 = t, d2 = 1.0 - t, d3 = 2.0 - t;
  w[0] = ((a * d0 - 5.0 * a) * d0 + 8.0 * a) * d0 - 4.0 * a;
  w[1] = ((a + 2.0) * d1 - (a + 3.0)) * d1 * d1 + 1.0;
  w[2] = ((a + 2.0) * d2 - (a + 3.0)) * d2 * d2 + 1.0;
  w[3] = ((a * d3 - 5.0 * a) * d3 + 8.0 * a) * d3 - 4.0 * a;
  // d/dt of W(d(t)): dd0/dt = dd1/dt = 1, dd2/dt = dd3/dt = -1
  dw[0] = (3.0 * a * d0 - 10.0 * a) * d0 + 8.0 * a;
  dw[1] = (3.0 * (a + 2.0) * d1 - 2.0 * (a + 3.0)) * d1;
  dw[2] = -(3.0 * (a + 2.0) * d2 - 2.0 * (a + 3.0)) * d2;
  dw[3] = -((3.0 * a * d3 - 10.0 * a) * d3 + 8.0 * a);
}

// value and gradient (d/dy, d/dx) of the Keys interpolation at (py, px)
template <int DT>
__device__ void sample(const piv_image& im, double m, double py, double px, double& val,
                       double& gy, double& gx) {
  const double fy = floor(py), fx = floor(px);
  double wy[4], dwy[4], wx[4], dwx[4];
  keys(py - fy, wy, dwy);
  keys(px - fx, wx, dwx);
  const int iy = int(fy) - 1, ix = int(fx) - 1;
  double v = 0.0, vy = 0.0, vx = 0.0;
  for (int r = 0; r < 4; ++r) {
    const int yy = iy + r;
    if (yy < 0 || yy >= im.height) continue;
    const int64_t row = int64_t(yy) * im.pitch;
    double s = 0.0, sd = 0.0;
    for (int c = 0; c < 4; ++c) {
      const int xx = ix + c;
      if (xx < 0 || xx >= im.width) continue;
      const double p = pixel<DT>(im.data, row + xx) - m;
      s += wx[c] * p;
      sd += dwx[c] * p;
    }
    v += wy[r] * s;
    vy += dwy[r] * s;
    vx += wy[r] * sd;
  }
  val = v;
  gy = vy;
  gx = vx;
}

struct Pixel {
  int jy, jx, jy1, jx1;
  double ty, tx;
};

__device__ __forceinline__ Pixel locate(const piv_deform_grid& g, int y, int x) {
  Pixel p;
  p.jy = g.cell_y[y];
  p.jx = g.cell_x[x];
  p.ty = g.t_y[y];
  p.tx = g.t_x[x];
  p.jy1 = min(p.jy + 1, g.ny - 1);
  p.jx1 = min(p.jx + 1, g.nx - 1);
  return p;
}

__device__ __forceinline__ double bilinear(const double* F, int nx, const Pixel& p) {
  return (1.0 - p.ty) * ((1.0 - p.tx) * F[p.jy * nx + p.jx] + p.tx * F[p.jy * nx + p.jx1]) +
         p.ty * ((1.0 - p.tx) * F[p.jy1 * nx + p.jx] + p.tx * F[p.jy1 * nx + p.jx1]);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kRefThreads / 64; ++w) s += red[w];
  return s;
}

struct Window {
  int pair, yc, xc, y0, x0;
};

__device__ __forceinline__ Window window_of(const piv_deform_grid& g, int w, int nvec, int n0,
                                            int n1) {
  Window wi;
  wi.pair = w / nvec;
  const int node = w - wi.pair * nvec;
  wi.yc = g.iyvecs[node / g.nx];
  wi.xc = g.ixvecs[node % g.nx];
  wi.y0 = wi.yc - n0 / 2;
  wi.x0 = wi.xc - n1 / 2;
  return wi;
}

// forward: per window the sums (Sa, Sb, Saa, Sbb, Sab) and the loss of the pair
template <int DT>
__global__ __launch_bounds__(kRefThreads) void k_refine_forward(
    const piv_image* __restrict__ images, piv_deform_grid g, const double* __restrict__ dx,
    const double* __restrict__ dy, int n0, int n1, int nvec, double* __restrict__ stats,
    double* __restrict__ loss) {
  __shared__ double red[kRefThreads / 64];
  const int w = blockIdx.x;
  const Window wi = window_of(g, w, nvec, n0, n1);
  const piv_image im0 = images[2 * wi.pair], im1 = images[2 * wi.pair + 1];
  const double m0 = im0.minval ? double(*im0.minval) : 0.0;
  const double m1 = im1.minval ? double(*im1.minval) : 0.0;
  const double* DX = dx + int64_t(wi.pair) * nvec;
  const double* DY = dy + int64_t(wi.pair) * nvec;
  double sa = 0.0, sb = 0.0, saa = 0.0, sbb = 0.0, sab = 0.0;
  for (int e = threadIdx.x; e < n0 * n1; e += kRefThreads) {
    const int y = wi.y0 + e / n1, x = wi.x0 + e % n1;
    const Pixel p = locate(g, y, x);
    const double u = bilinear(DX, g.nx, p), v = bilinear(DY, g.nx, p);
    double a, b, gy, gx;
    sample<DT>(im0, m0, y - 0.5 * v, x - 0.5 * u, a, gy, gx);
    sample<DT>(im1, m1, y + 0.5 * v, x + 0.5 * u, b, gy, gx);
    sa += a;
    sb += b;
    saa += a * a;
    sbb += b * b;
    sab += a * b;
  }
  sa = block_sum(sa, red);
  sb = block_sum(sb, red);
  saa = block_sum(saa, red);
  sbb = block_sum(sbb, red);
  sab = block_sum(sab, red);
  if (threadIdx.x == 0) {
    const double N = double(n0 * n1);
    const double va = saa - sa * sa / N, vb = sbb - sb * sb / N, cab = sab - sa * sb / N;
    double* st = stats + 5 * int64_t(w);
    st[0] = sa;
    st[1] = sb;
    st[2] = va;
    st[3] = vb;
    st[4] = cab;
    if (va > 0.0 && vb > 0.0) atomicAdd(&loss[wi.pair], 1.0 - cab / sqrt(va * vb));
  }
}

// backward: d loss / d (Dx, Dy) scaled by grad_out[pair], added into gx, gy
template <int DT>
__global__ __launch_bounds__(kRefThreads) void k_refine_backward(
    const piv_image* __restrict__ images, piv_deform_grid g, const double* __restrict__ dx,
    const double* __restrict__ dy, int n0, int n1, int nvec, const double* __restrict__ stats,
    const double* __restrict__ grad_out, double* __restrict__ gdx, double* __restrict__ gdy) {
  __shared__ double acc[2][kNodeSpan][kNodeSpan];
  const int w = blockIdx.x;
  const Window wi = window_of(g, w, nvec, n0, n1);
  const double* st = stats + 5 * int64_t(w);
  const double N = double(n0 * n1);
  const double ma = st[0] / N, mb = st[1] / N, va = st[2], vb = st[3], cab = st[4];
  if (!(va > 0.0 && vb > 0.0)) return;  // uniform across the block
  const double go = grad_out[wi.pair];
  const double sab = sqrt(va * vb), ncc = cab / sab;
  // node rows / columns of the window's cells
  const int ny0 = g.cell_y[wi.y0], nx0 = g.cell_x[wi.x0];
  for (int e = threadIdx.x; e < 2 * kNodeSpan * kNodeSpan; e += kRefThreads)
    (&acc[0][0][0])[e] = 0.0;
  __syncthreads();
  const piv_image im0 = images[2 * wi.pair], im1 = images[2 * wi.pair + 1];
  const double m0 = im0.minval ? double(*im0.minval) : 0.0;
  const double m1 = im1.minval ? double(*im1.minval) : 0.0;
  const double* DX = dx + int64_t(wi.pair) * nvec;
  const double* DY = dy + int64_t(wi.pair) * nvec;
  double* GX = gdx + int64_t(wi.pair) * nvec;
  double* GY = gdy + int64_t(wi.pair) * nvec;
  for (int e = threadIdx.x; e < n0 * n1; e += kRefThreads) {
    const int y = wi.y0 + e / n1, x = wi.x0 + e % n1;
    const Pixel p = locate(g, y, x);
    const double u = bilinear(DX, g.nx, p), v = bilinear(DY, g.nx, p);
    double a, ay, ax, b, by, bx;
    sample<DT>(im0, m0, y - 0.5 * v, x - 0.5 * u, a, ay, ax);
    sample<DT>(im1, m1, y + 0.5 * v, x + 0.5 * u, b, by, bx);
    // loss_w = 1 - NCC: dL/da = -[(b - mb) / sab - NCC (a - ma) / va], same for b
    const double la = -((b - mb) / sab - ncc * (a - ma) / va);
    const double lb = -((a - ma) / sab - ncc * (b - mb) / vb);
    // a = im0(p - d/2), b = im1(p + d/2)
    const double lu = go * 0.5 * (lb * bx - la * ax);
    const double lv = go * 0.5 * (lb * by - la * ay);
    const double w00 = (1.0 - p.ty) * (1.0 - p.tx), w01 = (1.0 - p.ty) * p.tx;
    const double w10 = p.ty * (1.0 - p.tx), w11 = p.ty * p.tx;
    const int jys[2] = {p.jy, p.jy1}, jxs[2] = {p.jx, p.jx1};
    const double ws[2][2] = {{w00, w01}, {w10, w11}};
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) {
        const int ly = jys[r] - ny0, lx = jxs[c] - nx0;
        const double wt = ws[r][c];
        if (wt == 0.0) continue;
        if (ly >= 0 && ly < kNodeSpan && lx >= 0 && lx < kNodeSpan) {
          atomicAdd(&acc[0][ly][lx], wt * lu);
          atomicAdd(&acc[1][ly][lx], wt * lv);
        } else {
          atomicAdd(&GX[jys[r] * g.nx + jxs[c]], wt * lu);
          atomicAdd(&GY[jys[r] * g.nx + jxs[c]], wt * lv);
        }
      }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kNodeSpan * kNodeSpan; e += kRefThreads) {
    const int ly = e / kNodeSpan, lx = e % kNodeSpan;
    const int jy = ny0 + ly, jx = nx0 + lx;
    if (jy < g.ny && jx < g.nx) {
      if (acc[0][ly][lx] != 0.0) atomicAdd(&GX[jy * g.nx + jx], acc[0][ly][lx]);
      if (acc[1][ly][lx] != 0.0) atomicAdd(&GY[jy * g.nx + jx], acc[1][ly][lx]);
    }
  }
}

int refine_fail(int code, const std::string& msg) {
  fia::last_error() = msg;
  return code;
}

int refine_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return refine_fail(PIV_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return PIV_OK;
}

int refine_args(const piv_image* images, const piv_deform_grid* grid, const double* dx,
                const double* dy, int32_t npairs, int32_t n0, int32_t n1) {
  if (!images || !grid || !dx || !dy) return refine_fail(PIV_EINVAL, "NULL argument");
  if (npairs <= 0 || n0 <= 0 || n1 <= 0 || grid->ny <= 0 || grid->nx <= 0)
    return refine_fail(PIV_EINVAL, "empty refinement problem");
  if (!grid->iyvecs || !grid->ixvecs || !grid->cell_y || !grid->t_y || !grid->cell_x ||
      !grid->t_x)
    return refine_fail(PIV_EINVAL, "NULL grid table");
  if (int64_t(npairs) * grid->ny * grid->nx > 0x7fffffff)
    return refine_fail(PIV_EINVAL, "too many windows");
  return PIV_OK;
}

}  // namespace

extern "C" {

int piv_refine_forward(int32_t dtype, const piv_image* images, int32_t npairs,
                       const piv_deform_grid* grid, int32_t n0, int32_t n1, const double* dx,
                       const double* dy, double* stats, double* loss, void* stream) {
  if (int rc = refine_args(images, grid, dx, dy, npairs, n0, n1)) return rc;
  if (!stats || !loss) return refine_fail(PIV_EINVAL, "NULL argument");
  const int nvec = grid->ny * grid->nx;
  const dim3 gr(static_cast<unsigned>(npairs * nvec)), bl(kRefThreads);
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case PIV_U8:
      hipLaunchKernelGGL(k_refine_forward<PIV_U8>, gr, bl, 0, s, images, *grid, dx, dy, n0, n1,
                         nvec, stats, loss);
      break;
    case PIV_U16:
      hipLaunchKernelGGL(k_refine_forward<PIV_U16>, gr, bl, 0, s, images, *grid, dx, dy, n0, n1,
                         nvec, stats, loss);
      break;
    case PIV_F32:
      hipLaunchKernelGGL(k_refine_forward<PIV_F32>, gr, bl, 0, s, images, *grid, dx, dy, n0, n1,
                         nvec, stats, loss);
      break;
    default:
      return refine_fail(PIV_EINVAL, "unknown image dtype");
  }
  return refine_check("piv_refine_forward");
}

int piv_refine_backward(int32_t dtype, const piv_image* images, int32_t npairs,
                        const piv_deform_grid* grid, int32_t n0, int32_t n1, const double* dx,
                        const double* dy, const double* stats, const double* grad_out,
                        double* gdx, double* gdy, void* stream) {
  if (int rc = refine_args(images, grid, dx, dy, npairs, n0, n1)) return rc;
  if (!stats || !grad_out || !gdx || !gdy) return refine_fail(PIV_EINVAL, "NULL argument");
  const int nvec = grid->ny * grid->nx;
  const dim3 gr(static_cast<unsigned>(npairs * nvec)), bl(kRefThreads);
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case PIV_U8:
      hipLaunchKernelGGL(k_refine_backward<PIV_U8>, gr, bl, 0, s, images, *grid, dx, dy, n0, n1,
                         nvec, stats, grad_out, gdx, gdy);
      break;
    case PIV_U16:
      hipLaunchKernelGGL(k_refine_backward<PIV_U16>, gr, bl, 0, s, images, *grid, dx, dy, n0,
                         n1, nvec, stats, grad_out, gdx, gdy);
      break;
    case PIV_F32:
      hipLaunchKernelGGL(k_refine_backward<PIV_F32>, gr, bl, 0, s, images, *grid, dx, dy, n0,
                         n1, nvec, stats, grad_out, gdx, gdy);
      break;
    default:
      return refine_fail(PIV_EINVAL, "unknown image dtype");
  }
  return refine_check("piv_refine_backward");
}

}  // extern "C"

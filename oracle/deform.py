"""Oracle of the image deformation between passes (FP64, numpy).

TEST INFRASTRUCTURE (see oracle/__init__.py). The reference has no image
deformation: its later passes shift the windows by the rounded predictor
(src/fluidimage/works/piv/singlepass.py:752-753,822-825, TODO at :736-738).
This module states the method the GPU path implements, so that it can be
checked; it is not a restatement of reference code.

Method (central-difference window deformation with bicubic resampling):

* the predictor (u, v) is known on the vector grid of the pass
  (iyvecs x ixvecs, row-major); at a pixel (y, x) it is the bilinear
  interpolation of the four surrounding grid nodes, clamped to the grid's
  bounding box (constant extension outside it);
* window k, centred on grid node (yc, xc) = (iyvecs[i], ixvecs[j]), covers
  the pixels y = yc - n0//2 + r, x = xc - n1//2 + c (the reference's crop
  convention, singlepass.py:363-407);
* its deformed crops are im0(y - v/2, x - u/2) and im1(y + v/2, x + u/2)
  sampled by Keys cubic convolution (a = -0.5; R. Keys, "Cubic convolution
  interpolation for digital image processing", IEEE TASSP 29(6), 1981) on
  the image minus its minimum, with zero outside the image (the reference's
  zero padding, singlepass.py:196-208);
* the correlation of the two deformed crops gives a residual displacement;
  the vector is the residual plus the predictor at the grid node.

Pinned by closed forms in tests/test_oracle.py: the Keys weights sum to one
and reproduce polynomials of degree <= 2 exactly, sampling at integer
positions returns the pixels, the bilinear field reproduces affine fields,
a zero field gives the plain crops, an even integer field gives integer
shifted crops.
"""

import numpy as np

KEYS_A = -0.5


def keys_weights(t, a=KEYS_A):
    """Weights of the taps at offsets -1, 0, 1, 2 for the fractional
    position t in [0, 1) (array): shape t.shape + (4,)."""
    t = np.asarray(t, dtype=np.float64)
    d = np.stack([1.0 + t, t, 1.0 - t, 2.0 - t], axis=-1)
    near = (a + 2.0) * d**3 - (a + 3.0) * d**2 + 1.0
    far = a * d**3 - 5.0 * a * d**2 + 8.0 * a * d - 4.0 * a
    return np.where(d <= 1.0, near, np.where(d < 2.0, far, 0.0))


def bicubic_sample(img, py, px):
    """Keys cubic convolution of the float64 image ``img`` at (py, px)
    (arrays of one shape), zero outside the image."""
    img = np.asarray(img, dtype=np.float64)
    H, W = img.shape
    py = np.asarray(py, dtype=np.float64)
    px = np.asarray(px, dtype=np.float64)
    fy = np.floor(py)
    fx = np.floor(px)
    wy = keys_weights(py - fy)
    wx = keys_weights(px - fx)
    fy = fy.astype(np.int64)
    fx = fx.astype(np.int64)
    out = np.zeros(py.shape)
    for a in range(4):
        yy = fy + a - 1
        iny = (yy >= 0) & (yy < H)
        for b in range(4):
            xx = fx + b - 1
            ok = iny & (xx >= 0) & (xx < W)
            val = np.where(ok, img[np.clip(yy, 0, H - 1), np.clip(xx, 0, W - 1)], 0.0)
            out += wy[..., a] * wx[..., b] * val
    return out


def _cell(nodes, p):
    """Index j of the grid cell [nodes[j], nodes[j+1]] holding p and the
    fraction t in [0, 1], with p clamped to [nodes[0], nodes[-1]]."""
    nodes = np.asarray(nodes, dtype=np.float64)
    if nodes.size == 1:
        return np.zeros(np.shape(p), np.int64), np.zeros(np.shape(p))
    q = np.clip(p, nodes[0], nodes[-1])
    j = np.clip(np.searchsorted(nodes, q, side="right") - 1, 0, nodes.size - 2)
    t = (q - nodes[j]) / (nodes[j + 1] - nodes[j])
    return j, t


def bilinear_field(iyvecs, ixvecs, u, y, x):
    """Bilinear interpolation of the grid field u (ny, nx) at (y, x)."""
    u = np.asarray(u, dtype=np.float64).reshape(len(iyvecs), len(ixvecs))
    jy, ty = _cell(iyvecs, np.asarray(y, dtype=np.float64))
    jx, tx = _cell(ixvecs, np.asarray(x, dtype=np.float64))
    jy1 = np.minimum(jy + 1, u.shape[0] - 1)
    jx1 = np.minimum(jx + 1, u.shape[1] - 1)
    return ((1 - ty) * ((1 - tx) * u[jy, jx] + tx * u[jy, jx1])
            + ty * ((1 - tx) * u[jy1, jx] + tx * u[jy1, jx1]))


def deformed_crops(im0, im1, iyvecs, ixvecs, u, v, shape):
    """Deformed crops (nvec, n0, n1) float64 of both images for every grid
    node (row-major), with the images' minima subtracted."""
    n0, n1 = shape
    a0 = np.asarray(im0, dtype=np.float64)
    a1 = np.asarray(im1, dtype=np.float64)
    a0 = a0 - a0.min()
    a1 = a1 - a1.min()
    yc, xc = np.meshgrid(np.asarray(iyvecs), np.asarray(ixvecs), indexing="ij")
    yc = yc.ravel()[:, None, None]
    xc = xc.ravel()[:, None, None]
    y = yc - n0 // 2 + np.arange(n0)[None, :, None] + 0.0 * xc
    x = xc - n1 // 2 + np.arange(n1)[None, None, :] + 0.0 * yc
    uu = bilinear_field(iyvecs, ixvecs, u, y, x)
    vv = bilinear_field(iyvecs, ixvecs, v, y, x)
    c0 = bicubic_sample(a0, y - 0.5 * vv, x - 0.5 * uu)
    c1 = bicubic_sample(a1, y + 0.5 * vv, x + 0.5 * uu)
    return c0, c1

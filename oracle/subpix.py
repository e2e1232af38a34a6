"""Oracle: sub-pixel peak refinement (test infrastructure).

Restates src/fluidimage/calcul/subpix.py:16-218 (``compute_subpix_2d_gaussian2``
and ``SubPix``). Arithmetic is float64 on the float32 map values.
"""

import numpy as np


class SubPixError(Exception):
    def __init__(self, explanation, code):
        super().__init__(explanation)
        self.explanation = explanation
        self.code = code


def subpix_2d_gaussian2(correl, ix, iy):
    """src/fluidimage/calcul/subpix.py:16-61. Returns (deplx, deply)."""
    crop = np.array(correl[iy - 1 : iy + 2, ix - 1 : ix + 2], dtype=np.float64)
    crop = crop - crop.min()
    crop = crop / crop.max()
    crop[crop == 0] = 1e-8
    logc = np.log(crop)
    x = np.array([-1.0, 0.0, 1.0])
    X, Y = np.meshgrid(x, x)  # X along columns (i1), Y along rows (i0)
    c10 = np.sum(X * logc) / 6
    c01 = np.sum(Y * logc) / 6
    c11 = np.sum(X * Y * logc) / 4
    c20 = np.sum((3 * X**2 - 2) * logc) / 6
    c02 = np.sum((3 * Y**2 - 2) * logc) / 6
    den = 4 * c20 * c02 - c11**2
    with np.errstate(divide="ignore", invalid="ignore"):
        deplx = (c11 * c01 - 2 * c10 * c02) / den
        deply = (c11 * c10 - 2 * c01 * c20) / den
    return float(deplx), float(deply)


def gaussian_fit_matrix(nsubpix):
    """The least-squares fit matrix of SubPix.prepare_subpix
    (src/fluidimage/calcul/subpix.py:86-105): the pseudo-inverse of
    M = [X^2, Y^2, X, Y, 1] on the (2 n + 1)^2 grid, shape (5, (2 n + 1)^2)."""
    xs = np.arange(-nsubpix, nsubpix + 1, dtype=float)
    X, Y = np.meshgrid(xs, xs)
    X = X.ravel()
    Y = Y.ravel()
    M = np.stack([X**2, Y**2, X, Y, np.ones_like(X)], axis=1)
    return np.linalg.pinv(M)


def subpix_2d_gaussian(correl, ix, iy, nsubpix=1):
    """src/fluidimage/calcul/subpix.py:157-181: the log-quadratic least-squares
    fit on the (2 n + 1)^2 crop (values <= 0 set to 1e-6 in a copy, float32
    log, float64 fit). Returns (deplx, deply), or None where the reference
    falls back to the centroid (a coefficient of a square term > 0, or a NaN
    displacement)."""
    crop = np.array(
        correl[iy - nsubpix : iy + nsubpix + 1, ix - nsubpix : ix + nsubpix + 1],
        dtype=np.float32,
    ).ravel()
    crop[crop <= 0.0] = 1e-6
    coef = gaussian_fit_matrix(nsubpix) @ np.log(crop).astype(np.float64)
    if coef[0] > 0 or coef[1] > 0:
        return None
    with np.errstate(divide="ignore", invalid="ignore"):
        sigmax = 1 / np.sqrt(-2 * coef[0])
        sigmay = 1 / np.sqrt(-2 * coef[1])
        deplx = coef[2] * sigmax**2
        deply = coef[3] * sigmay**2
    if np.isnan(deplx) or np.isnan(deply):
        return None
    return float(deplx), float(deply)


def subpix_centroid(correl, ix, iy, nsubpix=1):
    """src/fluidimage/calcul/subpix.py:189-198."""
    crop = np.array(
        correl[iy - nsubpix : iy + nsubpix + 1, ix - nsubpix : ix + nsubpix + 1],
        dtype=np.float64,
    )
    xs = np.arange(-nsubpix, nsubpix + 1, dtype=float)
    X, Y = np.meshgrid(xs, xs)
    s = np.sum(crop)
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.sum(X * crop) / s), float(np.sum(Y * crop) / s)


class SubPix:
    """src/fluidimage/calcul/subpix.py:64-218 (methods 2d_gaussian,
    2d_gaussian2, centroid and no_subpix)."""

    methods = ["2d_gaussian", "2d_gaussian2", "centroid", "no_subpix"]

    def __init__(self, method="centroid", nsubpix=None):
        if method == "2d_gaussian2" and nsubpix is not None and nsubpix != 1:
            raise ValueError(
                "Subpixel method '2d_gaussian2' doesn't require nsubpix. "
                "In this case, nsubpix has to be equal to None."
            )
        if method not in self.methods:
            raise ValueError(f"method has to be in {self.methods}")
        self.method = method
        self.n = 1 if nsubpix is None else nsubpix

    def compute_subpix(self, correl, ix, iy, method=None):
        """src/fluidimage/calcul/subpix.py:107-218 -> (ix + deplx, iy + deply)."""
        if method is None:
            method = self.method
        nsubpix = self.n
        ny, nx = correl.shape
        if (
            iy - nsubpix < 0
            or iy + nsubpix + 1 > ny
            or ix - nsubpix < 0
            or ix + nsubpix + 1 > nx
        ):
            raise SubPixError("close boundary", 3)
        if method == "2d_gaussian":
            d = subpix_2d_gaussian(correl, int(ix), int(iy), nsubpix)
            if d is None:
                return self.compute_subpix(correl, ix, iy, method="centroid")
            deplx, deply = d
        elif method == "2d_gaussian2":
            deplx, deply = subpix_2d_gaussian2(correl, int(ix), int(iy))
        elif method == "centroid":
            deplx, deply = subpix_centroid(correl, ix, iy, nsubpix)
        else:
            deplx = deply = 0.0
        if deplx**2 + deply**2 > 2 * (0.5 + nsubpix) ** 2:
            if method == "2d_gaussian2":
                return self.compute_subpix(correl, ix, iy, method="centroid")
            raise SubPixError("wrong subpix", 4)
        return deplx + ix, deply + iy

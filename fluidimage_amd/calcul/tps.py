"""Thin-plate-spline smoothing on subdomains, FP64 on the GPU.

Same method and parameters as the reference's ``ThinPlateSplineSubdom``
(src/fluidimage/calcul/interpolate/thin_plate_spline_subdom.py:25-280) and
``compute_tps_weights`` (src/fluidimage/calcul/interpolate/thin_plate_spline.py:168-216):

* subdomain grid, rescaled smoothing, 20 % buffers, half-open membership,
  partition-of-unity weights (1 - x^2)(1 - y^2) + 0.001 -- integer/geometry
  bookkeeping done here on the host (O(number of vectors));
* per subdomain, the saddle-point system [[K + sI, P], [P^T, 0]] with
  K = r^2 ln(r^2) / 2 and P = [1, y, x] is solved in FP64 by libpiv_hip:
  Householder QR of P, congruence of K + sI onto the null space of P^T
  (symmetric positive definite there) and a blocked Cholesky factorisation,
  reused for the x and y components and for every outlier iteration
  (threshold_tps, at most 10 solves, thin_plate_spline_subdom.py:236-280);
* blending of the smoothed values and evaluation at the new positions on
  the GPU.

The reference solves the same systems with LAPACK's LU (``np.linalg.solve``);
both give the unique solution up to FP64 rounding.
"""

import ctypes
from math import sqrt

import numpy as np

from .. import _native


class SingularSystemError(ArithmeticError):
    """A subdomain system is singular (the reference's LinAlgError)."""


def _membership(xs, ys, lmin_x, lmax_x, lmin_y, lmax_y):
    """(d, i) pairs of points i inside subdomain d = iy * nsx + ix, with the
    reference's half-open tests, sorted by d then i."""
    nsx = len(lmin_x)
    in_x = (xs[None, :] >= lmin_x[:, None]) & (xs[None, :] < lmax_x[:, None])
    in_y = (ys[None, :] >= lmin_y[:, None]) & (ys[None, :] < lmax_y[:, None])
    ix_i, ix_x = np.nonzero(in_x.T)  # (point, ix) sorted by point
    iy_i, iy_y = np.nonzero(in_y.T)
    npts = len(xs)
    cx = np.bincount(ix_i, minlength=npts)
    cy = np.bincount(iy_i, minlength=npts)
    sx = np.concatenate([[0], np.cumsum(cx)])
    sy = np.concatenate([[0], np.cumsum(cy)])
    reps = cx * cy
    tot = int(reps.sum())
    pts = np.repeat(np.arange(npts), reps)
    # position of each pair inside its point's cx * cy block
    start = np.concatenate([[0], np.cumsum(reps)])[:-1]
    k = np.arange(tot) - np.repeat(start, reps)
    cyr = np.repeat(cy, reps)
    kx = k // np.maximum(cyr, 1)
    ky = k % np.maximum(cyr, 1)
    dx_ = ix_x[sx[pts] + kx]
    dy_ = iy_y[sy[pts] + ky]
    dom = dy_ * nsx + dx_
    order = np.lexsort((pts, dom))
    return dom[order], pts[order]


class ThinPlateSplineSubdomHIP:
    """GPU version of ThinPlateSplineSubdom (``centers = vstack([ys, xs])``)."""

    def __init__(self, centers, subdom_size, smoothing_coef, threshold=None,
                 percent_buffer_area=20):
        centers = np.asarray(centers, dtype=np.float64)
        self.centers = centers
        self.subdom_size = subdom_size
        self.threshold = threshold
        xs = centers[1]
        ys = centers[0]
        self.num_centers = xs.size
        x_min, x_max = xs.min(), xs.max()
        y_min, y_max = ys.min(), ys.max()
        range_x = x_max - x_min
        range_y = y_max - y_min
        if range_x == 0 or range_y == 0:
            raise SingularSystemError("degenerate set of centers")
        aspect_ratio = range_y / range_x
        nb_subdom = xs.size / subdom_size
        nb_subdomx = int(np.floor(sqrt(nb_subdom / aspect_ratio))) or 1
        nb_subdomy = int(np.ceil(sqrt(nb_subdom * aspect_ratio))) or 1
        self.nb_subdomx = nb_subdomx
        self.nb_subdomy = nb_subdomy
        self.nb_subdom = nb_subdomx * nb_subdomy
        x_dom = np.linspace(x_min, x_max, nb_subdomx + 1)
        y_dom = np.linspace(y_min, y_max, nb_subdomy + 1)
        num_vectors_per_subdom = xs.size / self.nb_subdom
        self.smoothing_coef = smoothing_coef * sqrt(
            (x_dom[1] - x_dom[0]) * (y_dom[1] - y_dom[0]) / num_vectors_per_subdom
        )
        coef_buffer = percent_buffer_area / 100
        bx = coef_buffer * range_x / nb_subdomx
        by = coef_buffer * range_y / nb_subdomy
        self.limits_min_x = x_dom[:-1] - bx
        self.limits_max_x = x_dom[1:] + bx
        self.limits_min_y = y_dom[:-1] - by
        self.limits_max_y = y_dom[1:] + by
        ix = np.tile(np.arange(nb_subdomx), nb_subdomy)
        iy = np.repeat(np.arange(nb_subdomy), nb_subdomx)
        self.xmax_domains = self.limits_max_x[ix]
        self.ymax_domains = self.limits_max_y[iy]
        self.xc_domains = 0.5 * (self.limits_min_x[ix] + self.limits_max_x[ix])
        self.yc_domains = 0.5 * (self.limits_min_y[iy] + self.limits_max_y[iy])
        dom, pts = _membership(xs, ys, self.limits_min_x, self.limits_max_x,
                               self.limits_min_y, self.limits_max_y)
        self.slot_dom = dom
        self.slot_pt = pts
        self.dom_offsets = np.searchsorted(dom, np.arange(self.nb_subdom + 1)).astype(np.int64)
        self.indices_domains = [pts[self.dom_offsets[d]:self.dom_offsets[d + 1]]
                                for d in range(self.nb_subdom)]
        self.slot_coef = self._coef(xs[pts], ys[pts], dom)
        self.norm_coefs = np.zeros(self.num_centers)
        # accumulate in subdomain order like the reference (subdom.py:118-126)
        for d in range(self.nb_subdom):
            a, b = self.dom_offsets[d], self.dom_offsets[d + 1]
            self.norm_coefs[pts[a:b]] += self.slot_coef[a:b]
        if np.any(np.diff(self.dom_offsets) < 1):
            raise SingularSystemError("empty subdomain")

    def _coef(self, xs_d, ys_d, dom):
        """thin_plate_spline_subdom.py:208-222"""
        xc = self.xc_domains[dom]
        yc = self.yc_domains[dom]
        dx_max = self.xmax_domains[dom] - xc
        dy_max = self.ymax_domains[dom] - yc
        dx2 = (xs_d - xc) ** 2 / dx_max**2
        dy2 = (ys_d - yc) ** 2 / dy_max**2
        return (1 - dx2) * (1 - dy2) + 0.001

    # ------------------------------------------------------------------
    def smooth_and_interpolate(self, deltaxs, deltays, new_positions, stream=None):
        """compute_tps_weights_subdom for x and y, then init_with_new_positions
        + interpolate on ``new_positions = vstack([ys, xs])``.

        Returns (dxs_smooth, dys_smooth, dxs_interp, dys_interp, summary);
        the per-subdomain weights are kept in ``weights_x`` / ``weights_y``.
        """
        import torch

        dev = torch.device("cuda")
        lib = _native.load()
        nd = self.nb_subdom
        sizes = np.diff(self.dom_offsets)
        n_max = int(sizes.max())
        # per-subdomain matrix storage: n_d x n_d doubles
        mat_off = np.concatenate([[0], np.cumsum(sizes.astype(np.int64) ** 2)])
        new_positions = np.asarray(new_positions, dtype=np.float64)
        nys, nxs = new_positions[0], new_positions[1]
        ndom, npt = _membership(nxs, nys, self.limits_min_x, self.limits_max_x,
                                self.limits_min_y, self.limits_max_y)
        new_off = np.searchsorted(ndom, np.arange(nd + 1)).astype(np.int64)
        new_coef = self._coef(nxs[npt], nys[npt], ndom)
        # slots of every point, in subdomain order (for the ordered blends)
        c_order = np.lexsort((self.slot_dom, self.slot_pt))
        c_ptr = np.searchsorted(self.slot_pt[c_order], np.arange(self.num_centers + 1))
        n_order = np.lexsort((ndom, npt))
        n_ptr = np.searchsorted(npt[n_order], np.arange(len(nxs) + 1))

        def t(a, dtype):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(dev)

        xs = t(self.centers[1], np.float64)
        ys = t(self.centers[0], np.float64)
        vx = t(deltaxs, np.float32)
        vy = t(deltays, np.float32)
        d_off = t(self.dom_offsets, np.int64)
        d_pt = t(self.slot_pt, np.int32)
        d_coef = t(self.slot_coef, np.float64)
        d_matoff = t(mat_off, np.int64)
        mats = torch.empty(int(mat_off[-1]), dtype=torch.float64, device=dev)
        refl = torch.empty((nd, 16), dtype=torch.float64, device=dev)
        vwork = torch.empty(int(self.dom_offsets[-1]) * 6, dtype=torch.float64, device=dev)
        smooth_slots = torch.empty((2, int(self.dom_offsets[-1])), dtype=torch.float64,
                                   device=dev)
        weights = torch.empty((2, int(self.dom_offsets[-1]) + 3 * nd), dtype=torch.float64,
                              device=dev)
        stats = torch.zeros((nd, 6), dtype=torch.float64, device=dev)
        status = torch.zeros(nd, dtype=torch.int32, device=dev)
        c_slots = t(c_order, np.int32)
        c_ptr_t = t(c_ptr, np.int64)
        norm_c = t(self.norm_coefs, np.float64)
        smooth_out = torch.empty((2, self.num_centers), dtype=torch.float64, device=dev)
        n_slots = t(n_order, np.int32)
        n_ptr_t = t(n_ptr, np.int64)
        n_dom = t(ndom, np.int32)
        n_coef = t(new_coef, np.float64)
        n_x = t(nxs, np.float64)
        n_y = t(nys, np.float64)
        interp_out = torch.empty((2, len(nxs)), dtype=torch.float64, device=dev)
        thr = -1.0 if self.threshold is None else float(self.threshold)
        P = ctypes.c_void_p
        _native.check(
            lib.piv_tps_smooth(
                nd, n_max, P(xs.data_ptr()), P(ys.data_ptr()), P(vx.data_ptr()),
                P(vy.data_ptr()), P(d_off.data_ptr()), P(d_pt.data_ptr()),
                P(d_matoff.data_ptr()), float(self.smoothing_coef), thr, P(mats.data_ptr()),
                P(refl.data_ptr()), P(vwork.data_ptr()), P(smooth_slots.data_ptr()),
                int(self.dom_offsets[-1]), P(weights.data_ptr()),
                int(self.dom_offsets[-1]) + 3 * nd, P(stats.data_ptr()), P(status.data_ptr()),
                _native.stream_ptr(stream)),
            "piv_tps_smooth",
        )
        _native.check(
            lib.piv_tps_blend(
                self.num_centers, P(c_slots.data_ptr()), P(c_ptr_t.data_ptr()),
                P(d_coef.data_ptr()), P(norm_c.data_ptr()), P(smooth_slots.data_ptr()),
                int(self.dom_offsets[-1]), P(smooth_out.data_ptr()),
                _native.stream_ptr(stream)),
            "piv_tps_blend",
        )
        _native.check(
            lib.piv_tps_interpolate(
                len(nxs), P(n_x.data_ptr()), P(n_y.data_ptr()), P(n_slots.data_ptr()),
                P(n_ptr_t.data_ptr()), P(n_dom.data_ptr()), P(n_coef.data_ptr()),
                P(xs.data_ptr()), P(ys.data_ptr()), P(d_off.data_ptr()), P(d_pt.data_ptr()),
                P(weights.data_ptr()), int(self.dom_offsets[-1]) + 3 * nd,
                P(interp_out.data_ptr()), _native.stream_ptr(stream)),
            "piv_tps_interpolate",
        )
        torch.cuda.current_stream().synchronize()
        st = status.cpu().numpy()
        if st.any():
            raise SingularSystemError(
                f"{int((st != 0).sum())} subdomain system(s) not positive definite")
        w = weights.cpu().numpy()
        offs = self.dom_offsets
        self.weights_x = []
        self.weights_y = []
        for d in range(nd):
            a, b = offs[d], offs[d + 1]
            base_w = a + 3 * d
            self.weights_x.append(w[0, base_w:base_w + (b - a) + 3].copy())
            self.weights_y.append(w[1, base_w:base_w + (b - a) + 3].copy())
        s = stats.cpu().numpy()
        summary = {
            "nb_fixed_vectors": [int(v) for v in s[:, 3]],
            "max(Udiff)": [float(v) for v in s[:, 4]],
            "nb_iterations": [int(v) for v in s[:, 5]],
            "nb_fixed_vectors_tot": int(s[:, 3].sum()),
        }
        self.summary_x = {"nb_fixed_vectors": [int(v) for v in s[:, 0]],
                          "max(Udiff)": [float(v) for v in s[:, 1]],
                          "nb_iterations": [int(v) for v in s[:, 2]],
                          "nb_fixed_vectors_tot": int(s[:, 0].sum())}
        out = smooth_out.cpu().numpy()
        interp = interp_out.cpu().numpy()
        return out[0], out[1], interp[0], interp[1], summary

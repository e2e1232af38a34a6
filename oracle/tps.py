"""Oracle: thin-plate-spline smoothing on subdomains (test infrastructure).

Restates src/fluidimage/calcul/interpolate/thin_plate_spline.py:38-216 and
src/fluidimage/calcul/interpolate/thin_plate_spline_subdom.py:25-280, with
``numpy.linalg.solve`` (LAPACK dgesv) for the FP64 saddle-point systems, as
the reference does.
"""

from math import sqrt

import numpy as np


def compute_tps_matrix(new_pos, centers):
    """thin_plate_spline.py:38-104: [(N + 3), M] with r^2 ln(r^2) / 2."""
    new_pos = np.asarray(new_pos, np.float64)
    centers = np.asarray(centers, np.float64)
    d, m = new_pos.shape
    em = np.zeros((centers.shape[1], m))
    for ind_d in range(d):
        em += (new_pos[ind_d][None, :] - centers[ind_d][:, None]) ** 2
    nz = em != 0
    em[nz] = em[nz] * np.log(em[nz]) / 2
    return np.vstack([em, np.ones(m), new_pos])


def compute_tps_weights(centers, values, smoothing_coef):
    """thin_plate_spline.py:168-216 -> (smoothed_values, tps_weights)."""
    nb_dim, num = centers.shape
    rhs = np.hstack([values, np.zeros(nb_dim + 1)])
    em = compute_tps_matrix(centers, centers).T
    smoothing = np.hstack(
        [smoothing_coef * np.eye(num), np.zeros([num, nb_dim + 1])]
    )
    pm = np.hstack([np.ones([num, 1]), centers.T])
    im = np.vstack(
        [em + smoothing, np.hstack([pm.T, np.zeros([nb_dim + 1, nb_dim + 1])])]
    )
    weights = np.linalg.solve(im, rhs)
    return em @ weights, weights


class ThinPlateSplineSubdom:
    """thin_plate_spline_subdom.py:25-280 (``centers = vstack([ys, xs])``)."""

    def __init__(self, centers, subdom_size, smoothing_coef, threshold=None,
                 percent_buffer_area=20):
        self.centers = np.asarray(centers, np.float64)
        self.subdom_size = subdom_size
        self.threshold = threshold
        xs = self.centers[1]
        ys = self.centers[0]
        self.num_centers = xs.size
        x_min, x_max = xs.min(), xs.max()
        y_min, y_max = ys.min(), ys.max()
        range_x = x_max - x_min
        range_y = y_max - y_min
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
        buffer_length_x = coef_buffer * range_x / nb_subdomx
        buffer_length_y = coef_buffer * range_y / nb_subdomy
        self.limits_min_x = x_dom[:-1] - buffer_length_x
        self.limits_max_x = x_dom[1:] + buffer_length_x
        self.limits_min_y = y_dom[:-1] - buffer_length_y
        self.limits_max_y = y_dom[1:] + buffer_length_y
        self.xmax_domains = np.empty(self.nb_subdom)
        self.ymax_domains = np.empty(self.nb_subdom)
        self.xc_domains = np.empty(self.nb_subdom)
        self.yc_domains = np.empty(self.nb_subdom)
        self.indices_domains = []
        i_dom = 0
        for iy in range(nb_subdomy):
            for ix in range(nb_subdomx):
                xmin = self.limits_min_x[ix]
                xmax = self.limits_max_x[ix]
                ymin = self.limits_min_y[iy]
                ymax = self.limits_max_y[iy]
                self.indices_domains.append(
                    np.where((xs >= xmin) & (xs < xmax) & (ys >= ymin) & (ys < ymax))[0]
                )
                self.xmax_domains[i_dom] = xmax
                self.ymax_domains[i_dom] = ymax
                self.xc_domains[i_dom] = 0.5 * (xmin + xmax)
                self.yc_domains[i_dom] = 0.5 * (ymin + ymax)
                i_dom += 1
        self.norm_coefs = np.zeros(self.num_centers)
        self.norm_coefs_domains = []
        for i_dom in range(self.nb_subdom):
            ind = self.indices_domains[i_dom]
            coefs = self._compute_coef_norm(xs[ind], ys[ind], i_dom)
            self.norm_coefs_domains.append(coefs)
            self.norm_coefs[ind] += coefs

    def _compute_coef_norm(self, xs_domain, ys_domain, i_domain):
        """thin_plate_spline_subdom.py:208-222"""
        xc = self.xc_domains[i_domain]
        yc = self.yc_domains[i_domain]
        dx_max = self.xmax_domains[i_domain] - xc
        dy_max = self.ymax_domains[i_domain] - yc
        dx2 = (xs_domain - xc) ** 2 / dx_max**2
        dy2 = (ys_domain - yc) ** 2 / dy_max**2
        return (1 - dx2) * (1 - dy2) + 0.001

    def compute_tps_weights_iter(self, centers, values, smoothing_coef):
        """thin_plate_spline_subdom.py:236-280"""
        # a copy that keeps the caller's dtype: the reference passes the
        # float32 displacements, so replaced values are rounded to float32
        values = np.array(values)
        summary = {"nb_fixed_vectors": 0}
        smoothed, weights = compute_tps_weights(centers, values, smoothing_coef)
        count = 1
        if self.threshold is not None:
            differences = np.sqrt((smoothed - values) ** 2)
            ind = np.argwhere(differences > self.threshold)
            summary["max(Udiff)"] = max(differences)
            nb_fixed = 0
            while ind.size != 0:
                nb_fixed += ind.size
                values[ind] = smoothed[ind]
                smoothed, weights = compute_tps_weights(centers, values, smoothing_coef)
                differences = np.sqrt((smoothed - values) ** 2)
                ind = np.argwhere(differences > self.threshold)
                count += 1
                if count > 10:
                    break
            summary["nb_fixed_vectors"] = nb_fixed
        summary["nb_iterations"] = count
        return smoothed, weights, summary

    def compute_tps_weights_subdom(self, values):
        """thin_plate_spline_subdom.py:128-159"""
        smoothed_domains = [None] * self.nb_subdom
        weights_domains = [None] * self.nb_subdom
        summaries = [None] * self.nb_subdom
        for idx in range(self.nb_subdom):
            ind = self.indices_domains[idx]
            (smoothed_domains[idx], weights_domains[idx], summaries[idx]) = (
                self.compute_tps_weights_iter(
                    self.centers[:, ind], np.asarray(values)[ind], self.smoothing_coef
                )
            )
        smoothed = np.zeros(self.num_centers)
        summary = {"nb_fixed_vectors": [], "max(Udiff)": [], "nb_iterations": []}
        for idx in range(self.nb_subdom):
            ind = self.indices_domains[idx]
            smoothed[ind] += self.norm_coefs_domains[idx] * smoothed_domains[idx]
            for key in summary:
                summary[key].append(summaries[idx].get(key, 0))
        summary["nb_fixed_vectors_tot"] = sum(summary["nb_fixed_vectors"])
        smoothed /= self.norm_coefs
        return smoothed, weights_domains, summary

    def init_with_new_positions(self, new_positions):
        """thin_plate_spline_subdom.py:161-206"""
        new_positions = np.asarray(new_positions)
        xs = new_positions[1]
        ys = new_positions[0]
        self.num_new_positions = xs.size
        self.ind_new_positions_domains = []
        for iy in range(self.nb_subdomy):
            for ix in range(self.nb_subdomx):
                self.ind_new_positions_domains.append(
                    np.where(
                        (xs >= self.limits_min_x[ix])
                        & (xs < self.limits_max_x[ix])
                        & (ys >= self.limits_min_y[iy])
                        & (ys < self.limits_max_y[iy])
                    )[0]
                )
        self.norm_coefs_new_pos = np.zeros(self.num_new_positions)
        self.norm_coefs_new_pos_domains = []
        for i_dom in range(self.nb_subdom):
            ind = self.ind_new_positions_domains[i_dom]
            coefs = self._compute_coef_norm(xs[ind], ys[ind], i_dom)
            self.norm_coefs_new_pos_domains.append(coefs)
            self.norm_coefs_new_pos[ind] += coefs
        self.tps_matrices = [None] * self.nb_subdom
        for i_dom in range(self.nb_subdom):
            self.tps_matrices[i_dom] = compute_tps_matrix(
                new_positions[:, self.ind_new_positions_domains[i_dom]],
                self.centers[:, self.indices_domains[i_dom]],
            )

    def interpolate(self, weights_domains):
        """thin_plate_spline_subdom.py:224-234"""
        values = np.zeros(self.num_new_positions)
        for i_dom in range(self.nb_subdom):
            values[self.ind_new_positions_domains[i_dom]] += (
                self.norm_coefs_new_pos_domains[i_dom]
                * np.dot(weights_domains[i_dom], self.tps_matrices[i_dom])
            )
        values /= self.norm_coefs_new_pos
        return values


def smoothing_operator(centers, subdom_size, smoothing_coef, percent_buffer_area=20):
    """Dense T (num_centers x num_centers) with T v = the smoothed values of
    ThinPlateSplineSubdom.compute_tps_weights_subdom(v) without outlier
    iterations (thin_plate_spline_subdom.py:128-159, threshold None): per
    subdomain, the rows of compute_tps_weights' solve (thin_plate_spline.py:
    168-216) give smoothed = v - s w with w = inv(M)[:n, :n] v; the blend
    weights as the reference's."""
    tps = ThinPlateSplineSubdom(centers, subdom_size, smoothing_coef, None,
                                percent_buffer_area)
    num = tps.num_centers
    out = np.zeros((num, num))
    s = tps.smoothing_coef
    for i_dom in range(tps.nb_subdom):
        ind = tps.indices_domains[i_dom]
        n = ind.size
        c = tps.centers[:, ind]
        em = compute_tps_matrix(c, c).T
        smoothing = np.hstack([s * np.eye(n), np.zeros([n, 3])])
        pm = np.hstack([np.ones([n, 1]), c.T])
        im = np.vstack([em + smoothing, np.hstack([pm.T, np.zeros([3, 3])])])
        sw = np.linalg.inv(im)[:n, :n]
        blk = tps.norm_coefs_domains[i_dom][:, None] * (np.eye(n) - s * sw)
        out[np.ix_(ind, ind)] += blk
    out /= tps.norm_coefs[:, None]
    return out

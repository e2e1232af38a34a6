"""Pins of the oracle functions that restate WorkFIX, the TPS subdomain
geometry, the TPS outlier loop, the interpolation selector and the
secondary-peak search (CPU only).

Each expected value below is either a hand-computed literal (the arithmetic
is in the comment beside it) or the output of an independent library routine
(scipy's RBFInterpolator for the thin-plate spline, scipy's griddata), chosen
so that a plausible slip in the restatement (floor/ceil swapped, a closed
instead of half-open limit, a squared instead of unsquared threshold, a
missing division, x and y transposed) fails one of them.
"""

import math

import numpy as np
import pytest
from scipy.interpolate import RBFInterpolator, griddata

from oracle import correl as oc
from oracle import fix as ofix
from oracle import tps as otps
from oracle.piv import OracleWorkPIV, OracleWorkPIVPass, default_params

# ---------------------------------------------------------------- TPS geometry


def _grid_centers(x0, x1, y0, y1, step):
    ys, xs = np.mgrid[y0:y1 + step / 2:step, x0:x1 + step / 2:step]
    return np.vstack([ys.ravel().astype(float), xs.ravel().astype(float)])


def test_subdomain_counts_2048_last_pass():
    """2048 x 2048 at 16 px windows, 50 % overlap: centres 8, 16, ..., 2040
    (255 x 255 = 65,025). nb = 65025 / 200 = 325.125, sqrt = 18.03:
    floor -> 18 columns, ceil -> 19 rows, 342 subdomains."""
    centers = _grid_centers(8, 2040, 8, 2040, 8)
    assert centers.shape[1] == 65025
    t = otps.ThinPlateSplineSubdom(centers, 200, 2.0, threshold=1.5)
    assert (t.nb_subdomx, t.nb_subdomy, t.nb_subdom) == (18, 19, 342)
    # buffers 0.2 * 2032 / 18 = 22.5777... and 0.2 * 2032 / 19 = 21.3894...
    assert t.limits_min_x[0] == pytest.approx(-14.577777777777778, abs=1e-12)
    assert t.limits_max_x[-1] == pytest.approx(2062.5777777777776, abs=1e-12)
    assert t.limits_min_y[0] == pytest.approx(-13.389473684210527, abs=1e-12)
    assert t.limits_max_y[-1] == pytest.approx(2061.3894736842103, abs=1e-12)
    # x_dom[1] = 8 + 2032 / 18, so limits_max_x[0] = 120.888... + 22.577...
    assert t.limits_max_x[0] == pytest.approx(143.46666666666667, abs=1e-12)
    assert t.limits_min_x[1] == pytest.approx(98.31111111111111, abs=1e-12)
    # smoothing 2.0 * sqrt((2032/18)(2032/19) / (65025/342)) = 2 * 2032 / 255
    assert t.smoothing_coef == pytest.approx(15.937254901960784, rel=1e-13)


def test_subdomain_counts_with_missing_vectors():
    """64,180 valid vectors (845 interior ones removed, same bounds):
    nb = 320.9, sqrt = 17.91: floor -> 17, ceil -> 18, 306 subdomains."""
    centers = _grid_centers(8, 2040, 8, 2040, 8)
    rng = np.random.default_rng(0)
    ys, xs = centers
    interior = np.flatnonzero((xs > 8) & (xs < 2040) & (ys > 8) & (ys < 2040))
    drop = rng.choice(interior, 845, replace=False)
    keep = np.setdiff1d(np.arange(centers.shape[1]), drop)
    t = otps.ThinPlateSplineSubdom(centers[:, keep], 200, 2.0)
    assert t.num_centers == 64180
    assert (t.nb_subdomx, t.nb_subdomy, t.nb_subdom) == (17, 18, 306)


def test_subdomain_counts_follow_the_aspect_ratio():
    """21 x 11 centres on x in [0, 200], y in [0, 100]: aspect 0.5,
    nb = 231 / 20 = 11.55; x: floor(sqrt(23.1)) = 4, y: ceil(sqrt(5.775)) = 3
    (a swap of the two rules or of x and y gives 2 x 5)."""
    t = otps.ThinPlateSplineSubdom(_grid_centers(0, 200, 0, 100, 10), 20, 1.0)
    assert (t.nb_subdomx, t.nb_subdomy) == (4, 3)


def _tiny():
    """11 x 11 centres 0, 10, ..., 100 and subdom_size 30.25: nb = 4 exactly,
    so 2 x 2 subdomains; x_dom = y_dom = [0, 50, 100] and buffers
    0.2 * 100 / 2 = 10: limits [-10, 60) and [40, 110), all exact."""
    return otps.ThinPlateSplineSubdom(_grid_centers(0, 100, 0, 100, 10), 30.25, 5.0)


def test_half_open_membership_on_exact_limits():
    t = _tiny()
    assert (t.nb_subdomx, t.nb_subdomy) == (2, 2)
    assert np.array_equal(t.limits_min_x, [-10.0, 40.0])
    assert np.array_equal(t.limits_max_x, [60.0, 110.0])
    # columns of domain 0: x = 0..50 (x = 60 is its open limit): 6;
    # domain 1: x = 40..100: 7; the counts of the 4 domains are the products
    sizes = [len(ind) for ind in t.indices_domains]
    assert sizes == [36, 42, 42, 49]
    ys, xs = t.centers
    d0 = set(t.indices_domains[0].tolist())
    d3 = set(t.indices_domains[3].tolist())
    at = lambda x, y: int(np.flatnonzero((xs == x) & (ys == y))[0])  # noqa: E731
    assert at(60, 0) not in d0 and at(50, 50) in d0 and at(40, 40) in d0 and at(40, 40) in d3
    assert at(100, 100) in d3 and at(30, 30) not in d3
    # smoothing rescaled by sqrt(50 * 50 / (121 / 4)) = 50 / 5.5
    assert t.smoothing_coef == pytest.approx(5.0 * 50 / 5.5, rel=1e-14)


def test_partition_weights_literals():
    """(1 - x^2)(1 - y^2) + 0.001 with x, y relative to the buffered domain:
    domain 0 has centre (25, 25) and half-widths 35."""
    t = _tiny()
    w = t._compute_coef_norm(np.array([25.0, 0.0, 50.0]), np.array([25.0, 0.0, 10.0]), 0)
    # (25, 25): 1.001; (0, 0): (1 - 625/1225)^2 + 0.001;
    # (50, 10): (1 - 625/1225)(1 - 225/1225) + 0.001
    assert w == pytest.approx([1.001, 0.24090004164931278, 0.40083340274885465], rel=1e-14)
    # new positions take the same weights (init_with_new_positions)
    t.init_with_new_positions(np.array([[25.0, 5.0], [25.0, 45.0]]))
    assert t.ind_new_positions_domains[0].tolist() == [0, 1]
    assert t.ind_new_positions_domains[1].tolist() == [1]  # x = 45 is also in [40, 110)
    assert t.norm_coefs_new_pos[0] == pytest.approx(1.001, rel=1e-14)


# ---------------------------------------------------------------- TPS solve


def _rbf(centers, values, s):
    """scipy's thin-plate spline: kernel r^2 log r (= r^2 ln(r^2) / 2), the
    smoothing added to the kernel diagonal, degree-1 polynomial: the system
    the reference assembles (thin_plate_spline.py:204-214)."""
    return RBFInterpolator(np.asarray(centers).T, values, kernel="thin_plate_spline",
                           smoothing=s, degree=1)


@pytest.mark.parametrize("s", [0.0, 0.7, 6.0])
def test_compute_tps_weights_vs_scipy_rbf(s):
    rng = np.random.default_rng(5)
    centers = rng.uniform(0, 50, (2, 40))
    values = np.sin(centers[0] / 9) + 0.1 * centers[1] + rng.normal(0, 0.2, 40)
    smoothed, weights = otps.compute_tps_weights(centers, values, s)
    rbf = _rbf(centers, values, s)
    assert np.allclose(smoothed, rbf(centers.T), atol=1e-9)
    new = rng.uniform(0, 50, (2, 25))
    ev = weights @ otps.compute_tps_matrix(new, centers)
    assert np.allclose(ev, rbf(new.T), atol=1e-9)
    # the centre weights sum to zero and are orthogonal to x and y
    w = weights[:40]
    assert abs(w.sum()) < 1e-9 and abs(w @ centers[0]) < 1e-7 and abs(w @ centers[1]) < 1e-7


def test_subdomain_interpolation_vs_scipy_rbf_per_domain():
    """interpolate = sum over domains of weight * (that domain's spline at the
    point) / sum of weights, each domain's spline being scipy's RBF on its own
    members (no outlier iteration)."""
    t = otps.ThinPlateSplineSubdom(_grid_centers(0, 200, 0, 100, 10), 20, 1.5)
    ys, xs = t.centers
    values = np.cos(xs / 40) * (ys / 50) + 0.002 * xs * ys
    _, wdoms, _ = t.compute_tps_weights_subdom(values)
    rng = np.random.default_rng(1)
    new = np.vstack([rng.uniform(0, 100, 30), rng.uniform(0, 200, 30)])
    t.init_with_new_positions(new)
    got = t.interpolate(wdoms)
    num = np.zeros(30)
    den = np.zeros(30)
    for d in range(t.nb_subdom):
        ind = t.indices_domains[d]
        sel = t.ind_new_positions_domains[d]
        spline = _rbf(t.centers[:, ind], values[ind], t.smoothing_coef)
        w = t.norm_coefs_new_pos_domains[d]
        num[sel] += w * spline(new[:, sel].T)
        den[sel] += w
    assert np.allclose(got, num / den, atol=1e-8)


# ------------------------------------------------------------- outlier loop


def _one_domain():
    # 11 x 11 centres on [0, 100]^2 with subdom_size 200: a single subdomain
    return _grid_centers(0, 100, 0, 100, 10)


def test_outlier_loop_without_threshold_and_with_a_large_one():
    c = _one_domain()
    values = 100 + 0.5 * c[1] - 0.3 * c[0]
    t = otps.ThinPlateSplineSubdom(c, 200, 2.0, threshold=None)
    assert t.nb_subdom == 1
    _, _, s = t.compute_tps_weights_iter(c, values, 1.0)
    assert s["nb_iterations"] == 1 and s["nb_fixed_vectors"] == 0
    t.threshold = 1e9
    _, _, s = t.compute_tps_weights_iter(c, values, 1.0)
    assert s["nb_iterations"] == 1 and s["nb_fixed_vectors"] == 0


def test_smoothing_operator_is_the_subdomain_smoothing():
    """The dense operator T of the differentiable node field: T v equals
    compute_tps_weights_subdom(v)'s smoothed values (threshold None) on a
    grid of several subdomains, and reproduces an affine field exactly
    (w = 0 for affine values, whatever s)."""
    c = _grid_centers(0, 240, 0, 200, 8)
    t = otps.ThinPlateSplineSubdom(c, 150, 2.0, threshold=None)
    assert t.nb_subdom >= 4
    T = otps.smoothing_operator(c, 150, 2.0)
    rng = np.random.default_rng(12)
    for _ in range(2):
        v = rng.normal(0, 2, c.shape[1])
        sm, _, _ = t.compute_tps_weights_subdom(v)
        assert np.allclose(T @ v, sm, rtol=0, atol=1e-10 * np.abs(sm).max())
    aff = 3.0 + 0.25 * c[1] - 0.4 * c[0]
    assert np.allclose(T @ aff, aff, rtol=0, atol=1e-8)
    # a value's weight in its own smoothed value: below one for s > 0
    assert (np.diag(T) < 1).all()


def test_outlier_loop_stops_after_eleven_solves():
    """threshold 0 with noisy values and s > 0: the loop never converges and
    stops when its counter passes 10, i.e. after the first solve and ten
    re-solves (thin_plate_spline_subdom.py:260-278)."""
    c = _one_domain()
    rng = np.random.default_rng(2)
    values = rng.normal(0, 1, c.shape[1])
    t = otps.ThinPlateSplineSubdom(c, 200, 2.0, threshold=0.0)
    _, _, s = t.compute_tps_weights_iter(c, values, 5.0)
    assert s["nb_iterations"] == 11


def test_outlier_loop_replaces_a_spike_by_the_smoothed_value():
    """Affine field (reproduced exactly for any s) plus one spike of +40:
    the spike is counted as fixed and the loop pulls the field back to the
    affine one; replacing it by anything but the smoothed value (0, say)
    would leave a large outlier."""
    c = _one_domain()
    ys, xs = c
    affine = 100 + 0.5 * xs - 0.3 * ys
    values = affine.copy()
    spike = int(np.flatnonzero((xs == 50) & (ys == 50))[0])
    values[spike] += 40
    t = otps.ThinPlateSplineSubdom(c, 200, 2.0, threshold=1.5)
    smoothed, _, s = t.compute_tps_weights_iter(c, values, 20.0)
    assert s["nb_fixed_vectors"] >= 1 and 2 <= s["nb_iterations"] <= 11
    assert s["max(Udiff)"] > 1.5
    far = np.hypot(xs - 50, ys - 50) > 30
    assert np.abs(smoothed[far] - affine[far]).max() < 0.5
    # the spike's excess over the affine field drops from 40 to under 12 (the
    # loop stops once every |smoothed - value| <= 1.5, not at the fixed point)
    assert abs(smoothed[spike] - affine[spike]) < 12.0


# ---------------------------------------------------------------- WorkFIX


def _fix(dx, dy, cmax, errors=None, secondary=None, correl_min=0.2, threshold=3.0, dmax=None,
         nb_peaks=1, maps=None, correl=None, dx_in=None, dy_in=None):
    res = {"deltaxs": np.asarray(dx, np.float32).copy(),
           "deltays": np.asarray(dy, np.float32).copy(),
           "correls_max": np.asarray(cmax, np.float32).copy(),
           "errors": dict(errors or {}),
           "secondary_peaks": secondary or [None] * len(dx)}
    params = {"correl_min": correl_min, "threshold_diff_neighbour": threshold,
              "displacement_max": dmax}
    n = int(round(math.sqrt(len(dx))))
    return ofix.work_fix(res, params, n, n, nb_peaks, correl, dx_in, dy_in, maps)


def test_fix_neighbour_spike_and_corner_literals():
    """3 x 3 grid of (1, 0) with (5, 0) in the centre. Centre: neighbour
    mean 1, |diff|^2 = 16 > 3^2 -> NaN, "(diff = 4.00)". Corner (0, 0):
    neighbours 1, 1, 5 -> mean 7/3, |diff|^2 = (4/3)^2 = 1.78 -> kept."""
    dx = [1, 1, 1, 1, 5, 1, 1, 1, 1]
    r = _fix(dx, [0] * 9, [0.9] * 9)
    assert np.isnan(r["deltaxs"][4]) and np.isnan(r["deltays"][4])
    assert r["errors"] == {4: "diff neighbour too large (diff = 4.00)"}
    assert r["dxs_neighbors"][0] == pytest.approx(7 / 3, rel=1e-6)
    assert r["dxs_neighbors"][4] == pytest.approx(1.0)
    assert np.count_nonzero(np.isnan(r["deltaxs"])) == 1
    assert r["deltaxs_wrong"] == {4: np.float32(5)}


def test_fix_threshold_compares_squares():
    """|diff| = 3 exactly with threshold 3 is kept (strict >, on squares:
    9 > 9 is false); |diff| = 3.5 is rejected."""
    for centre, rejected in ((4.0, False), (4.5, True)):
        dx = [1, 1, 1, 1, centre, 1, 1, 1, 1]
        r = _fix(dx, [0] * 9, [0.9] * 9)
        assert bool(np.isnan(r["deltaxs"][4])) == rejected


def test_fix_correl_min_and_error_composition():
    """correl exactly 0.2 (float32) is kept (strict <); 0.19 is rejected; a
    vector with a PIV error that also fails correl_min gets " + " joined
    explanations and keeps its error value in deltaxs_wrong."""
    dx = [1.0] * 9
    cm = [0.9, np.float32(0.2), 0.19, 0.9, 0.9, 0.9, 0.9, 0.9, 0.1]
    errs = {8: "Correlation peak touching boundary."}
    dx[8] = 7.0
    r = _fix(dx, [0] * 9, cm, errors=errs, threshold=None)
    assert not np.isnan(r["deltaxs"][1])
    assert np.isnan(r["deltaxs"][2]) and r["errors"][2] == "correl < correl_min"
    assert r["errors"][8] == "Correlation peak touching boundary. + correl < correl_min"
    assert np.isnan(r["deltaxs"][8])


def test_fix_displacement_max_is_squared_and_strict():
    """fix.displacement_max 3: |d|^2 = 9 kept (9 > 9 false), 10.25 rejected,
    with the reference's explanation text."""
    dx = [3.0, 2.5, 0, 0, 0, 0, 0, 0, 0]
    dy = [0.0, 2.0, 0, 0, 0, 0, 0, 0, 0]
    r = _fix(dx, dy, [0.9] * 9, dmax=3.0, threshold=None)
    assert not np.isnan(r["deltaxs"][0])
    assert np.isnan(r["deltaxs"][1]) and r["errors"][1] == "delta2 < displacement_max2"


class _CorrelStub:
    """apply_subpix that returns the integer displacement (no sub-pixel
    shift), so the replacement values are exact literals."""

    def apply_subpix(self, dx, dy, correl):
        return float(dx), float(dy)


def _secondary_case(sec, dmax=None, main=(5.0, 0.0), correl0=0.9):
    """3 x 3 grid of (1, 0); the centre vector is `main` (diff to the mean
    of its neighbours: 4 px in x when main is (5, 0)) with one secondary
    peak list."""
    dx = [1, 1, 1, 1, main[0], 1, 1, 1, 1]
    dy = [0, 0, 0, 0, main[1], 0, 0, 0, 0]
    cm = [0.9] * 9
    cm[4] = correl0
    secondary = [None] * 9
    secondary[4] = sec
    return _fix(dx, dy, cm, secondary=secondary, threshold=10.0, dmax=dmax, nb_peaks=2,
                maps=[None] * 9, correl=_CorrelStub())


def test_secondary_peak_replaces_when_closer_to_the_neighbours():
    r = _secondary_case([(1.0, 0.0, np.float32(0.7))])
    assert r["deltaxs"][4] == 1.0 and r["correls_max"][4] == np.float32(0.7)
    assert r["replaced_vectors"] == {4: (np.float32(5.0), np.float32(0.0), np.float32(0.9))}


def test_secondary_peak_ratio_and_correl_min_are_strict():
    # 0.54 / 0.9 = 0.6 in float32: not > 0.6 -> no candidate
    r = _secondary_case([(1.0, 0.0, np.float32(0.54))])
    assert r["deltaxs"][4] == 5.0 and r["replaced_vectors"] == {}
    # ratio 0.61 / 0.9 > 0.6 but correl must also exceed correl_min (0.2)
    r = _secondary_case([(1.0, 0.0, np.float32(0.19))], correl0=0.3)
    assert r["deltaxs"][4] == 5.0


def test_secondary_peak_displacement_max_quirk_compares_unsquared():
    """fix.py:192 compares dx^2 + dy^2 with displacement_max itself: with
    displacement_max 3, the candidate (2, 0) (|d| = 2 < 3, but 4 > 3) is
    refused, while (1, 1) (2 < 3) is taken."""
    r = _secondary_case([(2.0, 0.0, np.float32(0.7))], dmax=3.0, main=(2.9, 0.0))
    assert r["deltaxs"][4] == np.float32(2.9) and r["replaced_vectors"] == {}
    r = _secondary_case([(1.0, 1.0, np.float32(0.7))], dmax=3.0, main=(2.9, 0.0))
    assert r["deltaxs"][4] == 1.0 and r["deltays"][4] == 1.0


def test_secondary_peak_tie_keeps_the_main_peak():
    """A candidate exactly as far from the neighbour mean as the main peak:
    argmin takes the first (the main peak), nothing is replaced."""
    r = _secondary_case([(-3.0, 0.0, np.float32(0.8))], main=(5.0, 0.0))
    # main: |5 - 1| = 4; candidate: |-3 - 1| = 4
    assert r["deltaxs"][4] == 5.0 and r["replaced_vectors"] == {}


def test_secondary_peak_of_a_nan_vector_is_not_used():
    """A vector rejected earlier (NaN) has a NaN distance to its neighbours;
    numpy's argmin returns the first NaN (index 0): no replacement."""
    dx = [1, 1, 1, 1, 5, 1, 1, 1, 1]
    cm = [0.9] * 9
    secondary = [None] * 9
    secondary[4] = [(1.0, 0.0, np.float32(0.8))]
    errors = {4: "Correlation peak touching nan."}  # NaN before the neighbour test
    r = _fix(dx, [0] * 9, cm, errors=errors, secondary=secondary, threshold=10.0,
             nb_peaks=2, maps=[None] * 9, correl=_CorrelStub())
    assert r["replaced_vectors"] == {} and np.isnan(r["deltaxs"][4])
    # the same candidate on a valid vector is taken
    r = _fix(dx, [0] * 9, cm, secondary=secondary, threshold=10.0, nb_peaks=2,
             maps=[None] * 9, correl=_CorrelStub())
    assert r["deltaxs"][4] == 1.0


# -------------------------------------------------------- secondary peaks


def test_secondary_peak_search_literals():
    """16 x 16 map, (iy0, ix0) = (7, 7), particle_radius 3, norm 2, no
    displacement_max (search box = whole map): main peak 10 at (7, 7); 9 at
    (7, 9) lies in its NaN square; then 8 at (4, 12) -> (dx, dy) =
    (7 - 12, 7 - 4) = (-5, 3), correl 4; then 6 at (12, 3) -> (4, -5),
    correl 3; then only zeros are left and the last maximum of the scan is
    (15, 15), on the border: the error ends the list (correl.py:261-266)."""
    m = np.zeros((16, 16), np.float32)
    m[7, 7], m[7, 9], m[4, 12], m[12, 3] = 10, 9, 8, 6
    c = oc.OracleCorrelFFT((16, 16), (16, 16), method_subpix="2d_gaussian2",
                           nb_peaks_to_search=4, particle_radius=3)
    dx, dy, cmax, other = c.compute_displacements_from_correl(m.copy(), norm=2.0)
    assert (dx, dy, cmax) == (0, 0, np.float32(5.0))
    assert [(a, b, float(v)) for a, b, v in other] == [(-5, 3, 4.0), (4, -5, 3.0)]


def test_secondary_peak_square_with_a_negative_start_is_empty():
    """The NaN square around a peak at row 2 starts at row -1: the slice
    [-1:6] is empty in numpy, so nothing is masked and the same secondary
    peak is found again (correl.py:255-259, the reference's quirk)."""
    m = np.zeros((16, 16), np.float32)
    m[7, 7], m[2, 12], m[12, 3] = 10, 8, 6
    c = oc.OracleCorrelFFT((16, 16), (16, 16), method_subpix="2d_gaussian2",
                           nb_peaks_to_search=4, particle_radius=3)
    _, _, _, other = c.compute_displacements_from_correl(m.copy(), norm=1.0)
    assert [(a, b, float(v)) for a, b, v in other] == [(-5, 5, 8.0)] * 3


def test_secondary_peak_search_stops_at_a_border_peak():
    """The next peak on the map border raises inside the secondary loop,
    which ends the list (correl.py:261-266)."""
    m = np.zeros((16, 16), np.float32)
    m[7, 7], m[0, 12], m[12, 3] = 10, 8, 6
    c = oc.OracleCorrelFFT((16, 16), (16, 16), method_subpix="2d_gaussian2",
                           nb_peaks_to_search=3, particle_radius=3)
    _, _, _, other = c.compute_displacements_from_correl(m.copy(), norm=1.0)
    assert other == []


# --------------------------------------------------- interpolation selector


def _field_result(work, rng):
    """A pass result on `work`'s grid: quadratic field (TPS and piecewise
    linear interpolation differ), a few NaN vectors."""
    xs = work.ixvecs_grid.astype(float)
    ys = work.iyvecs_grid.astype(float)
    dx = (0.002 * (xs - 60) ** 2 - 0.01 * ys).astype(np.float32)
    dy = (0.001 * xs * ys / 50).astype(np.float32)
    bad = rng.choice(dx.size, 5, replace=False)
    dx[bad] = np.nan
    return {"deltaxs": dx, "deltays": dy, "xs": xs, "ys": ys}, ~np.isnan(dx)


def _passes(use_tps):
    p = default_params()
    p["piv0"]["shape_crop_im0"] = 32
    p["multipass"]["number"] = 2
    p["multipass"]["use_tps"] = use_tps
    p["multipass"]["subdom_size"] = 1000  # one subdomain: compare with one RBF
    p["multipass"]["smoothing_coef"] = 0.5
    p["multipass"]["threshold_tps"] = None
    w = OracleWorkPIV(p)
    for work in w.works_piv:
        work.prepare_with_image((160, 192))
    return w.works_piv


def _expected_griddata(res, sel, work):
    gx, gy = np.meshgrid(work.ixvecs, work.iyvecs)
    pts = (res["xs"][sel], res["ys"][sel])
    out = []
    for v in (res["deltaxs"][sel], res["deltays"][sel]):
        lin = griddata(pts, v, (gx, gy), "linear")
        near = griddata(pts, v, (gx, gy), "nearest")
        lin[np.isnan(lin)] = near[np.isnan(lin)]
        out.append(lin.ravel())
    return out


def _expected_tps(res, sel, work, count):
    xs, ys = res["xs"][sel], res["ys"][sel]
    # one subdomain: smoothing 0.5 sqrt(range_x range_y / count)
    s = 0.5 * math.sqrt((xs.max() - xs.min()) * (ys.max() - ys.min()) / count)
    new = np.vstack([work.ixvecs_grid, work.iyvecs_grid]).T.astype(float)
    out = []
    for v in (res["deltaxs"][sel], res["deltays"][sel]):
        spline = RBFInterpolator(np.vstack([xs, ys]).T, v.astype(float),
                                 kernel="thin_plate_spline", smoothing=s, degree=1)
        out.append((spline(new), spline(np.vstack([xs, ys]).T)))
    return out


@pytest.mark.parametrize("use_tps, last, kind", [
    (True, False, "tps"), ("last", False, "griddata"), ("last", True, "tps"),
    (False, True, "griddata"), (False, False, "griddata")])
def test_apply_interp_selector(use_tps, last, kind):
    rng = np.random.default_rng(4)
    w0, w1 = _passes(use_tps)
    res, sel = _field_result(w0, rng)
    target = w1 if not last else w0
    target.apply_interp(res, last=last)
    key = "final" if last else "approx"
    got_x, got_y = res[f"deltaxs_{key}"], res[f"deltays_{key}"]
    assert len(got_x) == len(target.ixvecs_grid)
    if kind == "griddata":
        ex, ey = _expected_griddata(res, sel, target)
        assert np.allclose(got_x, ex, atol=1e-6) and np.allclose(got_y, ey, atol=1e-6)
        assert "tps" not in res
    else:
        (ex, sx), (ey, sy) = _expected_tps(res, sel, target, int(sel.sum()))
        assert res["tps"].nb_subdom == 1
        assert np.allclose(got_x, ex, atol=1e-7) and np.allclose(got_y, ey, atol=1e-7)
        if last:
            # smoothed values at the selected vectors, the interpolation elsewhere
            assert np.allclose(res["deltaxs_smooth"][sel], sx, atol=1e-7)
            assert np.allclose(res["deltaxs_smooth"][~sel], ex[~sel], atol=1e-7)
            assert res["xs_smooth"].dtype == np.float32


def test_multipass_window_schedule():
    """Pass windows: previous / coeff_zoom by true division then int
    (multipass.py:153-170, singlepass.py:86-87): 64 -> 32 -> 16 with zoom 2,
    48 -> 24 -> 12, and 64 / 3 = 21.33 -> 21 with zoom [3, 2] -> 10."""
    for s0, zoom, number, expected in ((64, 2, 3, [64, 32, 16]), (48, 2, 3, [48, 24, 12]),
                                       (64, [3, 2], 3, [64, 21, 10])):
        p = default_params()
        p["piv0"]["shape_crop_im0"] = s0
        p["multipass"]["number"] = number
        p["multipass"]["coeff_zoom"] = zoom
        w = OracleWorkPIV(p)
        assert [wk.shape_crop_im0[0] for wk in w.works_piv] == expected
    p = default_params()
    p["multipass"]["number"] = 3
    p["multipass"]["coeff_zoom"] = [2]
    with pytest.raises(ValueError, match="coeff_zoom"):
        OracleWorkPIV(p)


def test_pass_class_for_later_passes():
    p = default_params()
    w = OracleWorkPIVPass(p, 1, (21.333333333333332, 21.333333333333332))
    assert w.shape_crop_im0 == (21, 21)

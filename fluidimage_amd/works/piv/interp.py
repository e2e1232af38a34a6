"""Interpolation selector between passes (src/fluidimage/works/piv/singlepass.py:409-563).

TPS subdomains run on the GPU in FP64 (``calcul.tps.ThinPlateSplineSubdomHIP``);
``griddata`` is scipy's Delaunay-linear interpolation with nearest fill
(src/fluidimage/calcul/interpolate/griddata.py:11-27), exactly the reference's
host call, used when ``params.multipass.use_tps`` selects it or when a TPS
system is singular.
"""

import numpy as np
from scipy import interpolate


class InterpError(ValueError):
    pass


def griddata(centers, values, new_positions):
    """src/fluidimage/calcul/interpolate/griddata.py:11-27"""
    grid_x, grid_y = np.meshgrid(new_positions[0], new_positions[1])
    values_new = interpolate.griddata(centers.T, values, (grid_x, grid_y), "linear")
    inds = np.where(np.isnan(values_new))
    nearest = interpolate.griddata(centers.T, values, (grid_x, grid_y), "nearest")
    values_new[inds] = nearest[inds]
    return values_new.flatten()


def apply_interp(work, piv_results, last=False):
    if not last and not hasattr(piv_results, "ixvecs_approx"):
        (piv_results.ixvecs_approx,
         piv_results.iyvecs_approx) = work._xyoriginalimage_from_xymasked(
            work.ixvecs_grid, work.iyvecs_grid)
    if last and not hasattr(piv_results, "ixvecs_final"):
        (piv_results.ixvecs_final,
         piv_results.iyvecs_final) = work._xyoriginalimage_from_xymasked(
            work.ixvecs_grid, work.iyvecs_grid)

    selection = ~(np.isnan(piv_results.deltaxs) | np.isnan(piv_results.deltays))
    if not any(selection):
        raise InterpError("Only nan.")
    xs = piv_results.xs[selection]
    ys = piv_results.ys[selection]
    xs_tmp, ys_tmp = work._xymasked_from_xyoriginalimage(xs, ys)
    centers = np.vstack([ys_tmp, xs_tmp])
    deltaxs = piv_results.deltaxs[selection]
    deltays = piv_results.deltays[selection]

    mp = work.params.multipass
    done = False
    if mp.use_tps is True or (mp.use_tps == "last" and last):
        from ...calcul.tps import ThinPlateSplineSubdomHIP, SingularSystemError

        tps = ThinPlateSplineSubdomHIP(centers, mp.subdom_size, mp.smoothing_coef,
                                       threshold=mp.threshold_tps, percent_buffer_area=20)
        try:
            new_positions = np.vstack([work.iyvecs_grid, work.ixvecs_grid])
            out = tps.smooth_and_interpolate(deltaxs, deltays, new_positions)
        except SingularSystemError:
            print("LinAlgError in TPS: compute delta_approx with griddata")
        else:
            done = True
            (dxs_smooth, dys_smooth, approx_x, approx_y, summary) = out
            piv_results.deltaxs_tps = tps.weights_x
            piv_results.deltays_tps = tps.weights_y
            if last:
                xs_smooth = piv_results.ixvecs_final.copy().astype(np.float32)
                xs_smooth[selection] = xs
                ys_smooth = piv_results.iyvecs_final.copy().astype(np.float32)
                ys_smooth[selection] = ys
                full_x = approx_x.copy()
                full_x[selection] = dxs_smooth
                full_y = approx_y.copy()
                full_y[selection] = dys_smooth
                piv_results.xs_smooth = xs_smooth
                piv_results.ys_smooth = ys_smooth
                piv_results.deltaxs_smooth = full_x
                piv_results.deltays_smooth = full_y
            piv_results.tps_summary = summary
    if not done:
        approx_x = griddata(centers[::-1], deltaxs, (work.ixvecs, work.iyvecs))
        approx_y = griddata(centers[::-1], deltays, (work.ixvecs, work.iyvecs))
    if last:
        piv_results.deltaxs_final = approx_x
        piv_results.deltays_final = approx_y
    else:
        piv_results.deltaxs_approx = approx_x
        piv_results.deltays_approx = approx_y

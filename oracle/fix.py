"""Oracle: outlier detection WorkFIX and the 8-neighbour mean (test infrastructure).

Restates src/fluidimage/works/piv/fix.py:66-243 and
src/fluidimage/calcul/mean_neighbors.py:5-155.
"""

from math import sqrt

import numpy as np

from .subpix import SubPixError


def mean_neighbors(arr, arr_is_not_nan):
    """src/fluidimage/calcul/mean_neighbors.py:5-125.

    ``arr`` holds zeros where NaN. The reference treats corners, edges and
    the core separately; each branch sums the valid values among the (up
    to 8) neighbours inside the grid, so one padded sum restates all
    branches. Sums are float32 like the reference arrays.
    """
    n0, n1 = arr.shape
    a = np.zeros((n0 + 2, n1 + 2), dtype=arr.dtype)
    b = np.zeros((n0 + 2, n1 + 2), dtype=np.uint8)
    a[1:-1, 1:-1] = arr
    b[1:-1, 1:-1] = arr_is_not_nan
    tot = np.zeros_like(arr)
    cnt = np.zeros(arr.shape, dtype=np.int32)
    for d0 in (-1, 0, 1):
        for d1 in (-1, 0, 1):
            if d0 == 0 and d1 == 0:
                continue
            tot = tot + a[1 + d0 : 1 + d0 + n0, 1 + d1 : 1 + d1 + n1]
            cnt = cnt + b[1 + d0 : 1 + d0 + n0, 1 + d1 : 1 + d1 + n1]
    out = np.full(arr.shape, np.nan, dtype=arr.dtype)
    good = arr_is_not_nan & (cnt > 0)
    out[good] = tot[good] / cnt[good].astype(arr.dtype)
    return out


def mean_neighbors_xy(deltaxs, deltays, ny, nx):
    """src/fluidimage/calcul/mean_neighbors.py:128-155 on a (ny, nx) grid."""
    shape = (ny, nx)
    dx2 = np.asarray(deltaxs, np.float32).copy().reshape(shape)
    dy2 = np.asarray(deltays, np.float32).copy().reshape(shape)
    bad = np.isnan(dx2) | np.isnan(dy2)
    dx2[bad] = 0.0
    dy2[bad] = 0.0
    ok = ~bad
    return (
        mean_neighbors(dx2, ok).reshape(ny * nx),
        mean_neighbors(dy2, ok).reshape(ny * nx),
    )


def work_fix(result, params_fix, ny, nx, nb_peaks_to_search, correl, deltaxs_input=None,
             deltays_input=None, maps=None):
    """WorkFIX.calcul: src/fluidimage/works/piv/fix.py:66-243.

    ``result`` is a dict with ``deltaxs``, ``deltays`` (float32, modified in
    place), ``correls_max``, ``errors`` (dict ivec -> str) and
    ``secondary_peaks``. ``correl`` is the oracle correlation object (for
    ``apply_subpix`` on secondary peaks), ``maps`` the per-window maps.
    """
    deltaxs_wrong = {}
    deltays_wrong = {}
    deltaxs = result["deltaxs"]
    deltays = result["deltays"]
    errors = result["errors"]
    for ierr in errors.keys():
        deltaxs_wrong[ierr] = deltaxs[ierr]
        deltays_wrong[ierr] = deltays[ierr]
        deltaxs[ierr] = np.nan
        deltays[ierr] = np.nan

    def put_to_nan(inds, explanation):
        for ind in inds:
            ind = int(ind)
            if ind not in deltaxs_wrong:
                deltaxs_wrong[ind] = deltaxs[ind]
                deltays_wrong[ind] = deltays[ind]
                deltaxs[ind] = np.nan
                deltays[ind] = np.nan
                errors[ind] = explanation
            else:
                errors[ind] += " + " + explanation

    correls_max = result["correls_max"]
    with np.errstate(invalid="ignore"):
        inds = (correls_max < params_fix["correl_min"]).nonzero()[0]
    put_to_nan(inds, "correl < correl_min")

    dmax = params_fix.get("displacement_max")
    if dmax:
        displacement_max2 = dmax**2
        delta2s = deltaxs**2 + deltays**2
        with np.errstate(invalid="ignore"):
            inds = (delta2s > displacement_max2).nonzero()[0]
        put_to_nan(inds, "delta2 < displacement_max2")

    threshold = params_fix.get("threshold_diff_neighbour")
    dxs_neighbors = dys_neighbors = differences2 = None
    if threshold is not None:
        dxs_neighbors, dys_neighbors = mean_neighbors_xy(deltaxs, deltays, ny, nx)
        diff_x = dxs_neighbors - deltaxs
        diff_y = dys_neighbors - deltays
        differences2 = diff_x**2 + diff_y**2
        with np.errstate(invalid="ignore"):
            inds = (differences2 > threshold**2).nonzero()[0]
        put_to_nan(inds, "diff neighbour too large")
        for ivec in inds:
            errors[ivec] += f" (diff = {sqrt(differences2[ivec]):.2f})"
        result["deltaxs_wrong"] = deltaxs_wrong
        result["deltays_wrong"] = deltays_wrong

    result["dxs_neighbors"] = dxs_neighbors
    result["dys_neighbors"] = dys_neighbors
    if nb_peaks_to_search < 2:
        return result

    ratio_correl_peaks = 0.6
    replaced = {}
    for ivec, other_peaks in enumerate(result["secondary_peaks"]):
        if other_peaks is None or len(other_peaks) == 0:
            continue
        correl0 = correls_max[ivec]
        if deltaxs_input is not None:
            dx_input = deltaxs_input[ivec]
            dy_input = deltays_input[ivec]
        else:
            dx_input = 0
            dy_input = 0
        good = []
        for dx, dy, corr in other_peaks:
            if corr / correl0 > ratio_correl_peaks and corr > params_fix["correl_min"]:
                good.append((dx + dx_input, dy + dy_input, corr))
        if not good:
            continue
        diff_neighbours = np.empty(len(good) + 1)
        diff_neighbours[0] = sqrt(float(differences2[ivec]))  # NaN stays NaN
        for i, (dx, dy, corr) in enumerate(good):
            diff_neighbours[i + 1] = np.sqrt(
                (dxs_neighbors[ivec] - dx) ** 2 + (dys_neighbors[ivec] - dy) ** 2
            )
        argmin = diff_neighbours.argmin()
        if argmin == 0:
            continue
        dx, dy, corr = good[argmin - 1]
        if threshold is not None and diff_neighbours[argmin] > threshold:
            continue
        if dmax and dx**2 + dy**2 > dmax:
            continue
        dx -= dx_input
        dy -= dy_input
        try:
            dx, dy = correl.apply_subpix(dx, dy, maps[ivec])
        except SubPixError:
            continue
        dx += dx_input
        dy += dy_input
        old = (deltaxs_wrong.get(ivec, deltaxs[ivec]), deltays_wrong.get(ivec, deltays[ivec]))
        replaced[ivec] = (old[0], old[1], correls_max[ivec])
        deltaxs[ivec] = dx
        deltays[ivec] = dy
        correls_max[ivec] = corr
        if ivec in errors:
            del deltaxs_wrong[ivec], deltays_wrong[ivec], errors[ivec]
    result["replaced_vectors"] = replaced
    return result

"""Oracle: single-pass and multipass PIV works (test infrastructure).

Restates src/fluidimage/works/piv/singlepass.py:52-870 (grid, padding,
window positions, crop, ``_loop_vectors``, ``apply_interp``),
src/fluidimage/works/piv/multipass.py:107-190 and
src/fluidimage/works/with_mask.py:15-32. Parameters are plain nested dicts
with the reference's names and defaults (``default_params``).
"""

import copy

import numpy as np
from scipy import interpolate

from .correl import EXPLANATIONS, ERR_BAD_CROP, OracleCorrelFFT, PIVError
from .fix import work_fix
from .subpix import SubPixError
from .tps import ThinPlateSplineSubdom


def default_params():
    """Defaults of singlepass.py:583-596,651-653, fix.py:32-39,
    multipass.py:58-68, with_mask.py:36."""
    return {
        "piv0": {
            "shape_crop_im0": 48,
            "shape_crop_im1": None,
            "displacement_max": None,
            "method_correl": "fftw",
            "method_subpix": "2d_gaussian2",
            "nsubpix": None,
            "nb_peaks_to_search": 1,
            "particle_radius": 3,
            "grid": {"overlap": 0.5, "from": "overlap"},
        },
        "fix": {"correl_min": 0.2, "threshold_diff_neighbour": 10, "displacement_max": None},
        "multipass": {
            "number": 1,
            "coeff_zoom": 2,
            "use_tps": "last",
            "subdom_size": 200,
            "smoothing_coef": 5.0,
            "threshold_tps": 1.5,
            # extension (no reference counterpart): None or "bicubic"
            "image_deformation": None,
        },
        "mask": {"strcrop": None},
    }


def slices_from_strcrop(strcrop):
    """src/fluidimage/data_objects/piv.py:71-75"""
    return tuple(
        slice(*(int(i) if i else None for i in part.strip().split(":")))
        for part in strcrop.split(",")
    )


def griddata(centers, values, new_positions):
    """src/fluidimage/calcul/interpolate/griddata.py:11-27 (scipy Qhull)."""
    grid_x, grid_y = np.meshgrid(new_positions[0], new_positions[1])
    values_new = interpolate.griddata(centers.T, values, (grid_x, grid_y), "linear")
    inds = np.where(np.isnan(values_new))
    nearest = interpolate.griddata(centers.T, values, (grid_x, grid_y), "nearest")
    values_new[inds] = nearest[inds]
    return values_new.flatten()


def _as_shape(shape):
    if isinstance(shape, (int, np.integer)):
        return (shape, shape)
    if isinstance(shape, (tuple, list)) and len(shape) == 2:
        return tuple(shape)
    raise NotImplementedError("For now, shape_crop_im0 has to be one or two integer!")


class OracleWorkPIVPass:
    """One pass: FirstWorkPIV (index_pass 0) or WorkPIVFromDisplacement."""

    def __init__(self, params, index_pass=0, shape_crop_im0=None, shape_crop_im1=None,
                 precision="f64"):
        self.params = params
        self.index_pass = index_pass
        self.overlap = params["piv0"]["grid"]["overlap"]
        if self.overlap >= 1:
            raise ValueError("params.piv0.grid.overlap has to be smaller than 1")
        if shape_crop_im0 is None:
            shape_crop_im0 = params["piv0"]["shape_crop_im0"]
        if shape_crop_im1 is None:
            shape_crop_im1 = params["piv0"]["shape_crop_im1"]
        # singlepass.py:52-87
        if shape_crop_im1 is None:
            shape_crop_im1 = shape_crop_im0
        s0 = _as_shape(shape_crop_im0)
        s1 = _as_shape(shape_crop_im1)
        if s1[0] > s0[0] or s1[1] > s0[1]:
            raise ValueError("shape_crop_im1 must be inferior or equal to shape_crop_im0")
        self.shape_crop_im0 = tuple(int(n) for n in s0)
        self.shape_crop_im1 = tuple(int(n) for n in s1)
        # singlepass.py:363-389
        n0, n1 = self.shape_crop_im0
        self._start_for_crop0 = (n0 // 2, n1 // 2)
        self._stop_for_crop0 = (n0 // 2 + n0 % 2, n1 // 2 + n1 % 2)
        m0, m1 = self.shape_crop_im1
        self._start_for_crop1 = (m0 // 2, m1 // 2)
        self._stop_for_crop1 = (m0 // 2 + m0 % 2, m1 // 2 + m1 % 2)
        p = params["piv0"]
        self.correl = OracleCorrelFFT(
            self.shape_crop_im0,
            self.shape_crop_im1,
            method_subpix=p["method_subpix"],
            nsubpix=p["nsubpix"],
            displacement_max=p["displacement_max"],
            particle_radius=p["particle_radius"],
            nb_peaks_to_search=p["nb_peaks_to_search"],
            precision=precision,
        )

    # -- masks (with_mask.py:15-32)
    def _shift(self, xs, ys, sign):
        strcrop = self.params["mask"]["strcrop"]
        if strcrop is not None:
            slices = slices_from_strcrop(strcrop)
            if slices[1].start is not None:
                xs = xs + sign * slices[1].start
            if slices[0].start is not None:
                ys = ys + sign * slices[0].start
        return xs, ys

    def xyoriginal_from_xymasked(self, xs, ys):
        return self._shift(xs, ys, +1)

    def xymasked_from_xyoriginal(self, xs, ys):
        return self._shift(xs, ys, -1)

    def prepare_with_image(self, imshape):
        """singlepass.py:108-141"""
        self.imshape0 = (len_y, len_x) = tuple(imshape)
        scim = self.shape_crop_im0
        stepy = scim[0] - int(np.round(self.overlap * scim[0]))
        stepx = scim[1] - int(np.round(self.overlap * scim[1]))
        assert stepy >= 1 and stepx >= 1
        ixvec_max = len_x - self._stop_for_crop0[1]
        ixvecs = np.arange(self._start_for_crop0[1], ixvec_max, stepx, dtype=int)
        iyvec_max = len_y - self._stop_for_crop0[0]
        iyvecs = np.arange(self._start_for_crop0[0], iyvec_max, stepy, dtype=int)
        if ixvec_max - ixvecs[-1] > stepx // 2.5:
            ixvecs = np.append(ixvecs, ixvec_max)
        if iyvec_max - iyvecs[-1] > stepy // 2.5:
            iyvecs = np.append(iyvecs, iyvec_max)
        self.ixvecs = ixvecs
        self.iyvecs = iyvecs
        gx, gy = np.meshgrid(ixvecs, iyvecs)
        self.ixvecs_grid = gx.flatten()
        self.iyvecs_grid = gy.flatten()

    def pad_images(self, im0, im1):
        """singlepass.py:196-208"""
        npad = self.npad = max(self._start_for_crop0 + self._stop_for_crop0)
        tmp = [(npad, npad), (npad, npad)]
        im0pad = np.pad(im0 - im0.min(), tmp, "constant", constant_values=0)
        im1pad = np.pad(im1 - im1.min(), tmp, "constant", constant_values=0)
        return im0pad, im1pad

    def positions(self, deltaxs_input=None, deltays_input=None):
        """singlepass.py:210-260 (first pass) and :787-870 (later passes)."""
        if self.index_pass == 0 or deltaxs_input is None:
            ixs0 = self.ixvecs_grid
            iys0 = self.iyvecs_grid
            return (ixs0, iys0, ixs0 + self.npad, iys0 + self.npad,
                    ixs0 + self.npad, iys0 + self.npad)
        ixs0 = self.ixvecs_grid - deltaxs_input // 2
        iys0 = self.iyvecs_grid - deltays_input // 2
        ixs1 = ixs0 + deltaxs_input
        iys1 = iys0 + deltays_input
        H, W = self.imshape0
        out0 = (ixs0 > W) | (ixs0 < 0) | (iys0 > H) | (iys0 < 0)
        out1 = (ixs1 > W) | (ixs1 < 0) | (iys1 > H) | (iys1 < 0)
        case_a = out1
        case_b = out0 & ~out1
        ixs0 = np.where(case_a, self.ixvecs_grid - deltaxs_input, ixs0)
        iys0 = np.where(case_a, self.iyvecs_grid - deltays_input, iys0)
        ixs1 = np.where(case_a, self.ixvecs_grid, ixs1)
        iys1 = np.where(case_a, self.iyvecs_grid, iys1)
        ixs0 = np.where(case_b, self.ixvecs_grid, ixs0)
        iys0 = np.where(case_b, self.iyvecs_grid, iys0)
        ixs1 = np.where(case_b, self.ixvecs_grid + deltaxs_input, ixs1)
        iys1 = np.where(case_b, self.iyvecs_grid + deltays_input, iys1)
        xs = (ixs0 + ixs1) / 2.0
        ys = (iys0 + iys1) / 2.0
        return (xs, ys, ixs0 + self.npad, iys0 + self.npad,
                ixs1 + self.npad, iys1 + self.npad)

    def crop_valid(self, ixs_pad, iys_pad, shape_crop, start, stop, padshape):
        """True where the slice of singlepass.py:391-407 has the full shape."""
        top = iys_pad - start[0]
        left = ixs_pad - start[1]
        return (top >= 0) & (iys_pad + stop[0] <= padshape[0]) & (left >= 0) & (
            ixs_pad + stop[1] <= padshape[1]
        )

    def loop_vectors(self, im0, im1, deltaxs_input=None, deltays_input=None, keep_maps=True,
                     deform=False):
        """singlepass.py:262-361 with the FFTs batched over windows.

        ``deform``: the later-pass windows are the bicubic deformed crops of
        oracle/deform.py around the grid nodes, with the float predictor
        (deltaxs_input, deltays_input) on the grid; the vector is the
        residual plus the predictor (no reference counterpart)."""
        im0pad, im1pad = self.pad_images(im0, im1)
        if deform and deltaxs_input is not None:
            xs = self.ixvecs_grid.astype(np.float64)
            ys = self.iyvecs_grid.astype(np.float64)
        else:
            deform = False
            xs, ys, ixs0, iys0, ixs1, iys1 = self.positions(deltaxs_input, deltays_input)
        nb_vec = len(xs)
        n0, n1 = self.shape_crop_im0
        m0, m1 = self.shape_crop_im1
        if deform:
            if (m0, m1) != (n0, n1):
                raise ValueError("image deformation needs shape_crop_im1 == shape_crop_im0")
            valid = np.ones(nb_vec, dtype=bool)
        else:
            valid = self.crop_valid(ixs0, iys0, self.shape_crop_im0, self._start_for_crop0,
                                    self._stop_for_crop0, im0pad.shape) & self.crop_valid(
                ixs1, iys1, self.shape_crop_im1, self._start_for_crop1, self._stop_for_crop1,
                im1pad.shape)
        deltaxs = np.empty(nb_vec, dtype=np.float32)
        deltays = np.empty_like(deltaxs)
        correls_max = np.empty_like(deltaxs)
        errors = {}
        codes = np.zeros(nb_vec, np.int32)
        secondary_peaks = [None] * nb_vec
        maps = [None] * nb_vec
        norms = np.full(nb_vec, np.nan)
        has_to_apply_subpix = self.index_pass == self.params["multipass"]["number"] - 1
        iv = np.flatnonzero(valid)
        if iv.size and deform:
            from .deform import deformed_crops

            c0, c1 = deformed_crops(im0, im1, self.iyvecs, self.ixvecs, deltaxs_input,
                                    deltays_input, (n0, n1))
            cmaps, cnorms = self.correl(c0.astype(np.float32), c1.astype(np.float32))
        elif iv.size:
            r0 = np.arange(n0)[None, :, None]
            c0 = np.arange(n1)[None, None, :]
            r1 = np.arange(m0)[None, :, None]
            c1 = np.arange(m1)[None, None, :]
            crops0 = im0pad[(iys0[iv] - self._start_for_crop0[0])[:, None, None] + r0,
                            (ixs0[iv] - self._start_for_crop0[1])[:, None, None] + c0]
            crops1 = im1pad[(iys1[iv] - self._start_for_crop1[0])[:, None, None] + r1,
                            (ixs1[iv] - self._start_for_crop1[1])[:, None, None] + c1]
            cmaps, cnorms = self.correl(crops0.astype(np.float32), crops1.astype(np.float32))
        for k, ivec in enumerate(iv):
            correl = np.array(cmaps[k])
            norm = float(cnorms[k])
            maps[ivec] = correl
            norms[ivec] = norm
            try:
                dx, dy, cmax, other = self.correl.compute_displacements_from_correl(
                    correl, norm=norm)
            except PIVError as e:
                errors[ivec] = e.explanation
                codes[ivec] = e.code
                deltaxs[ivec], deltays[ivec], correls_max[ivec] = e.results
                continue
            if has_to_apply_subpix:
                try:
                    dx, dy = self.correl.apply_subpix(dx, dy, correl)
                except SubPixError as e:
                    errors[ivec] = e.explanation
                    codes[ivec] = e.code
            deltaxs[ivec] = dx
            deltays[ivec] = dy
            correls_max[ivec] = cmax
            secondary_peaks[ivec] = other
        for ivec in np.flatnonzero(~valid):
            deltaxs[ivec] = deltays[ivec] = correls_max[ivec] = np.nan
            errors[ivec] = EXPLANATIONS[ERR_BAD_CROP]
            codes[ivec] = ERR_BAD_CROP
        if deform:
            deltaxs[:] = (deltaxs.astype(np.float64) + deltaxs_input).astype(np.float32)
            deltays[:] = (deltays.astype(np.float64) + deltays_input).astype(np.float32)
        elif deltaxs_input is not None:
            deltaxs += deltaxs_input
            deltays += deltays_input
        xs, ys = self.xyoriginal_from_xymasked(xs, ys)
        return {
            "deltaxs": deltaxs,
            "deltays": deltays,
            "xs": xs,
            "ys": ys,
            "correls_max": correls_max,
            "correls": maps if keep_maps else None,
            "norms": norms,
            "errors": errors,
            "codes": codes,
            "secondary_peaks": secondary_peaks,
            "deltaxs_input": deltaxs_input,
            "deltays_input": deltays_input,
        }

    def apply_interp(self, res, last=False):
        """singlepass.py:409-563"""
        ixg, iyg = self.xyoriginal_from_xymasked(self.ixvecs_grid, self.iyvecs_grid)
        if not last:
            res.setdefault("ixvecs_approx", ixg)
            res.setdefault("iyvecs_approx", iyg)
        else:
            res.setdefault("ixvecs_final", ixg)
            res.setdefault("iyvecs_final", iyg)
        selection = ~(np.isnan(res["deltaxs"]) | np.isnan(res["deltays"]))
        if not any(selection):
            raise ValueError("Only nan.")
        xs = res["xs"][selection]
        ys = res["ys"][selection]
        xs_tmp, ys_tmp = self.xymasked_from_xyoriginal(xs, ys)
        centers = np.vstack([ys_tmp, xs_tmp])
        deltaxs = res["deltaxs"][selection]
        deltays = res["deltays"][selection]
        mp = self.params["multipass"]
        use_tps = mp["use_tps"] is True or (mp["use_tps"] == "last" and last)
        done = False
        if use_tps:
            tps = ThinPlateSplineSubdom(centers, mp["subdom_size"], mp["smoothing_coef"],
                                        threshold=mp["threshold_tps"], percent_buffer_area=20)
            try:
                dxs_smooth, dxs_tps, summary = tps.compute_tps_weights_subdom(deltaxs)
                dys_smooth, dys_tps, summary = tps.compute_tps_weights_subdom(deltays)
            except np.linalg.LinAlgError:
                pass
            else:
                done = True
                res["deltaxs_tps"] = dxs_tps
                res["deltays_tps"] = dys_tps
                res["tps_summary"] = summary
                tps.init_with_new_positions(np.vstack([self.iyvecs_grid, self.ixvecs_grid]))
                approx_x = tps.interpolate(dxs_tps)
                approx_y = tps.interpolate(dys_tps)
                res["tps"] = tps
                if last:
                    xs_smooth = res["ixvecs_final"].copy().astype(np.float32)
                    xs_smooth[selection] = xs
                    ys_smooth = res["iyvecs_final"].copy().astype(np.float32)
                    ys_smooth[selection] = ys
                    full_x = approx_x.copy()
                    full_x[selection] = dxs_smooth
                    full_y = approx_y.copy()
                    full_y[selection] = dys_smooth
                    res["xs_smooth"] = xs_smooth
                    res["ys_smooth"] = ys_smooth
                    res["deltaxs_smooth"] = full_x
                    res["deltays_smooth"] = full_y
        if not done:
            approx_x = griddata(centers[::-1], deltaxs, (self.ixvecs, self.iyvecs))
            approx_y = griddata(centers[::-1], deltays, (self.ixvecs, self.iyvecs))
        if last:
            res["deltaxs_final"] = approx_x
            res["deltays_final"] = approx_y
        else:
            res["deltaxs_approx"] = approx_x
            res["deltays_approx"] = approx_y


def apply_strcrop(im, strcrop):
    if strcrop is None:
        return im
    return im[slices_from_strcrop(strcrop)]


class OracleWorkPIV:
    """WorkPIV: src/fluidimage/works/piv/multipass.py:107-190."""

    def __init__(self, params=None, precision="f64"):
        if params is None:
            params = default_params()
        self.params = params
        first = OracleWorkPIVPass(params, 0, precision=precision)
        self.works_piv = [first]
        number = params["multipass"]["number"]
        coeff_zoom = params["multipass"]["coeff_zoom"]
        if isinstance(coeff_zoom, int):
            coeffs_zoom = [coeff_zoom] * (number - 1)
        elif len(coeff_zoom) == number - 1:
            coeffs_zoom = coeff_zoom
        else:
            raise ValueError("params.multipass.coeff_zoom has to be an integer or "
                             "an iterable of length params.multipass.number - 1")
        s0 = params["piv0"]["shape_crop_im0"]
        s1 = params["piv0"]["shape_crop_im1"]
        if s1 is None:
            s1 = s0
        s0 = _as_shape(copy.copy(s0))
        s1 = _as_shape(copy.copy(s1))
        for i in range(1, number):
            s0 = (s0[0] / coeffs_zoom[i - 1], s0[1] / coeffs_zoom[i - 1])
            s1 = (s1[0] / coeffs_zoom[i - 1], s1[1] / coeffs_zoom[i - 1])
            self.works_piv.append(OracleWorkPIVPass(params, i, s0, s1, precision=precision))

    def calcul(self, im0, im1):
        strcrop = self.params["mask"]["strcrop"]
        im0 = apply_strcrop(np.asarray(im0), strcrop)
        im1 = apply_strcrop(np.asarray(im1), strcrop)
        results = []
        res = None
        p = self.params["piv0"]
        for work in self.works_piv:
            if not hasattr(work, "ixvecs_grid"):
                work.prepare_with_image(im0.shape)
            if res is None:
                res = work.loop_vectors(im0, im1)
            else:
                work.apply_interp(res)
                if self.params["multipass"].get("image_deformation"):
                    res = work.loop_vectors(im0, im1, np.asarray(res["deltaxs_approx"], float),
                                            np.asarray(res["deltays_approx"], float),
                                            deform=True)
                else:
                    dxin = np.round(res["deltaxs_approx"]).astype("int32")
                    dyin = np.round(res["deltays_approx"]).astype("int32")
                    res = work.loop_vectors(im0, im1, dxin, dyin)
            ny, nx = len(work.iyvecs), len(work.ixvecs)
            res = work_fix(res, self.params["fix"], ny, nx, p["nb_peaks_to_search"],
                           work.correl, res["deltaxs_input"], res["deltays_input"],
                           res["correls"])
            results.append(res)
        try:
            work.apply_interp(res, last=True)
        except ValueError:
            pass
        return results

"""Multipass PIV (src/fluidimage/works/piv/multipass.py:18-198) on the GPU."""

import time
from copy import copy

from ...data_objects import MultipassPIVResults
from ...params import Params
from ..with_mask import BaseWorkWithMask
from .fix import WorkFIX
from .interp import InterpError
from .singlepass import FirstWorkPIV, WorkPIVFromDisplacement


class WorkPIV:
    """Main work for PIV with multipass (same parameters as the reference)."""

    @classmethod
    def create_default_params(cls):
        params = Params(tag="params")
        cls._complete_params_with_default(params)
        return params

    @classmethod
    def _complete_params_with_default(cls, params):
        params._set_child("series", attribs={"path": "", "str_subset": None, "ind_start": 0,
                                             "ind_stop": None, "ind_step": 1})
        FirstWorkPIV._complete_params_with_default(params)
        WorkFIX._complete_params_with_default(params)
        params._set_child(
            "multipass",
            attribs={
                "number": 1,
                "coeff_zoom": 2,
                "use_tps": "last",
                "subdom_size": 200,
                "smoothing_coef": 5.0,
                "threshold_tps": 1.5,
                # extensions (no counterpart in the reference): image
                # deformation between passes (None or "bicubic") and gradient
                # refinement steps after the last pass (0: none)
                "image_deformation": None,
                "refine_steps": 0,
            },
        )

    def __init__(self, params=None):
        if params is None:
            params = self.create_default_params()
        self.params = params
        work_piv = FirstWorkPIV(params)
        self.works_piv = [work_piv]
        self.works_fix = [WorkFIX(params.fix, work_piv)]
        coeff_zoom = params.multipass.coeff_zoom
        if isinstance(coeff_zoom, int):
            coeffs_zoom = [coeff_zoom] * (params.multipass.number - 1)
        elif len(coeff_zoom) == params.multipass.number - 1:
            coeffs_zoom = coeff_zoom
        else:
            raise ValueError(
                "params.multipass.coeff_zoom has to be an integer or "
                "an iterable of length params.multipass.number - 1"
            )
        shape_crop_im0 = copy(params.piv0.shape_crop_im0)
        shape_crop_im1 = copy(params.piv0.shape_crop_im1)
        if shape_crop_im1 is None:
            shape_crop_im1 = shape_crop_im0
        if isinstance(shape_crop_im0, int):
            shape_crop_im0 = (shape_crop_im0, shape_crop_im0)
        elif isinstance(shape_crop_im0, (list, tuple)) and len(shape_crop_im0) == 2:
            shape_crop_im0 = tuple(shape_crop_im0)
        else:
            raise NotImplementedError("For now, shape_crop_im0 has to be one or two integer!")
        if isinstance(shape_crop_im1, int):
            shape_crop_im1 = (shape_crop_im1, shape_crop_im1)
        elif isinstance(shape_crop_im1, (list, tuple)) and len(shape_crop_im1) == 2:
            shape_crop_im1 = tuple(shape_crop_im1)
        else:
            raise NotImplementedError("For now, shape_crop_im1 has to be one or two integer!")
        for i in range(1, params.multipass.number):
            shape_crop_im0 = (shape_crop_im0[0] / coeffs_zoom[i - 1],
                              shape_crop_im0[1] / coeffs_zoom[i - 1])
            shape_crop_im1 = (shape_crop_im1[0] / coeffs_zoom[i - 1],
                              shape_crop_im1[1] / coeffs_zoom[i - 1])
            work_piv = WorkPIVFromDisplacement(params, index_pass=i,
                                               shape_crop_im0=shape_crop_im0,
                                               shape_crop_im1=shape_crop_im1)
            self.works_piv.append(work_piv)
            self.works_fix.append(WorkFIX(params.fix, work_piv))

    keep_maps = True  # results carry every correlation map, like the reference

    def calcul(self, couple):
        """multipass.py:172-190, all passes on the GPU (pipeline.py)."""
        from ...data_objects import ArrayCouple
        from .pipeline import run_pairs

        if isinstance(couple, dict):
            couple = ArrayCouple(**couple)
        elif isinstance(couple, (tuple, list)):
            couple = ArrayCouple(arrays=couple)
        if not isinstance(couple, ArrayCouple):
            raise ValueError(f"not isinstance(couple, ArrayCouple), {type(couple) = }")
        couple.apply_mask(self.params.mask)
        im0, im1 = couple.get_arrays()
        pipe, passes = run_pairs(self, [(im0, im1)], keep_maps=self.keep_maps)
        return pipe.to_host(passes, 0, couple=couple)

    def calcul_batch(self, pairs, keep_maps=False):
        """Several same-shape pairs in one pipeline: every pass is one
        correlation launch over the windows of all pairs."""
        from .pipeline import run_pairs

        pipe, passes = run_pairs(self, list(pairs), keep_maps=keep_maps)
        return [pipe.to_host(passes, k) for k in range(len(pairs))]

    def calcul_passes_per_work(self, couple):
        """The reference's pass-by-pass chain (FirstWorkPIV, WorkFIX,
        WorkPIVFromDisplacement ... with host results between passes)."""
        results = MultipassPIVResults()
        piv_result = couple
        for work_piv, work_fix in zip(self.works_piv, self.works_fix):
            piv_result = work_piv.calcul(piv_result)
            piv_result = work_fix.calcul(piv_result)
            results.append(piv_result)
        try:
            work_piv.apply_interp(piv_result, last=True)
        except InterpError as e:
            print("Warning: InterpError at the end of the last piv pass:", e)
        return results

    def calcul_from_arrays(self, *arrays, names=None):
        """src/fluidimage/works/__init__.py:173-179"""
        names = [f"array{i}" for i in range(len(arrays))]
        t_start = time.perf_counter()
        result = self.calcul({"arrays": arrays, "names": names})
        print(f"Calcul done in {time.perf_counter() - t_start:.2f} s")
        return result

    def _prepare_with_image(self, im=None, imshape=None):
        if imshape is None:
            imshape = im.shape
        for work_piv in self.works_piv:
            work_piv._prepare_with_image(imshape=imshape)


Work = WorkPIV

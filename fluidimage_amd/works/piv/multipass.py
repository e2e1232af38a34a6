"""Multipass PIV (src/fluidimage/works/piv/multipass.py:18-198) on the GPU."""

import time

from ...data_objects import MultipassPIVResults
from ...params import Params
from ..with_mask import BaseWorkWithMask
from .fix import WorkFIX
from .interp import InterpError
from .schedule import pass_schedule
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
                # deformation between passes (None or "bicubic"), gradient
                # refinement steps after the last pass (0: none) and their
                # unknowns ("tps": TPS node values smoothed by the TPS of
                # subdom_size and smoothing_coef; "bilinear": the node
                # displacements themselves)
                "image_deformation": None,
                "refine_steps": 0,
                "refine_field": "tps",
            },
        )

    def __init__(self, params=None):
        self.params = params = params if params is not None else self.create_default_params()
        sizes = pass_schedule(params.piv0.shape_crop_im0, params.piv0.shape_crop_im1,
                              params.multipass.number, params.multipass.coeff_zoom)
        self.works_piv = [FirstWorkPIV(params)] + [
            WorkPIVFromDisplacement(params, index_pass=k, shape_crop_im0=s0, shape_crop_im1=s1)
            for k, (s0, s1) in enumerate(sizes[1:], start=1)]
        self.works_fix = [WorkFIX(params.fix, w) for w in self.works_piv]

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
        """The per-work chain of the reference (multipass.py:172-190): each
        pass's work and WorkFIX with host results in between, then the final
        interpolation of the last result."""
        results = MultipassPIVResults()
        state = couple
        for stage in zip(self.works_piv, self.works_fix):
            state = stage[1].calcul(stage[0].calcul(state))
            results.append(state)
        try:
            self.works_piv[-1].apply_interp(state, last=True)
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
        shape = im.shape if imshape is None else imshape
        for w in self.works_piv:
            w._prepare_with_image(imshape=shape)


Work = WorkPIV

"""Single PIV pass on the GPU with the reference's WorkPIV semantics.

Mirrors src/fluidimage/works/piv/singlepass.py: ``FirstWorkPIV`` (first
pass) and ``WorkPIVFromDisplacement`` (later passes) with the same
parameters, grid (``_prepare_with_image``, :108-141), padding and crop
geometry (:196-208,363-407), window positions (:210-260,787-870), per-window
results (:262-361) and interpolation selector (``apply_interp``, :409-563).

The whole per-window loop is one launch of the fused HIP kernel over all
windows (``run_device``); ``_loop_vectors`` keeps the reference's host
signature on top of it.
"""

import ctypes
from copy import deepcopy

import numpy as np

from ... import _native
from ...calcul.correl import correlation_classes
from ...data_objects import ArrayCouple, HeavyPIVResults, error_strings
from ...device import ImageSet, piv_dtype_of, to_device_image
from ..with_mask import BaseWorkWithMask
from .schedule import grid_positions, grid_step, window_shapes


class InterpError(ValueError):
    pass


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise _native.NativeLibraryError("the PIV works need a ROCm GPU (torch.cuda)")
    return torch


class BaseWorkPIV(BaseWorkWithMask):
    """Base of the single-pass works (singlepass.py:41-563)."""

    index_pass = 0

    def _init_shape_crop(self, shape_crop_im0, shape_crop_im1):
        """Integer window shapes of this pass (singlepass.py:52-87)."""
        self.shape_crop_im0, self.shape_crop_im1 = window_shapes(shape_crop_im0, shape_crop_im1)

    def _init_correl(self):
        """The correlation object of this pass, chosen in the registry by
        params.piv0.method_correl (singlepass.py:89-106)."""
        p0 = self.params.piv0
        cls = correlation_classes.get(p0.method_correl)
        if cls is None:
            raise ValueError(
                "params.piv0.method_correl should be in " + str(list(correlation_classes.keys())))
        if not hasattr(cls, "correlate_windows"):
            raise NotImplementedError(
                f"method_correl {p0.method_correl!r}: the batched WorkPIV runs "
                "the fused FFT correlation ('fftw' / 'hip'); the direct correlation is "
                "available through its per-window and window-stack API "
                f"({cls.__name__})")
        options = {key: getattr(p0, key) for key in (
            "method_subpix", "nsubpix", "displacement_max", "particle_radius",
            "nb_peaks_to_search")}
        self.correl = cls(im0_shape=self.shape_crop_im0, im1_shape=self.shape_crop_im1,
                          **options)

    def _init_crop(self):
        """singlepass.py:363-389"""
        s0 = self.shape_crop_im0
        s1 = self.shape_crop_im1
        self._start_for_crop0 = (s0[0] // 2, s0[1] // 2)
        self._stop_for_crop0 = (s0[0] // 2 + s0[0] % 2, s0[1] // 2 + s0[1] % 2)
        self._start_for_crop1 = (s1[0] // 2, s1[1] // 2)
        self._stop_for_crop1 = (s1[0] // 2 + s1[0] % 2, s1[1] // 2 + s1[1] % 2)
        # _pad_images (singlepass.py:204): largest half window
        self.npad = max(self._start_for_crop0 + self._stop_for_crop0)

    @property
    def crop_geometry(self):
        return np.array(
            [*self._start_for_crop0, *self._stop_for_crop0, *self._start_for_crop1,
             *self._stop_for_crop1], dtype=np.int32)

    @property
    def has_to_apply_subpix(self):
        return self.index_pass == self.params.multipass.number - 1

    def _prepare_with_image(self, im0=None, imshape=None):
        """Grid of vectors (singlepass.py:108-141): window centres along y
        and x, flattened row-major (ivec = iy_index nx + ix_index)."""
        shape = im0.shape if imshape is None else imshape
        self.imshape0 = tuple(int(n) for n in shape)
        axes = [grid_positions(self.imshape0[a], self._start_for_crop0[a],
                               self._stop_for_crop0[a],
                               grid_step(self.shape_crop_im0[a], self.overlap))
                for a in (0, 1)]
        self.iyvecs, self.ixvecs = axes
        self.iyvecs_grid = np.repeat(self.iyvecs, len(self.ixvecs))
        self.ixvecs_grid = np.tile(self.ixvecs, len(self.iyvecs))
        self._dev_grid = None
        self._dev_grid_full = None

    def _device_grid(self, device):
        if self._dev_grid is None or self._dev_grid[0].device != device:
            torch = _torch()
            self._dev_grid = (
                torch.from_numpy(self.iyvecs.astype(np.int32)).to(device),
                torch.from_numpy(self.ixvecs.astype(np.int32)).to(device),
            )
        return self._dev_grid

    def _device_grid_full(self, device):
        """Flattened grid (x, y) as float64 device tensors (TPS targets)."""
        g = getattr(self, "_dev_grid_full", None)
        if g is None or g[0].device != device:
            torch = _torch()
            g = (torch.from_numpy(self.ixvecs_grid.astype(np.float64)).to(device),
                 torch.from_numpy(self.iyvecs_grid.astype(np.float64)).to(device))
            self._dev_grid_full = g
        return g

    @property
    def nb_vectors(self):
        return len(self.ixvecs) * len(self.iyvecs)

    # ---------------------------------------------------------------- GPU
    def make_windows(self, image0, image1, dx_approx=None, dy_approx=None, windows=None,
                     device=None, stream=None, npairs=1):
        """Window records of this pass for the image pair (image0, image1) of
        an ImageSet (and the next npairs - 1 pairs, images image0 + 2k,
        image1 + 2k); ``dx_approx``/``dy_approx`` (float64 device tensors on
        this pass's grid, pair-major) give the predictor of later passes."""
        torch = _torch()
        if device is None:
            device = torch.device("cuda")
        iyv, ixv = self._device_grid(device)
        nvec = self.nb_vectors * npairs
        if windows is None:
            windows = torch.empty((nvec, 10), dtype=torch.int32, device=device)
        xs = torch.empty(nvec, dtype=torch.float64, device=device)
        ys = torch.empty(nvec, dtype=torch.float64, device=device)
        dxin = torch.empty(nvec, dtype=torch.int32, device=device)
        dyin = torch.empty(nvec, dtype=torch.int32, device=device)
        geom = self.crop_geometry
        H, W = self.imshape0
        lib = _native.load()
        _native.check(
            lib.piv_make_windows(
                int(npairs), len(self.iyvecs), len(self.ixvecs), ctypes.c_void_p(iyv.data_ptr()),
                ctypes.c_void_p(ixv.data_ptr()),
                None if dx_approx is None else ctypes.c_void_p(dx_approx.data_ptr()),
                None if dy_approx is None else ctypes.c_void_p(dy_approx.data_ptr()),
                H, W, self.npad, geom.ctypes.data_as(ctypes.c_void_p), int(image0), int(image1),
                ctypes.c_void_p(windows.data_ptr()), ctypes.c_void_p(xs.data_ptr()),
                ctypes.c_void_p(ys.data_ptr()), ctypes.c_void_p(dxin.data_ptr()),
                ctypes.c_void_p(dyin.data_ptr()), _native.stream_ptr(stream)),
            "piv_make_windows",
        )
        return {"windows": windows, "xs": xs, "ys": ys, "dx_in": dxin, "dy_in": dyin}

    def run_device(self, imageset, image0=0, image1=1, dx_approx=None, dy_approx=None,
                   keep_maps=False, stream=None):
        """One pass for one pair, all on the device. Returns a dict of
        device tensors (dx, dy, correl_max, err, xs, ys, dx_in, dy_in, ...)."""
        torch = _torch()
        dev = imageset.device
        out = self.make_windows(image0, image1, dx_approx, dy_approx, device=dev, stream=stream)
        nvec = self.nb_vectors
        maps = None
        if keep_maps:
            maps = torch.empty((nvec,) + self.shape_crop_im0, dtype=torch.float32, device=dev)
        res = self.correl.correlate_windows(imageset.table, out["windows"], nvec,
                                            imageset.dtype, self.has_to_apply_subpix, maps=maps,
                                            stream=stream)
        out.update(res)
        out["maps"] = maps
        return out

    # ---------------------------------------------------------------- host API
    def _pad_images(self, im0, im1):
        """Images keep their shape on the device: the padding is virtual
        (zeros outside the image, min subtracted inside)."""
        return im0, im1

    def _loop_vectors(self, im0, im1, deltaxs_input=None, deltays_input=None,
                      keep_maps=True):
        """singlepass.py:262-361: returns (deltaxs, deltays, xs, ys,
        correls_max, correls, errors, secondary_peaks)."""
        torch = _torch()
        dev = torch.device("cuda")
        images = ImageSet([to_device_image(im0, dev), to_device_image(im1, dev)])
        if not hasattr(self, "ixvecs_grid"):
            self._prepare_with_image(imshape=images.images[0].shape)
        dxa = dya = None
        if deltaxs_input is not None:
            dxa = torch.from_numpy(np.asarray(deltaxs_input, np.float64)).to(dev)
            dya = torch.from_numpy(np.asarray(deltays_input, np.float64)).to(dev)
        out = self.run_device(images, 0, 1, dxa, dya, keep_maps=keep_maps)
        torch.cuda.current_stream().synchronize()
        return self._host_results(out, first_pass=deltaxs_input is None)

    def _host_results(self, out, first_pass):
        h = {k: v.cpu().numpy() for k, v in out.items() if v is not None and k != "windows"}
        self._last_sec_sub = h.get("sec_sub")
        errors = error_strings(h["err"])
        if first_pass:
            xs = self.ixvecs_grid.copy()
            ys = self.iyvecs_grid.copy()
        else:
            xs = h["xs"]
            ys = h["ys"]
        correls = None
        if out.get("maps") is not None:
            correls = list(h["maps"])
            for i in np.flatnonzero(h["err"] == _native.WIN_BAD_CROP):
                correls[i] = None
        nb = self.params.piv0.nb_peaks_to_search
        secondary = [None] * len(xs)
        if nb > 1:
            for i in range(len(xs)):
                e = int(h["err"][i])
                if e in (_native.WIN_OK, _native.WIN_SUBPIX_WRONG, _native.WIN_SUBPIX_BOUNDARY):
                    secondary[i] = [
                        (int(a), int(b), np.float32(c))
                        for (a, b, c) in h["sec"][i, : int(h["nsec"][i])]
                    ]
        return (h["dx"], h["dy"], xs, ys, h["correl_max"], correls, errors, secondary)

    def _complete_result(self, result):
        result._sec_sub = getattr(self, "_last_sec_sub", None)
        result.indices_no_displacement = self.correl.get_indices_no_displacement()
        result.displacement_max = self.correl.displacement_max

    def apply_interp(self, piv_results, last=False):
        """singlepass.py:409-563. TPS subdomains (GPU, FP64) or griddata."""
        from .interp import apply_interp

        return apply_interp(self, piv_results, last=last)


class FirstWorkPIV(BaseWorkPIV):
    """First PIV pass (singlepass.py:566-682)."""

    index_pass = 0

    @classmethod
    def _complete_params_with_default(cls, params):
        params._set_child(
            "piv0",
            attribs={
                "shape_crop_im0": 48,
                "shape_crop_im1": None,
                "displacement_max": None,
                "displacement_mean": None,
                "method_correl": "fftw",
                "method_subpix": "2d_gaussian2",
                "nsubpix": None,
                "nb_peaks_to_search": 1,
                "particle_radius": 3,
            },
        )
        params.piv0._set_child("grid", attribs={"overlap": 0.5, "from": "overlap"})
        cls._complete_params_with_default_mask(params)

    def __init__(self, params):
        self.params = params
        self.overlap = params.piv0.grid.overlap
        if self.overlap >= 1:
            raise ValueError("params.piv0.grid.overlap has to be smaller than 1")
        self._init_shape_crop(params.piv0.shape_crop_im0, params.piv0.shape_crop_im1)
        self._init_crop()
        self._init_correl()

    def calcul(self, couple):
        """singlepass.py:143-190"""
        if isinstance(couple, dict):
            couple = ArrayCouple(**couple)
        if not isinstance(couple, ArrayCouple):
            raise ValueError(f"not isinstance(couple, ArrayCouple), {type(couple) = }")
        couple.apply_mask(self.params.mask)
        im0, im1 = couple.get_arrays()
        if not hasattr(self, "ixvecs_grid"):
            self._prepare_with_image(im0)
        (deltaxs, deltays, xs, ys, correls_max, correls, errors,
         secondary_peaks) = self._loop_vectors(im0, im1)
        xs, ys = self._xyoriginalimage_from_xymasked(xs, ys)
        result = HeavyPIVResults(deltaxs, deltays, xs, ys, errors, correls_max=correls_max,
                                 correls=correls, couple=deepcopy(couple), params=self.params,
                                 secondary_peaks=secondary_peaks)
        self._complete_result(result)
        return result


class WorkPIVFromDisplacement(BaseWorkPIV):
    """Later passes (singlepass.py:685-870)."""

    def __init__(self, params, index_pass=1, shape_crop_im0=None, shape_crop_im1=None):
        self.params = params
        self.index_pass = index_pass
        self.overlap = params.piv0.grid.overlap
        if shape_crop_im0 is None:
            shape_crop_im0 = params.piv0.shape_crop_im0
        if shape_crop_im1 is None:
            shape_crop_im1 = params.piv0.shape_crop_im1
        self._init_shape_crop(shape_crop_im0, shape_crop_im1)
        self._init_crop()
        self._init_correl()

    def calcul(self, piv_results):
        """singlepass.py:733-785"""
        if not isinstance(piv_results, HeavyPIVResults):
            raise ValueError
        couple = piv_results.couple
        im0, im1 = couple.get_arrays()
        if not hasattr(self, "ixvecs_grid"):
            self._prepare_with_image(im0)
        self.apply_interp(piv_results)
        deltaxs_input = np.round(piv_results.deltaxs_approx).astype("int32")
        deltays_input = np.round(piv_results.deltays_approx).astype("int32")
        (deltaxs, deltays, xs, ys, correls_max, correls, errors,
         secondary_peaks) = self._loop_vectors(im0, im1, deltaxs_input, deltays_input)
        xs, ys = self._xyoriginalimage_from_xymasked(xs, ys)
        result = HeavyPIVResults(deltaxs, deltays, xs, ys, errors, correls_max=correls_max,
                                 correls=correls, couple=deepcopy(couple), params=self.params,
                                 secondary_peaks=secondary_peaks)
        self._complete_result(result)
        result.deltaxs_input = deltaxs_input
        result.deltays_input = deltays_input
        return result

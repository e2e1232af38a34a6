"""Correlation classes backed by the gfx950 HIP library.

``CorrelHIP`` is a drop-in for the reference's ``CorrelFFTW``
(src/fluidimage/calcul/correl.py:167-281,513-568,639-692): same constructor,
same ``__call__(im0, im1) -> (correl, norm)``, same peak search, error and
sub-pixel semantics, same index convention ``(iy0, ix0) = (n0//2 - 1,
n1//2 - 1)``. Every numerical step runs on the GPU through libpiv_hip; the
host only prepares the search mask (integer geometry, once per plan).

``correlation_classes`` mirrors the reference registry
(correl.py:709-713); ``register(registry)`` inserts the HIP class into
fluidimage's own ``correlation_classes`` dict so that
``params.piv0.method_correl = "hip"`` selects it.
"""

import ctypes
from math import sqrt

import numpy as np

from .. import _native
from .errors import PIVError


def parse_displacement_max(displ_max, im0_shape):
    """Same rule as src/fluidimage/calcul/correl.py:68-73 ("N%" strings)."""
    if isinstance(displ_max, str) and displ_max.endswith("%"):
        return float(displ_max[:-1]) / 100 * max(im0_shape)
    return displ_max


def compute_search_mask(shape, displacement_max, ix0, iy0):
    """where_large_displacement and the search box of
    src/fluidimage/calcul/correl.py:518-547 (host-side geometry)."""
    n0, n1 = shape
    i1 = np.arange(n1)[None, :]
    i0 = np.arange(n0)[:, None]
    dx = ix0 - i1
    dy = iy0 - i0
    # the reference compares math.sqrt(dx**2 + dy**2) with displacement_max;
    # np.sqrt of the exact integer sum is the same correctly rounded value
    wld = np.sqrt((dx**2 + dy**2).astype(np.float64)) > displacement_max
    rows_all = wld.all(axis=1)
    cols_all = wld.all(axis=0)

    def first_false(flags):
        for k, f in enumerate(flags):
            if not f:
                return k
        return len(flags) - 1

    i0_start = first_false(rows_all)
    i1_start = first_false(cols_all)
    i0_stop = len(rows_all) - 1 - first_false(rows_all[::-1]) + 1
    i1_stop = len(cols_all) - 1 - first_false(cols_all[::-1]) + 1
    return wld, (i0_start, i0_stop), (i1_start, i1_stop)


class SubPixParams:
    """Validation of src/fluidimage/calcul/subpix.py:73-91."""

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
        self.n = 1 if nsubpix is None else int(nsubpix)


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise _native.NativeLibraryError("CorrelHIP needs a ROCm GPU (torch.cuda)")
    return torch


class CorrelHIP:
    """FFT cross-correlation on gfx950 behind the CorrelFFTW interface."""

    _tag = "hip"

    def __init__(
        self,
        im0_shape,
        im1_shape=None,
        method_subpix="centroid",
        nsubpix=1,
        displacement_max=None,
        particle_radius=3,
        nb_peaks_to_search=1,
        mode=None,
    ):
        im0_shape = tuple(int(n) for n in im0_shape)
        if im1_shape is None:
            im1_shape = im0_shape
        im1_shape = tuple(int(n) for n in im1_shape)
        self.mode = mode
        self.subpix = SubPixParams(method=method_subpix, nsubpix=nsubpix)
        self.im0_shape = im0_shape
        self.im1_shape = im1_shape
        self.iy0, self.ix0 = (i // 2 - 1 for i in im0_shape)
        self.displacement_max = parse_displacement_max(displacement_max, im0_shape)
        self.particle_radius = particle_radius
        self.nb_peaks_to_search = nb_peaks_to_search
        self.start_stop_for_search0 = [0, None]
        self.start_stop_for_search1 = [0, None]
        if self.displacement_max is not None:
            wld, ss0, ss1 = compute_search_mask(
                im0_shape, self.displacement_max, self.ix0, self.iy0
            )
            self.where_large_displacement = wld
            self.start_stop_for_search0 = ss0
            self.start_stop_for_search1 = ss1
        if im0_shape != im1_shape:
            raise ValueError(
                "with this correlation method the input images "
                "have to have the same shape."
            )
        if nb_peaks_to_search < 1:
            raise ValueError("nb_peaks_to_search has to be >= 1")
        self._plan = None
        self._create_plan()

    # ------------------------------------------------------------ plan
    def _create_plan(self):
        lib = _native.load()
        n0, n1 = self.im0_shape
        if not lib.piv_shape_supported(n0, n1):
            raise NotImplementedError(
                f"window shape {self.im0_shape} is not supported by libpiv_hip "
                "(rows and columns of 2 to 256, at most 16384 elements)"
            )
        plan = ctypes.c_void_p()
        if self.displacement_max is None:
            dmax, mask_p, box_p = -1.0, None, None
            self._mask_buf = None
        else:
            dmax = float(self.displacement_max)
            self._mask_buf = np.ascontiguousarray(self.where_large_displacement, np.uint8)
            self._box_buf = np.array(
                [*self.start_stop_for_search0, *self.start_stop_for_search1], np.int32
            )
            mask_p = self._mask_buf.ctypes.data_as(ctypes.c_void_p)
            box_p = self._box_buf.ctypes.data_as(ctypes.c_void_p)
            if dmax < 0:
                raise ValueError("displacement_max must be >= 0")
        _native.check(
            lib.piv_plan_create(
                ctypes.byref(plan), n0, n1, dmax, mask_p, box_p,
                _native.PIV_SUBPIX[self.subpix.method], self.subpix.n,
                int(self.particle_radius), int(self.nb_peaks_to_search),
            ),
            "piv_plan_create",
        )
        self._plan = plan
        self._lib = lib

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                self._lib.piv_plan_destroy(plan)
            except Exception:
                pass

    @property
    def plan(self):
        return self._plan

    # ------------------------------------------------------------ indices
    def compute_displacement_from_indices(self, ix, iy):
        """correl.py:209-211"""
        return self.ix0 - ix, self.iy0 - iy

    def compute_indices_from_displacement(self, dx, dy):
        """correl.py:213-215"""
        return self.ix0 - dx, self.iy0 - dy

    def get_indices_no_displacement(self):
        """correl.py:217-219"""
        return self.iy0, self.ix0

    # ------------------------------------------------------------ results
    def _alloc_results(self, nwin, device):
        torch = _torch()
        kw = dict(device=device)
        out = {
            "dx": torch.empty(nwin, dtype=torch.float32, **kw),
            "dy": torch.empty(nwin, dtype=torch.float32, **kw),
            "correl_max": torch.empty(nwin, dtype=torch.float32, **kw),
            "err": torch.empty(nwin, dtype=torch.int32, **kw),
            "ix": torch.empty(nwin, dtype=torch.int32, **kw),
            "iy": torch.empty(nwin, dtype=torch.int32, **kw),
            "norm": torch.empty(nwin, dtype=torch.float64, **kw),
        }
        ns = self.nb_peaks_to_search - 1
        if ns > 0:
            out["sec"] = torch.empty((nwin, ns, 3), dtype=torch.float32, **kw)
            out["nsec"] = torch.empty(nwin, dtype=torch.int32, **kw)
            out["sec_sub"] = torch.empty((nwin, ns, 2), dtype=torch.float32, **kw)
        return out

    @staticmethod
    def results_struct(out):
        def ptr(key):
            t = out.get(key)
            return None if t is None else t.data_ptr()

        return _native.PivResults(*(ptr(k) for k in (
            "dx", "dy", "correl_max", "err", "ix", "iy", "norm", "sec", "nsec", "sec_sub")))

    # ------------------------------------------------------------ batched API
    def correlate_windows(self, images, windows, nwin, dtype, apply_subpix, maps=None,
                          out=None, stream=None, map_square=True):
        """Run the fused kernel on ``nwin`` windows.

        ``images``: device tensor of ``PivImage`` records (uint8 bytes),
        ``windows``: device tensor of ``PivWindow`` records, ``dtype``: one of
        ``_native.PIV_U8/U16/F32``. With ``map_square`` (and
        displacement_max None) the maps keep the second-chance NaN square
        like the maps of the reference's ``_loop_vectors``. Returns the dict
        of result tensors.
        """
        torch = _torch()
        if out is None:
            out = self._alloc_results(nwin, images.device)
        res = self.results_struct(out)
        options = (_native.PIV_OPT_SUBPIX if apply_subpix else 0) | (
            _native.PIV_OPT_MAP_SQUARE if map_square else 0)
        _native.check(
            self._lib.piv_correlate_batch(
                self._plan, int(dtype), ctypes.c_void_p(images.data_ptr()),
                ctypes.c_void_p(windows.data_ptr()), int(nwin), int(options),
                None if maps is None else ctypes.c_void_p(maps.data_ptr()),
                ctypes.byref(res), _native.stream_ptr(stream),
            ),
            "piv_correlate_batch",
        )
        return out

    # ------------------------------------------------------------ per-window API
    def __call__(self, im0, im1):
        """Correlation map (reference layout, float32) and norm of one pair
        of windows (correl.py:648-662)."""
        torch = _torch()
        from ..device import ImageSet, window_table

        a0 = np.ascontiguousarray(im0, dtype=np.float32)
        a1 = np.ascontiguousarray(im1, dtype=np.float32)
        if a0.shape != self.im0_shape or a1.shape != self.im1_shape:
            raise ValueError("image shapes differ from the plan shapes")
        dev = torch.device("cuda")
        imgs = ImageSet([torch.from_numpy(a0).to(dev), torch.from_numpy(a1).to(dev)],
                        subtract_min=False)
        wins = window_table([(0, 1, 0, 0, 0, 0, 0, 0, 0)], dev)
        maps = torch.empty((1,) + self.im0_shape, dtype=torch.float32, device=dev)
        out = self.correlate_windows(imgs.table, wins, 1, _native.PIV_F32, False, maps=maps,
                                     map_square=False)
        torch.cuda.current_stream().synchronize()
        return maps[0].cpu().numpy(), float(out["norm"][0].item())

    def compute_displacements_from_correl(self, correl, norm=1.0):
        """Displacement of the correlation peak (correl.py:226-275,561-568).

        Returns ``(dx, dy, correl_max, other_peaks)`` or raises ``PIVError``
        with ``results = (dx, dy, correl_max)``. With ``displacement_max``
        None, the second-chance NaN square is written into ``correl`` in
        place, as in the reference.
        """
        torch = _torch()
        dev = torch.device("cuda")
        arr = np.asarray(correl)
        if arr.shape != self.im0_shape:
            raise ValueError("correl has not the plan shape")
        maps = torch.from_numpy(np.ascontiguousarray(arr, np.float32)).to(dev)[None]
        norms = torch.tensor([float(norm)], dtype=torch.float64, device=dev)
        out = self._alloc_results(1, dev)
        res = self.results_struct(out)
        _native.check(
            self._lib.piv_peaks_from_maps(
                self._plan, ctypes.c_void_p(maps.data_ptr()), ctypes.c_void_p(norms.data_ptr()),
                1, 0, ctypes.byref(res), _native.stream_ptr()),
            "piv_peaks_from_maps",
        )
        torch.cuda.current_stream().synchronize()
        host = {k: v.cpu().numpy() for k, v in out.items()}
        if self.displacement_max is None and isinstance(correl, np.ndarray):
            back = maps[0].cpu().numpy()
            nan = np.isnan(back) & ~np.isnan(correl)
            if nan.any() and correl.flags.writeable:
                correl[nan] = np.nan
        err = int(host["err"][0])
        dx = host["dx"][0]
        dy = host["dy"][0]
        cmax = host["correl_max"][0]
        if err != _native.WIN_OK:
            e = PIVError(explanation=_native.EXPLANATIONS[err])
            e.results = (int(dx) if dx == dx else dx, int(dy) if dy == dy else dy, cmax)
            raise e
        dx, dy = int(dx), int(dy)
        if self.nb_peaks_to_search == 1:
            other = None
        else:
            other = []
            for k in range(int(host["nsec"][0])):
                sdx, sdy, sc = host["sec"][0, k]
                other.append((int(sdx), int(sdy), np.float32(sc)))
        return dx, dy, np.float32(cmax), other

    def apply_subpix(self, dx, dy, correl):
        """Sub-pixel displacement (correl.py:277-281, subpix.py:107-218)."""
        torch = _torch()
        dev = torch.device("cuda")
        ix, iy = self.compute_indices_from_displacement(dx, dy)
        arr = np.ascontiguousarray(np.asarray(correl), np.float32)
        maps = torch.from_numpy(arr).to(dev)[None]
        tix = torch.tensor([int(ix)], dtype=torch.int32, device=dev)
        tiy = torch.tensor([int(iy)], dtype=torch.int32, device=dev)
        ox = torch.empty(1, dtype=torch.float32, device=dev)
        oy = torch.empty(1, dtype=torch.float32, device=dev)
        er = torch.empty(1, dtype=torch.int32, device=dev)
        _native.check(
            self._lib.piv_subpix_at(
                self._plan, ctypes.c_void_p(maps.data_ptr()), ctypes.c_void_p(tix.data_ptr()),
                ctypes.c_void_p(tiy.data_ptr()), 1, ctypes.c_void_p(ox.data_ptr()),
                ctypes.c_void_p(oy.data_ptr()), ctypes.c_void_p(er.data_ptr()),
                _native.stream_ptr()),
            "piv_subpix_at",
        )
        torch.cuda.current_stream().synchronize()
        err = int(er.item())
        if err != _native.WIN_OK:
            raise PIVError(explanation=_native.EXPLANATIONS[err],
                           result_compute_subpix=(iy, ix))
        return self.compute_displacement_from_indices(float(ox.item()), float(oy.item()))


def _displacements_from_indices(self, ix, iy):
    """compute_displacement_from_indices (correl.py:209-211) under the plural
    name some callers use, for scalars or arrays of indices (ix, iy):
    returns (ix0 - ix, iy0 - iy)."""
    if np.ndim(ix) or np.ndim(iy):
        return self.ix0 - np.asarray(ix), self.iy0 - np.asarray(iy)
    return self.compute_displacement_from_indices(ix, iy)


CorrelHIP.compute_displacements_from_indices = _displacements_from_indices


class CorrelDirectHIP:
    """Direct (spatial) correlation on gfx950 behind the interface of the
    reference's CorrelPythran / CorrelPyCuda (src/fluidimage/calcul/correl.py:
    284-451): maps of (2 dmax + 1)^2 with the zero displacement at (dmax,
    dmax), CorrelBase's peak search without mask (the second-chance square is
    written into the caller's map) and sub-pixel fit. The windows may differ
    in shape (im1 no larger than im0). Map, norm, search and sub-pixel fit run
    on the GPU (piv_correlate_direct, piv_peaks_any, piv_subpix_any).

    The reference's ``_finalize_init`` only binds its local
    ``displacement_max`` when the parameter is None (correl.py:388-408), so an
    explicit value raises there; here an explicit value is used as given."""

    _tag = "direct"
    _kind = "pythran"

    def __init__(self, im0_shape, im1_shape=None, method_subpix="centroid", nsubpix=1,
                 displacement_max=None, particle_radius=3, nb_peaks_to_search=1, mode=None):
        im0_shape = tuple(int(n) for n in im0_shape)
        im1_shape = im0_shape if im1_shape is None else tuple(int(n) for n in im1_shape)
        self.mode = mode
        self.subpix = SubPixParams(method=method_subpix, nsubpix=nsubpix)
        self.im0_shape, self.im1_shape = im0_shape, im1_shape
        dmax = parse_displacement_max(displacement_max, im0_shape)
        if dmax is None:
            dmax = self._default_displacement_max()
        if dmax <= 0:
            raise ValueError("displacement_max <= 0 : problem with images shapes?")
        self.displacement_max = int(dmax)
        self.ix0 = self.iy0 = self.displacement_max
        self.particle_radius = particle_radius
        self.nb_peaks_to_search = nb_peaks_to_search
        if nb_peaks_to_search < 1:
            raise ValueError("nb_peaks_to_search has to be >= 1")
        self.start_stop_for_search0 = [0, None]
        self.start_stop_for_search1 = [0, None]
        self._lib = _native.load()
        self._norm_code = (_native.PIV_DIRECT_PYCUDA if self._kind == "pycuda"
                           else _native.PIV_DIRECT_PYTHRAN)

    def _default_displacement_max(self):
        """correl.py:388-408 (pythran) / 418-442 (pycuda)."""
        if self.im0_shape == self.im1_shape:
            return min(self.im0_shape) // 2 - (0 if self._kind == "pycuda" else 1)
        return min(self.im0_shape[0] - self.im1_shape[0],
                   self.im0_shape[1] - self.im1_shape[1]) // 2 - 1

    @property
    def map_shape(self):
        n = 2 * self.displacement_max + 1
        return (n, n)

    def compute_displacement_from_indices(self, ix, iy):
        return self.ix0 - ix, self.iy0 - iy

    def compute_indices_from_displacement(self, dx, dy):
        return self.ix0 - dx, self.iy0 - dy

    def get_indices_no_displacement(self):
        return self.iy0, self.ix0

    _alloc_results = CorrelHIP._alloc_results
    results_struct = staticmethod(CorrelHIP.results_struct)

    # ------------------------------------------------------------ batched API
    def correlate_stack(self, win0, win1, stream=None):
        """Maps and norms of a stack of window pairs (device float32 tensors
        (nwin, n0y, n0x) and (nwin, n1y, n1x)): returns (maps (nwin, 2 dmax +
        1, 2 dmax + 1) float32, norms (nwin,) float64) on the device."""
        torch = _torch()
        nwin = int(win0.shape[0])
        if tuple(win0.shape[1:]) != self.im0_shape or tuple(win1.shape[1:]) != self.im1_shape \
                or int(win1.shape[0]) != nwin:
            raise ValueError("window stacks do not have the shapes of the correlation")
        win0 = win0.to(torch.float32).contiguous()
        win1 = win1.to(torch.float32).contiguous()
        maps = torch.empty((nwin,) + self.map_shape, dtype=torch.float32, device=win0.device)
        norms = torch.empty(nwin, dtype=torch.float64, device=win0.device)
        (n0y, n0x), (n1y, n1x) = self.im0_shape, self.im1_shape
        _native.check(
            self._lib.piv_correlate_direct(
                nwin, ctypes.c_void_p(win0.data_ptr()), n0y, n0x,
                ctypes.c_void_p(win1.data_ptr()), n1y, n1x, self.displacement_max,
                self._norm_code, ctypes.c_void_p(maps.data_ptr()),
                ctypes.c_void_p(norms.data_ptr()), _native.stream_ptr(stream)),
            "piv_correlate_direct")
        return maps, norms

    def peaks_stack(self, maps, norms, apply_subpix=True, stream=None):
        """Peak search (+ sub-pixel fit) of a stack of maps on the device
        (the second-chance squares are written into ``maps``); returns the
        dict of result tensors (dx, dy, correl_max, err, ix, iy ...)."""
        nwin = int(maps.shape[0])
        out = self._alloc_results(nwin, maps.device)
        res = self.results_struct(out)
        n0, n1 = self.map_shape
        _native.check(
            self._lib.piv_peaks_any(
                nwin, ctypes.c_void_p(maps.data_ptr()), ctypes.c_void_p(norms.data_ptr()),
                n0, n1, self.iy0, self.ix0, _native.PIV_SUBPIX[self.subpix.method],
                self.subpix.n, int(self.particle_radius), int(self.nb_peaks_to_search),
                int(bool(apply_subpix)), ctypes.byref(res), _native.stream_ptr(stream)),
            "piv_peaks_any")
        return out

    # ------------------------------------------------------------ per-window API
    def __call__(self, im0, im1):
        """(correl, norm) of one pair of windows (correl_numpy / correl_pycuda)."""
        torch = _torch()
        dev = torch.device("cuda")
        a0 = np.ascontiguousarray(im0, dtype=np.float32)
        a1 = np.ascontiguousarray(im1, dtype=np.float32)
        if a0.shape != self.im0_shape or a1.shape != self.im1_shape:
            raise ValueError("image shapes differ from the correlation's shapes")
        maps, norms = self.correlate_stack(torch.from_numpy(a0).to(dev)[None],
                                           torch.from_numpy(a1).to(dev)[None])
        torch.cuda.current_stream().synchronize()
        return maps[0].cpu().numpy(), float(norms[0].item())

    def compute_displacements_from_correl(self, correl, norm=1.0):
        """CorrelBase.compute_displacements_from_correl (correl.py:226-275):
        ``(dx, dy, correl_max, other_peaks)`` or ``PIVError``; the
        second-chance NaN square is written into ``correl``."""
        torch = _torch()
        dev = torch.device("cuda")
        arr = np.asarray(correl)
        if arr.shape != self.map_shape:
            raise ValueError("correl has not the map shape of the correlation")
        maps = torch.from_numpy(np.ascontiguousarray(arr, np.float32)).to(dev)[None]
        norms = torch.tensor([float(norm)], dtype=torch.float64, device=dev)
        out = self.peaks_stack(maps, norms, apply_subpix=False)
        torch.cuda.current_stream().synchronize()
        host = {k: v.cpu().numpy() for k, v in out.items()}
        if isinstance(correl, np.ndarray) and correl.flags.writeable:
            back = maps[0].cpu().numpy()
            nan = np.isnan(back) & ~np.isnan(correl)
            if nan.any():
                correl[nan] = np.nan
        err = int(host["err"][0])
        dx, dy, cmax = host["dx"][0], host["dy"][0], host["correl_max"][0]
        if err != _native.WIN_OK:
            e = PIVError(explanation=_native.EXPLANATIONS[err])
            e.results = (int(dx) if dx == dx else dx, int(dy) if dy == dy else dy, cmax)
            raise e
        other = None
        if self.nb_peaks_to_search > 1:
            other = [(int(a), int(b), np.float32(c))
                     for (a, b, c) in host["sec"][0, :int(host["nsec"][0])]]
        return int(dx), int(dy), np.float32(cmax), other

    def apply_subpix(self, dx, dy, correl):
        """correl.py:277-281 on the direct map."""
        torch = _torch()
        dev = torch.device("cuda")
        ix, iy = self.compute_indices_from_displacement(dx, dy)
        arr = np.ascontiguousarray(np.asarray(correl), np.float32)
        maps = torch.from_numpy(arr).to(dev)[None]
        tix = torch.tensor([int(ix)], dtype=torch.int32, device=dev)
        tiy = torch.tensor([int(iy)], dtype=torch.int32, device=dev)
        ox = torch.empty(1, dtype=torch.float32, device=dev)
        oy = torch.empty(1, dtype=torch.float32, device=dev)
        er = torch.empty(1, dtype=torch.int32, device=dev)
        n0, n1 = arr.shape
        _native.check(
            self._lib.piv_subpix_any(
                1, ctypes.c_void_p(maps.data_ptr()), n0, n1, ctypes.c_void_p(tix.data_ptr()),
                ctypes.c_void_p(tiy.data_ptr()), _native.PIV_SUBPIX[self.subpix.method],
                self.subpix.n, ctypes.c_void_p(ox.data_ptr()), ctypes.c_void_p(oy.data_ptr()),
                ctypes.c_void_p(er.data_ptr()), _native.stream_ptr()),
            "piv_subpix_any")
        torch.cuda.current_stream().synchronize()
        err = int(er.item())
        if err != _native.WIN_OK:
            raise PIVError(explanation=_native.EXPLANATIONS[err],
                           result_compute_subpix=(iy, ix))
        return self.compute_displacement_from_indices(float(ox.item()), float(oy.item()))


class CorrelPythranHIP(CorrelDirectHIP):
    """CorrelPythran's direct correlation (correl_numpy, correl.py:284-412)."""

    _tag = "pythran"
    _kind = "pythran"


class CorrelPyCudaHIP(CorrelDirectHIP):
    """CorrelPyCuda's direct correlation (correl_pycuda.py:26-240,
    correl.py:415-451): plain sums, norm times the window size."""

    _tag = "pycuda"
    _kind = "pycuda"


# "fftw" (the reference's default key, singlepass.py:590) selects the same
# FFT correlation, so reference parameter trees work unchanged; "pythran" and
# "pycuda" the direct correlation (per-window API)
correlation_classes = {CorrelHIP._tag: CorrelHIP, "fftw": CorrelHIP,
                       CorrelPythranHIP._tag: CorrelPythranHIP,
                       CorrelPyCudaHIP._tag: CorrelPyCudaHIP}


def register(registry):
    """Insert the HIP classes into a reference-style ``correlation_classes``
    (the FFT class as "hip", the direct ones under their reference tags)."""
    registry[CorrelHIP._tag] = CorrelHIP
    registry[CorrelPythranHIP._tag] = CorrelPythranHIP
    registry[CorrelPyCudaHIP._tag] = CorrelPyCudaHIP
    return registry

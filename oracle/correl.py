"""Oracle: FFT cross-correlation, search mask, peak search (test infrastructure).

Restates src/fluidimage/calcul/correl.py (CorrelBase, CorrelFFTBase,
CorrelFFTWithOperBase/CorrelFFTW, nan_indices_max, _compute_indices_max,
_like_fftshift) and src/fluidimage/calcul/fft.py (FFTW2DReal2Complex norms,
_compute_energy_from_fourier). FFTW is replaced by numpy's pocketfft, in
float32 (``precision="f32"``, like the reference's float32 FFTW plans) or in
float64 (``precision="f64"``, the accurate reference used for parity and tie
detection).
"""

from math import sqrt

import numpy as np

from .subpix import SubPix

# error codes shared with the C-ABI header (fluidimage_amd/include); the
# explanation strings are the reference's own (correl.py:128,130,
# subpix.py:154,215, singlepass.py:310)
NO_ERROR = 0
ERR_PEAK_BOUNDARY = 1
ERR_PEAK_NAN = 2
ERR_SUBPIX_BOUNDARY = 3
ERR_SUBPIX_WRONG = 4
ERR_BAD_CROP = 5
ERR_NO_PEAK = 6

EXPLANATIONS = {
    ERR_PEAK_BOUNDARY: "Correlation peak touching boundary.",
    ERR_PEAK_NAN: "Correlation peak touching nan.",
    ERR_SUBPIX_BOUNDARY: "close boundary",
    ERR_SUBPIX_WRONG: "wrong subpix",
    ERR_BAD_CROP: "Bad im_crop shape.",
    ERR_NO_PEAK: "No valid correlation value in the search box.",
}


class PIVError(Exception):
    """Restates src/fluidimage/calcul/errors.py:8-15."""

    def __init__(self, *args, **kargs):
        self.explanation = "General no peak error"
        self.code = None
        for k, v in kargs.items():
            self.__dict__[k] = v
        super().__init__(*args)


def parse_displacement_max(displ_max, im0_shape):
    """src/fluidimage/calcul/correl.py:68-73"""
    if isinstance(displ_max, str) and displ_max.endswith("%"):
        return float(displ_max[:-1]) / 100 * max(im0_shape)
    return displ_max


def fftshift_rows_cols(n0, n1):
    """Index maps of ``_like_fftshift``: out[i0, i1] = arr[r0[i0], r1[i1]].

    src/fluidimage/calcul/correl.py:596-636 computes
    ``fftshift(arr[::-1, ::-1])``; along one axis of length n this reads
    ``arr[n - 1 - ((i - n // 2) % n)]``.
    """
    r0 = n0 - 1 - (np.arange(n0) - n0 // 2) % n0
    r1 = n1 - 1 - (np.arange(n1) - n1 // 2) % n1
    return r0, r1


def like_fftshift(arr):
    """src/fluidimage/calcul/correl.py:596-636 (works on a leading batch axis)."""
    n0, n1 = arr.shape[-2:]
    r0, r1 = fftshift_rows_cols(n0, n1)
    return np.ascontiguousarray(arr[..., r0, :][..., r1])


def energy_from_fourier(field_fft, coef_norm):
    """src/fluidimage/calcul/fft.py:72-90, on the last two axes.

    Half-spectrum energy with weight 1 on the first and the last stored
    column and 2 on the others (for odd n1 the last column keeps weight 1,
    as in the reference). Accumulated in float64, like the Pythran double
    accumulator.
    """
    a2 = np.abs(field_fft.astype(np.complex128)) ** 2
    n1 = field_fft.shape[-1]
    total = a2[..., :, 0].sum(-1) + a2[..., :, n1 - 1].sum(-1)
    if n1 > 2:
        total = total + 2 * a2[..., :, 1 : n1 - 1].sum((-2, -1))
    return 0.5 / coef_norm * total


def correl_fft(im0, im1, precision="f64"):
    """CorrelFFTW.__call__: src/fluidimage/calcul/correl.py:648-662.

    ``im0``/``im1`` are (n0, n1) or (nwin, n0, n1). Returns the map after
    ``_like_fftshift`` (float32 for "f32", float64 for "f64") and the norm
    ``sqrt(4 E0 E1) * (n0 n1)**2`` (src/fluidimage/calcul/fft.py:100-109).
    The c2r transform is unnormalised (fft.py:243), hence the factor n0 n1
    on numpy's normalised inverse.
    """
    dtype = np.float32 if precision == "f32" else np.float64
    im0 = np.asarray(im0, dtype=dtype)
    im1 = np.asarray(im1, dtype=dtype)
    n0, n1 = im0.shape[-2:]
    coef_norm = n0 * n1
    f0 = np.fft.rfft2(im0)
    f0[..., 0, 0] = 0.0
    f1 = np.fft.rfft2(im1)
    f1[..., 0, 0] = 0.0
    energy0 = energy_from_fourier(f0, coef_norm**2)
    energy1 = energy_from_fourier(f1, coef_norm**2)
    norm = np.sqrt(4 * energy0 * energy1) * coef_norm**2
    correl = np.fft.irfft2(f0.conj() * f1, s=(n0, n1)) * dtype(coef_norm)
    correl = correl.astype(dtype, copy=False)
    return like_fftshift(correl), norm


def search_mask(shape, displacement_max):
    """CorrelFFTBase._finalize_init: src/fluidimage/calcul/correl.py:518-549.

    Returns ``(where_large_displacement, (i0_start, i0_stop),
    (i1_start, i1_stop))``; ``(None, (0, None), (0, None))`` when
    ``displacement_max`` is None (correl.py:198-199).
    """
    if displacement_max is None:
        return None, (0, None), (0, None)
    n0, n1 = shape
    iy0, ix0 = n0 // 2 - 1, n1 // 2 - 1
    wld = np.zeros(shape, dtype=bool)
    for i0 in range(n0):
        for i1 in range(n1):
            dx, dy = ix0 - i1, iy0 - i0
            if sqrt(dx**2 + dy**2) > displacement_max:
                wld[i0, i1] = True
    for i0_start in range(n0):
        if not all(wld[i0_start, :]):
            break
    for i1_start in range(n1):
        if not all(wld[:, i1_start]):
            break
    for i0_stop in range(n0 - 1, -1, -1):
        if not all(wld[i0_stop, :]):
            break
    i0_stop += 1
    for i1_stop in range(n1 - 1, -1, -1):
        if not all(wld[:, i1_stop]):
            break
    i1_stop += 1
    return wld, (i0_start, i0_stop), (i1_start, i1_stop)


def nan_indices_max_loop(correl, i0_start, i0_stop, i1_start, i1_stop):
    """Literal sequential scan of src/fluidimage/calcul/correl.py:88-132.

    Pure-Python loops; used to pin the vectorised ``nan_indices_max``.
    Returns ``(i0_max, i1_max, code)``.
    """
    n0, n1 = correl.shape
    flat = correl.ravel()
    correl_max = np.nan
    for i_flat in range(i0_start * n1 + i1_start, n0 * n1):
        value = flat[i_flat]
        if not np.isnan(value):
            correl_max = value
            break
    if np.isnan(correl_max):
        return 0, 0, ERR_NO_PEAK
    i0_max = 0
    i1_max = 0
    for i0 in range(i0_start, i0_stop):
        for i1 in range(i1_start, i1_stop):
            value = correl[i0, i1]
            if np.isnan(value):
                continue
            if value >= correl_max:
                correl_max = value
                i0_max = i0
                i1_max = i1
    i0, i1 = i0_max, i1_max
    if i0 == 0 or i0 == n0 - 1 or i1 == 0 or i1 == n1 - 1:
        return i0_max, i1_max, ERR_PEAK_BOUNDARY
    if np.isnan(correl[i0 - 1 : i0 + 2, i1 - 1 : i1 + 2]).any():
        return i0_max, i1_max, ERR_PEAK_NAN
    return i0_max, i1_max, NO_ERROR


def nan_indices_max(correl, i0_start, i0_stop, i1_start, i1_stop):
    """Vectorised src/fluidimage/calcul/correl.py:88-132.

    The scan seeds with the first non-NaN value at or after
    (i0_start, i1_start) in row-major order and keeps a candidate when
    ``value >= current``, so the LAST maximal element of the box wins.
    Returns ``(i0_max, i1_max, code)``; ``ERR_NO_PEAK`` replaces the
    reference's ``assert`` when no value is finite.
    """
    n0, n1 = correl.shape
    flat = correl.ravel()
    tail = flat[i0_start * n1 + i1_start :]
    finite = ~np.isnan(tail)
    if not finite.any():
        return 0, 0, ERR_NO_PEAK
    seed = tail[np.argmax(finite)]
    box = correl[i0_start:i0_stop, i1_start:i1_stop]
    ok = ~np.isnan(box)
    cand = ok & (box >= seed)
    if not cand.any():
        # only possible when the seed lies outside the box
        i0, i1 = 0, 0
    else:
        best = np.max(box[cand])
        hits = np.flatnonzero((box == best) & ok)
        i0b, i1b = divmod(int(hits[-1]), box.shape[1])
        i0, i1 = i0_start + i0b, i1_start + i1b
    if i0 == 0 or i0 == n0 - 1 or i1 == 0 or i1 == n1 - 1:
        return i0, i1, ERR_PEAK_BOUNDARY
    if np.isnan(correl[i0 - 1 : i0 + 2, i1 - 1 : i1 + 2]).any():
        return i0, i1, ERR_PEAK_NAN
    return i0, i1, NO_ERROR


def _python_slice(start, stop, n):
    """Bounds numpy uses for ``a[start:stop]`` on an axis of length n."""
    return slice(start, stop).indices(n)[:2]


class OracleCorrelFFT:
    """CorrelFFTW restated (src/fluidimage/calcul/correl.py:167-281,513-568,
    639-692) with numpy FFTs in ``precision``."""

    _tag = "fftw"

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
        precision="f64",
    ):
        im0_shape = tuple(int(n) for n in im0_shape)
        if im1_shape is None:
            im1_shape = im0_shape
        im1_shape = tuple(int(n) for n in im1_shape)
        self.mode = mode
        self.precision = precision
        self.subpix = SubPix(method=method_subpix, nsubpix=nsubpix)
        self.im0_shape = im0_shape
        self.im1_shape = im1_shape
        self.iy0, self.ix0 = (i // 2 - 1 for i in im0_shape)
        self.displacement_max = parse_displacement_max(displacement_max, im0_shape)
        self.particle_radius = particle_radius
        self.nb_peaks_to_search = nb_peaks_to_search
        (
            self.where_large_displacement,
            self.start_stop_for_search0,
            self.start_stop_for_search1,
        ) = search_mask(im0_shape, self.displacement_max)
        if im0_shape != im1_shape:
            raise ValueError(
                "with this correlation method the input images "
                "have to have the same shape."
            )

    def __call__(self, im0, im1):
        return correl_fft(im0, im1, self.precision)

    def compute_displacement_from_indices(self, ix, iy):
        return self.ix0 - ix, self.iy0 - iy

    def compute_indices_from_displacement(self, dx, dy):
        return self.ix0 - dx, self.iy0 - dy

    def get_indices_no_displacement(self):
        return self.iy0, self.ix0

    def _compute_indices_max(self, correl, norm):
        """src/fluidimage/calcul/correl.py:135-164"""
        i0_start, i0_stop = self.start_stop_for_search0
        i1_start, i1_stop = self.start_stop_for_search1
        if i0_stop is None:
            i0_stop, i1_stop = correl.shape
        iy, ix, code = nan_indices_max(correl, i0_start, i0_stop, i1_start, i1_stop)
        if code == ERR_NO_PEAK:
            err = PIVError(explanation=EXPLANATIONS[code], code=code)
            err.results = (np.nan, np.nan, np.nan)
            raise err
        if norm == 0:
            correl_max = 0.0
        else:
            # numpy 2 promotion: float32 scalar / Python float -> float32
            correl_max = np.float32(correl[iy, ix]) / np.float32(norm)
        if code:
            err = PIVError(explanation=EXPLANATIONS[code], code=code)
            err.results = (ix, iy, correl_max)
            raise err
        return ix, iy, correl_max

    def _nan_square(self, correl, ix, iy):
        r = self.particle_radius
        correl[iy - r : iy + r + 1, ix - r : ix + r + 1] = np.nan

    def compute_displacements_from_correl(self, correl, norm=1.0):
        """src/fluidimage/calcul/correl.py:561-568 then :226-275.

        With ``displacement_max`` None the second-chance NaN square is
        written into the caller's array, as in the reference.
        """
        if self.displacement_max is not None:
            correl = correl.copy()
            correl[self.where_large_displacement] = np.nan
        try:
            ix, iy, correl_max = self._compute_indices_max(correl, norm)
        except PIVError as piv_error:
            ix, iy, correl_max = piv_error.results
            if getattr(piv_error, "code", None) == ERR_NO_PEAK:
                raise
            self._nan_square(correl, ix, iy)
            try:
                ix2, iy2, correl_max2 = self._compute_indices_max(correl, norm)
            except PIVError as _piv_error:
                dx, dy = self.compute_displacement_from_indices(ix, iy)
                _piv_error.results = (dx, dy, correl_max)
                raise _piv_error
            else:
                ix, iy, correl_max = ix2, iy2, correl_max2

        dx, dy = self.compute_displacement_from_indices(ix, iy)

        if self.nb_peaks_to_search == 1:
            other_peaks = None
        elif self.nb_peaks_to_search >= 1:
            correl = correl.copy()
            other_peaks = []
            for _ in range(0, self.nb_peaks_to_search - 1):
                self._nan_square(correl, ix, iy)
                try:
                    ix, iy, correl_max_other = self._compute_indices_max(correl, norm)
                except PIVError:
                    break
                dx_other, dy_other = self.compute_displacement_from_indices(ix, iy)
                other_peaks.append((dx_other, dy_other, correl_max_other))
        else:
            raise ValueError
        return dx, dy, correl_max, other_peaks

    def apply_subpix(self, dx, dy, correl):
        """src/fluidimage/calcul/correl.py:277-281"""
        ix, iy = self.compute_indices_from_displacement(dx, dy)
        ix, iy = self.subpix.compute_subpix(correl, ix, iy)
        return self.compute_displacement_from_indices(ix, iy)


def brute_force_correl(im0, im1):
    """FP64 circular cross-correlation of the mean-removed windows, times
    n0 n1, laid out like the reference map (pins ``correl_fft``).

    ``out[d] = n0 n1 * sum_j x0'[j] x1'[j + d]`` at map index
    ``(iy0 - dy, ix0 - dx)``; entries are filled by direct summation.
    """
    x0 = np.asarray(im0, np.float64)
    x1 = np.asarray(im1, np.float64)
    x0 = x0 - x0.mean()
    x1 = x1 - x1.mean()
    n0, n1 = x0.shape
    iy0, ix0 = n0 // 2 - 1, n1 // 2 - 1
    out = np.empty((n0, n1))
    for i0 in range(n0):
        for i1 in range(n1):
            dy, dx = iy0 - i0, ix0 - i1
            shifted = np.roll(np.roll(x1, -dy, axis=0), -dx, axis=1)
            out[i0, i1] = n0 * n1 * np.sum(x0 * shifted)
    return out


def default_displacement_max_direct(im0_shape, im1_shape, kind):
    """displacement_max of CorrelPythran / CorrelPyCuda when the parameter is
    None (src/fluidimage/calcul/correl.py:388-408, 418-442): min(shape) // 2
    - 1 (pythran) or min(shape) // 2 (pycuda) for equal shapes, else
    min(n0 - n1 per axis) // 2 - 1."""
    if tuple(im0_shape) == tuple(im1_shape):
        return min(im0_shape) // 2 - (1 if kind == "pythran" else 0)
    return min(im0_shape[0] - im1_shape[0], im0_shape[1] - im1_shape[1]) // 2 - 1


def _direct_overlap(n0, n1, d):
    """(length, start in im1, start in im0) of the overlap on one axis for
    the displacement d (correl.py:327-333 for d <= 0, :341-344 for d > 0)."""
    if d <= 0:
        c = n0 // 2 - n1 // 2 + d
        return n1 + min(c, 0), -min(c, 0), max(0, c)
    return n1 - max(n0 // 2 + n1 // 2 + d - n0, 0), 0, n0 // 2 - n1 // 2 + d


def correl_direct(im0, im1, disp_max, normalisation="pythran"):
    """correl_numpy (src/fluidimage/calcul/correl.py:284-381, "pythran") and
    correl_pycuda (src/fluidimage/calcul/correl_pycuda.py:26-240, "pycuda"):
    map[dmax + dy, dmax + dx] = sum of im1[y, x] im0[y + oy, x + ox] over the
    overlap; "pythran" divides by the overlap area and multiplies by
    im1.size, norm sqrt(sum im1^2 sum im0^2); "pycuda" keeps the sum, norm
    times im0.size. FP64 sums, the map rounded to float32 (the reference
    sums in float32)."""
    a = np.asarray(im0, np.float64)
    b = np.asarray(im1, np.float64)
    (ny0, nx0), (ny1, nx1) = a.shape, b.shape
    n = 2 * int(disp_max) + 1
    out = np.empty((n, n), np.float64)
    for i in range(n):
        ly, y1, y0 = _direct_overlap(ny0, ny1, i - disp_max)
        for j in range(n):
            lx, x1, x0 = _direct_overlap(nx0, nx1, j - disp_max)
            s = np.sum(b[y1:y1 + ly, x1:x1 + lx] * a[y0:y0 + ly, x0:x0 + lx])
            out[i, j] = s / (lx * ly) * b.size if normalisation == "pythran" else s
    norm = sqrt(np.sum(b**2) * np.sum(a**2))
    if normalisation == "pycuda":
        norm *= a.size
    return out.astype(np.float32), norm


class OracleCorrelDirect(OracleCorrelFFT):
    """CorrelPythran / CorrelPyCuda restated: direct maps of (2 dmax + 1)^2
    with the zero displacement at (dmax, dmax) (correl.py:409-410,444-445)
    and CorrelBase's search without mask (correl.py:226-281)."""

    def __init__(self, im0_shape, im1_shape=None, method_subpix="centroid", nsubpix=1,
                 displacement_max=None, particle_radius=3, nb_peaks_to_search=1, mode=None,
                 kind="pythran"):
        im0_shape = tuple(int(n) for n in im0_shape)
        im1_shape = im0_shape if im1_shape is None else tuple(int(n) for n in im1_shape)
        self.kind = kind
        self.mode = mode
        self.subpix = SubPix(method=method_subpix, nsubpix=nsubpix)
        self.im0_shape, self.im1_shape = im0_shape, im1_shape
        dmax = parse_displacement_max(displacement_max, im0_shape)
        if dmax is None:
            dmax = default_displacement_max_direct(im0_shape, im1_shape, kind)
        if dmax <= 0:
            raise ValueError("displacement_max <= 0 : problem with images shapes?")
        self.displacement_max = int(dmax)
        self.ix0 = self.iy0 = self.displacement_max
        self.particle_radius = particle_radius
        self.nb_peaks_to_search = nb_peaks_to_search
        self.start_stop_for_search0 = [0, None]
        self.start_stop_for_search1 = [0, None]

    def __call__(self, im0, im1):
        return correl_direct(im0, im1, self.displacement_max, self.kind)

    def compute_displacements_from_correl(self, correl, norm=1.0):
        """CorrelBase.compute_displacements_from_correl (no search mask: the
        second-chance square lands in the caller's array)."""
        dmax, self.displacement_max = self.displacement_max, None
        try:
            return super().compute_displacements_from_correl(correl, norm)
        finally:
            self.displacement_max = dmax

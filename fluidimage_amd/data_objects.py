"""PIV result containers with the reference's attribute names.

``HeavyPIVResults`` / ``MultipassPIVResults`` follow
src/fluidimage/data_objects/piv.py:214-276,538-585 (attributes, ``passes``,
``piv0``/``piv1``... and item access). File formats (HDF5, UVmat) are not
part of the GPU path.
"""

import numpy as np


class ArrayCouple:
    """Couple of arrays (src/fluidimage/data_objects/piv.py:78-200, arrays only)."""

    def __init__(self, names=None, arrays=None, params_mask=None, serie=None):
        if arrays is None:
            raise ValueError("ArrayCouple needs arrays")
        self.arrays = tuple(arrays)
        self.names = tuple(names) if names is not None else tuple(
            f"array{i}" for i in range(len(self.arrays)))
        self.name = "-".join(self.names)
        self.params_mask = params_mask

    def apply_mask(self, params_mask):
        """with_mask.py / piv.py:71-75: crop both arrays by ``strcrop``."""
        if params_mask is None or getattr(params_mask, "strcrop", None) is None:
            return
        from .works.with_mask import slices_from_strcrop

        sl = slices_from_strcrop(params_mask.strcrop)
        self.arrays = tuple(a[sl] for a in self.arrays)

    def get_arrays(self):
        return self.arrays


class HeavyPIVResults:
    """One pass: deltaxs, deltays, xs, ys, errors, correls_max, correls,
    secondary_peaks, plus the interpolated fields set by apply_interp."""

    def __init__(self, deltaxs, deltays, xs, ys, errors, correls_max=None, correls=None,
                 couple=None, params=None, secondary_peaks=None):
        self.deltaxs = deltaxs
        self.deltays = deltays
        self.xs = xs
        self.ys = ys
        self.errors = errors
        self.correls_max = correls_max
        self.correls = correls
        self.couple = couple
        self.params = params
        self.secondary_peaks = secondary_peaks

    def get_grid_pixel(self):
        return self.xs, self.ys


class MultipassPIVResults:
    """Passes of a multipass computation (``passes``, ``piv0``...)."""

    def __init__(self):
        self.passes = []

    def __getitem__(self, key):
        return self.passes[key]

    def __len__(self):
        return len(self.passes)

    def append(self, results):
        i = len(self.passes)
        self.passes.append(results)
        self.__dict__[f"piv{i}"] = results

    @property
    def last(self):
        return self.passes[-1]


def error_strings(codes, flags=None, diff2=None, explanations=None):
    """Build the reference's ``errors`` dict (ivec -> explanation) from the
    per-window codes of the correlation and the WorkFIX flag bits, in the
    order WorkFIX composes them (fix.py:73-130)."""
    from . import _native

    explanations = explanations or _native.EXPLANATIONS
    errors = {}
    codes = np.asarray(codes)
    for i in np.flatnonzero(codes != 0):
        errors[int(i)] = explanations[int(codes[i])]
    if flags is None:
        return errors
    flags = np.asarray(flags)
    labels = ((2, "correl < correl_min"), (4, "delta2 < displacement_max2"),
              (8, "diff neighbour too large"))
    for bit, text in labels:
        for i in np.flatnonzero(flags & bit):
            i = int(i)
            if i in errors:
                errors[i] += " + " + text
            else:
                errors[i] = text
    if diff2 is not None:
        d2 = np.asarray(diff2)
        for i in np.flatnonzero(flags & 8):
            errors[int(i)] += f" (diff = {np.sqrt(np.float64(d2[i])):.2f})"
    return errors

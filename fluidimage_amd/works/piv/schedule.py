"""Window geometry of the passes of a multipass PIV.

The interface (parameter names, error texts) is the reference's
(src/fluidimage/works/piv/singlepass.py:52-87, multipass.py:107-170): a
window size is an integer or a pair of integers, the second image's window
defaults to the first's and may not exceed it, and each later pass divides
the previous pass's (unrounded) window size by its zoom factor, truncating
to an integer only when the pass builds its windows.

Everything is computed once per work, on the host.
"""

import numbers

import numpy as np

_SHAPE_ERRORS = {
    "shape_crop_im0": "For now, shape_crop_im0 has to be one or two integer!",
    "shape_crop_im1": "For now, shape_crop_im1 has to be one or two integer!",
}


def as_pair(value, name, integral=numbers.Integral):
    """(n0, n1) from an integer or a 2-sequence; the reference's
    NotImplementedError otherwise."""
    if isinstance(value, integral):
        return (value, value)
    if isinstance(value, (tuple, list)) and len(value) == 2:
        return (value[0], value[1])
    raise NotImplementedError(_SHAPE_ERRORS[name])


def window_shapes(shape_crop_im0, shape_crop_im1=None):
    """Integer window shapes of one pass from the (possibly fractional)
    sizes given to it."""
    s0 = as_pair(shape_crop_im0, "shape_crop_im0")
    s1 = s0 if shape_crop_im1 is None else as_pair(shape_crop_im1, "shape_crop_im1")
    if any(b > a for a, b in zip(s0, s1)):
        raise ValueError("shape_crop_im1 must be inferior or equal to shape_crop_im0")
    return tuple(int(n) for n in s0), tuple(int(n) for n in s1)


def zoom_factors(number, coeff_zoom):
    """One zoom factor per later pass: an int repeated, or a sequence of
    exactly number - 1 factors."""
    if isinstance(coeff_zoom, int):
        return [coeff_zoom] * (number - 1)
    if len(coeff_zoom) == number - 1:
        return list(coeff_zoom)
    raise ValueError(
        "params.multipass.coeff_zoom has to be an integer or "
        "an iterable of length params.multipass.number - 1"
    )


def pass_schedule(shape_crop_im0, shape_crop_im1, number, coeff_zoom):
    """Window sizes handed to every pass, as [(size0, size1)] (the first
    pass gets the parameters themselves): each later size is the previous
    one divided by the pass's zoom factor, in floating point."""
    factors = zoom_factors(number, coeff_zoom)
    # the multipass work accepts only plain ints as scalar sizes
    cur0 = as_pair(shape_crop_im0, "shape_crop_im0", integral=int)
    cur1 = cur0 if shape_crop_im1 is None else as_pair(shape_crop_im1, "shape_crop_im1",
                                                       integral=int)
    sizes = [(shape_crop_im0, shape_crop_im1)]
    for z in factors:
        cur0 = tuple(n / z for n in cur0)
        cur1 = tuple(n / z for n in cur1)
        sizes.append((cur0, cur1))
    return sizes


def schedule_params(params, windows):
    """Set ``params`` (piv0.shape_crop_im0, multipass.number and coeff_zoom)
    to the square window sizes ``windows`` (e.g. (128, 64, 32)) and check
    that the pass schedule reproduces them."""
    windows = [int(w) for w in windows]
    factors = [windows[k] // windows[k + 1] for k in range(len(windows) - 1)]
    params.piv0.shape_crop_im0 = windows[0]
    params.piv0.shape_crop_im1 = None
    params.multipass.number = len(windows)
    params.multipass.coeff_zoom = factors
    got = [window_shapes(*s)[0][0]
           for s in pass_schedule(windows[0], None, len(windows), factors)]
    if got != windows:
        raise ValueError(f"window schedule {windows} is not a chain of integer zooms")
    return params


def grid_positions(length, start, stop, step):
    """Window centres along one image axis of ``length`` px: every ``step``
    from ``start`` (the window's half size) up to ``length - stop``
    excluded, plus ``length - stop`` itself when the gap left after the last
    centre exceeds ``step // 2.5`` (singlepass.py:108-141)."""
    last = length - stop
    pos = np.arange(start, last, step, dtype=int)
    if last - pos[-1] > step // 2.5:
        pos = np.append(pos, last)
    return pos


def grid_step(size, overlap):
    """Distance between window centres for a window of ``size`` px."""
    step = size - int(np.round(overlap * size))
    assert step >= 1
    return step

"""Sliding-window preprocessing tools on the GPU.

Drop-in counterparts of the python toolbox's spatial filters
(src/fluidimage/works/preproc/_toolbox_py.py:66-165), with the same
signatures, the same scipy.ndimage filter conventions (window offsets
-(w // 2) .. w - 1 - w // 2, boundary modes, the rank of a median or a
percentile) and the same output types: float64 for integer images (numpy
promotes the float weight), float32 for a float32 image. Several images
(a 3-D array or a list of 2-D arrays) are filtered in one launch and
returned as a float64 array, as the reference's ``iterate_multiple_imgs``
decorator returns them (src/fluidimage/works/preproc/io.py:44-63).

Images given as torch tensors on the GPU stay there (the result is a
device tensor); numpy images are uploaded and the result copied back.
"""

import ctypes
import numbers

import numpy as np

from ... import _native

__all__ = ["sliding_median", "sliding_minima", "sliding_percentile"]

_MODE_ALIASES = {"grid-mirror": "reflect", "grid-wrap": "wrap", "grid-constant": "constant"}


def _sizes(window_size):
    """scipy.ndimage's ``size`` for a 2-D image: a scalar or one per axis."""
    if isinstance(window_size, numbers.Integral):
        return int(window_size), int(window_size)
    sizes = tuple(int(s) for s in window_size)
    if len(sizes) != 2:
        raise RuntimeError("sequence argument must have length equal to input rank")
    return sizes


def filter_rank(operation, wy, wx, percentile=None):
    """The rank scipy.ndimage's _rank_filter selects in a wy x wx footprint
    (median: size // 2; percentile: int(size * p / 100), size - 1 at 100,
    negative percentiles counted from 100)."""
    size = wy * wx
    if operation == "minimum":
        return 0
    if operation == "median":
        return size // 2
    p = float(percentile)
    if p < 0.0:
        p += 100.0
    if p < 0 or p > 100:
        raise RuntimeError("invalid percentile")
    return size - 1 if p == 100.0 else int(float(size) * p / 100.0)


_TYPES = (np.uint8, np.uint16, np.float32)


def _checked(a):
    if a.dtype not in _TYPES:
        raise TypeError(f"images of uint8, uint16 or float32 (got {a.dtype})")
    return a


def _device_images(img):
    """(list of 2-D device tensors, host input?, several images?)."""
    import torch

    from ...device import to_device_image

    if not torch.cuda.is_available():
        raise _native.NativeLibraryError("the sliding filters need a ROCm GPU (torch.cuda)")
    dev = torch.device("cuda", torch.cuda.current_device())
    if isinstance(img, torch.Tensor):
        if img.dim() == 2:
            return [to_device_image(img.to(dev), dev)], False, False
        return [to_device_image(t.to(dev), dev) for t in img], False, True
    if isinstance(img, np.ndarray) and img.ndim == 2:
        return [to_device_image(_checked(img), dev)], True, False
    return [to_device_image(_checked(np.asarray(a)), dev) for a in img], True, True


def _sliding(img, operation, weight, window_size, boundary_condition, percentile=None,
             cval=0.0):
    import torch

    from ...device import ImageSet

    mode = _MODE_ALIASES.get(boundary_condition, boundary_condition)
    if mode not in _native.PIV_BOUNDARY:
        raise RuntimeError("boundary mode not supported")
    wy, wx = _sizes(window_size)
    rank = filter_rank(operation, wy, wx, percentile)
    ims, host, several = _device_images(img)
    shapes = {tuple(t.shape) for t in ims}
    if len(shapes) != 1:
        raise ValueError("all images must have the same shape")
    (h, w), = shapes
    images = ImageSet(ims, subtract_min=False)
    f32 = images.dtype == _native.PIV_F32
    out = torch.empty((len(ims), h, w), dtype=torch.float32 if f32 else torch.float64,
                      device=ims[0].device)
    lib = _native.load()
    _native.check(
        lib.piv_sliding_filter(ctypes.c_void_p(images.table.data_ptr()), len(ims), h, w,
                               images.dtype, wy, wx, rank, _native.PIV_BOUNDARY[mode],
                               float(cval), float(weight), ctypes.c_void_p(out.data_ptr()),
                               _native.stream_ptr(None)),
        "piv_sliding_filter",
    )
    if several:
        out = out.to(torch.float64)
        return out.cpu().numpy() if host else out
    out = out[0]
    return out.cpu().numpy() if host else out


def sliding_median(img=None, weight=1.0, window_size=30, boundary_condition="reflect"):
    """img - weight * median over the sliding window, negatives set to 0
    (_toolbox_py.py:66-94)."""
    return _sliding(img, "median", weight, window_size, boundary_condition)


def sliding_percentile(img=None, percentile=10.0, weight=1.0, window_size=30,
                       boundary_condition="reflect"):
    """img - weight * percentile over the sliding window, negatives set to 0
    (_toolbox_py.py:97-131)."""
    return _sliding(img, "percentile", weight, window_size, boundary_condition,
                    percentile=percentile)


def sliding_minima(img=None, weight=1.0, window_size=30, boundary_condition="reflect"):
    """img - weight * minimum over the sliding window, negatives set to 0
    (_toolbox_py.py:134-160)."""
    return _sliding(img, "minimum", weight, window_size, boundary_condition)


class PreprocToolsHIP:
    """The tools above as static methods, like the reference's
    PreprocToolsPy / PreprocToolsCV backends
    (src/fluidimage/works/preproc/toolbox.py)."""

    sliding_median = staticmethod(sliding_median)
    sliding_percentile = staticmethod(sliding_percentile)
    sliding_minima = staticmethod(sliding_minima)
    available_tools = list(__all__)
